"""GPU parity: the HIP evaluator against the CPU oracle on the same frozen
compressed structure, through the C-ABI.

Bars: FP64 plan and drop-in entry points <= 1e-12 relative Frobenius error
(and exactly the same index handling); FP32 and split-BF16x3 <= 1e-5 against
the FP64 evaluation of the same K~.
"""
import sys

import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = {
    # name: builder arguments
    "h2b_small": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0),
    "hss_small": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.0, h=1.0),
    "h2b_ragged": dict(n=3001, d=5, leaf=100, rank=70, tol=1e-7, budget=0.3, h=1.2),
    "h2b_fullrank": dict(n=1024, d=5, leaf=64, rank=100000, tol=0.0, budget=0.25, h=1.3),
}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run_plan(hm, TW, nrhs, prec, dtype=None):
    import torch
    plan = hgemm.Plan(hm, precisions=(prec,))
    dt = dtype or (torch.float64 if prec == "fp64" else torch.float32)
    W = torch.from_numpy(np.ascontiguousarray(TW)).to(device="cuda", dtype=dt)
    Y = torch.full_like(W, float("nan"))
    plan.eval(W, Y, nrhs, precision=prec)
    torch.cuda.synchronize()
    return plan, Y.double().cpu().numpy()


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("nrhs", [1, 37, 64, 130])
def test_fp64_plan_matches_oracle(gpu, case, nrhs):
    _say(f"[fp64 plan] case={case} nrhs={nrhs}")
    X, hm = make_problem(**CASES[case])
    W = host.normal((hm.N, nrhs), seed=100 + nrhs)
    TW = hm.transform(W)
    Yo, wo, uo = oracle.hgemm(hm, TW, nrhs)
    plan, Yg = _run_plan(hm, TW, nrhs, "fp64")
    assert np.isfinite(Yg).all()
    assert rel(Yg, Yo) <= 1e-12
    import torch
    ws, us = plan.skeletons(nrhs, torch.float64)
    assert rel(ws.cpu().numpy(), wo) <= 1e-12
    assert rel(us.cpu().numpy(), uo) <= 1e-12


@pytest.mark.parametrize("nrhs", [5, 70])
def test_dropin_hgemm_h2b_accumulates_like_reference(gpu, nrhs):
    _say(f"[HGEMM] nrhs={nrhs}")
    X, hm = make_problem(**CASES["h2b_ragged"])
    W = host.normal((hm.N, nrhs), seed=7)
    TW = hm.transform(W)
    A = host.normal((hm.N * nrhs,), seed=8)
    U = host.normal((int(hm.skel_rows) * nrhs,), seed=9)
    Yo, wo, uo = oracle.hgemm(hm, TW, nrhs, apres=A, uskel=U, variant="h2")
    Yg, wg, ug = hgemm.hgemm(hm, TW, nrhs, apres=A, uskel=U)
    assert rel(Yg, Yo) <= 1e-12 and rel(wg, wo) <= 1e-12 and rel(ug, uo) <= 1e-12


@pytest.mark.parametrize("nrhs", [3, 66])
def test_dropin_hgemm_hss_overwrites_like_reference(gpu, nrhs):
    _say(f"[HGEMM_HSS] nrhs={nrhs}")
    X, hm = make_problem(**CASES["hss_small"])
    W = host.normal((hm.N, nrhs), seed=10)
    TW = hm.transform(W)
    A = host.normal((hm.N * nrhs,), seed=11)
    U = host.normal((int(hm.skel_rows) * nrhs,), seed=12)
    Yo, wo, uo = oracle.hgemm(hm, TW, nrhs, apres=A, uskel=U, variant="hss")
    Yg, wg, ug = hgemm.hgemm_hss(hm, TW, nrhs, apres=A, uskel=U)
    assert rel(Yg, Yo) <= 1e-12 and rel(wg, wo) <= 1e-12 and rel(ug, uo) <= 1e-12


@pytest.mark.parametrize("n", [128, 200, 700])
def test_smallest_trees(gpu, n):
    """Two, four and sixteen leaves of 64 points (the smallest trees the
    builder makes; with so few leaves the budget near lists degenerate to
    HSS): every precision and schedule, forward and adjoint, against the
    oracle."""
    import torch
    _say(f"[small tree] n={n}")
    X, hm = make_problem(n=n, d=3, leaf=64, rank=20, tol=1e-6, budget=0.5, h=1.0)
    nrhs = 7
    TW = host.normal((hm.N * nrhs,), seed=31)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    Go, _, _ = oracle.hgemm(hm.transposed(), TW, nrhs)
    plan = hgemm.Plan(hm, precisions=("fp64", "fp32", "bf16x3"), adjoint=True)
    W = torch.from_numpy(TW).cuda()
    for prec in ("fp64", "fp32", "bf16x3"):
        # the dependency-driven sweep and the coarse lists are BF16x3 schedules
        for sc in sorted(s for s in hgemm.SCHEDULES
                         if prec == "bf16x3" or not (s == "sweep" or s.startswith("coarse"))):
            for d in ("forward", "adjoint"):
                plan.set_schedule(d, prec, sc)
            Y = torch.full_like(W, float("nan"))
            G = torch.full_like(W, float("nan"))
            plan.eval(W, Y, nrhs, prec)
            plan.eval_adjoint(W, G, nrhs, prec)
            torch.cuda.synchronize()
            tol = 1e-12 if prec == "fp64" else 1e-5
            ey, eg = rel(Y.cpu().numpy(), Yo), rel(G.cpu().numpy(), Go)
            assert ey <= tol and eg <= tol, (prec, sc, ey, eg)
        _say(f"  {prec}: nleaf={hm.nleaf} forward {ey:.2e} adjoint {eg:.2e}")


def test_fp64_closed_form_equals_dense(gpu):
    _say("[fp64 closed form]")
    X, hm = make_problem(**CASES["h2b_fullrank"])
    W = host.normal((hm.N, 9), seed=13)
    _, Yg = _run_plan(hm, hm.transform(W), 9, "fp64")
    K = oracle.dense_rows(X, 1.3, np.arange(hm.N), W)
    assert rel(hm.untransform(Yg, 9), K) <= 1e-12


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", ["h2b_small", "hss_small", "h2b_ragged"])
@pytest.mark.parametrize("nrhs", [1, 96, 200])
def test_mixed_precision_within_1e5_of_fp64(gpu, prec, case, nrhs):
    _say(f"[{prec}] case={case} nrhs={nrhs}")
    import torch
    X, hm = make_problem(**CASES[case])
    W = host.normal((hm.N, nrhs), seed=200 + nrhs)
    TW = hm.transform(W)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    # FP64 I/O (conversion inside the evaluator) and FP32 I/O
    _, Yg = _run_plan(hm, TW, nrhs, prec, dtype=torch.float64)
    assert np.isfinite(Yg).all()
    assert rel(Yg, Yo) <= 1e-5
    _, Yf = _run_plan(hm, TW, nrhs, prec, dtype=torch.float32)
    assert rel(Yf, Yo) <= 1e-5


def test_plan_errors(gpu):
    import torch
    X, hm = make_problem(**CASES["h2b_small"])
    plan = hgemm.Plan(hm, precisions=("fp64",))
    W = torch.zeros(hm.N * 4, dtype=torch.float64, device="cuda")
    Y = torch.zeros_like(W)
    with pytest.raises(hgemm.MxError) as e:
        plan.eval(W, Y, 4, precision="fp32")
    assert e.value.code == -4
    # corrupted structure is rejected at plan creation
    import copy
    bad = copy.copy(hm)
    bad.arrays = dict(hm.arrays)
    lc = hm.arrays["lchildren"].copy()
    lc[0] = hm.numnodes + 5
    bad.arrays["lchildren"] = lc
    with pytest.raises(hgemm.MxError) as e:
        hgemm.Plan(bad)
    assert e.value.code == -1
    bad2 = copy.copy(hm)
    bad2.arrays = dict(hm.arrays)
    dp = hm.arrays["Dptr"].copy()
    dp[-1] = len(hm.arrays["D"])
    bad2.arrays["Dptr"] = dp
    with pytest.raises(hgemm.MxError):
        hgemm.Plan(bad2)


def test_repeated_and_interleaved_evaluations(gpu):
    """Work buffers are reused across evaluations with different nrhs and
    precisions; results stay identical."""
    import torch
    X, hm = make_problem(**CASES["h2b_small"])
    plan = hgemm.Plan(hm, precisions=("fp64", "fp32", "bf16x3"))
    outs = {}
    for nrhs in (64, 8, 64):
        W = host.normal((hm.N, nrhs), seed=400 + nrhs)
        TW = torch.from_numpy(hm.transform(W)).cuda()
        for prec in ("fp64", "fp32", "bf16x3"):
            Y = torch.empty_like(TW)
            plan.eval(TW, Y, nrhs, precision=prec)
            torch.cuda.synchronize()
            key = (nrhs, prec)
            y = Y.cpu().numpy()
            if key in outs:
                assert np.array_equal(outs[key], y)
            outs[key] = y
    Yo, _, _ = oracle.hgemm(hm, hm.transform(host.normal((hm.N, 8), seed=408)), 8)
    assert rel(outs[(8, "fp64")], Yo) <= 1e-12


def test_dense_rows_matches_oracle(gpu):
    """GPU dense reference rows (used to report each result's error) against
    the oracle's directly computed K(rows, :) W."""
    import torch
    _say("[dense rows]")
    X, hm = make_problem(**CASES["h2b_ragged"])
    plan = hgemm.Plan(hm)
    perm = hm.lids[: hm.N]
    plan.set_points(X[perm], 1.2)
    nrhs = 33
    W = host.normal((hm.N, nrhs), seed=500)
    Wd = torch.from_numpy(hm.transform(W)).cuda()
    rows = np.array([0, 1, 7, 150, 2999, 3000, 1234], dtype=np.int32)   # leaf-order ids
    got = plan.dense_rows(rows, Wd, nrhs).cpu().numpy()
    ref = oracle.dense_rows(X, 1.2, perm[rows], W)
    assert rel(got, ref) <= 1e-12
    # the row extraction helper addresses the leaf-blocked layout
    Y = torch.empty_like(Wd)
    plan.eval(Wd, Y, nrhs)
    Yo, _, _ = oracle.hgemm(hm, hm.transform(W), nrhs)
    Yrows = plan.rows_of(Y, rows, nrhs).cpu().numpy()
    assert rel(Yrows, hm.untransform(Yo, nrhs)[perm[rows]]) <= 1e-12


def test_flops_and_launch_timing(gpu):
    import torch
    _say("[flops/timing]")
    X, hm = make_problem(**CASES["h2b_small"])
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"))
    for prec in ("fp64", "bf16x3"):
        tot, parts = plan.flops(prec, 64)
        assert tot == pytest.approx(oracle.flops(hm, 64), rel=1e-12)
    W = torch.from_numpy(hm.transform(host.normal((hm.N, 64), seed=1))).cuda()
    Y = torch.empty_like(W)
    plan.set_profiling(True)
    plan.eval(W, Y, 64, precision="bf16x3")
    times = plan.launch_times()
    kinds = {k for k, _ in times}
    assert kinds == {"bf16x3", "convert"}
    assert all(ms > 0 for _, ms in times)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_graph_replays_match_eager(gpu, prec):
    """Captured evaluations replayed several times, with eager evaluations and
    synchronisations in between and W changed in place, equal eager results."""
    import torch
    _say(f"[graph] {prec}")
    X, hm = make_problem(**CASES["h2b_small"])
    plan = hgemm.Plan(hm, precisions=(prec,))
    nrhs = 80
    dt = torch.float64
    Ws = [torch.from_numpy(hm.transform(host.normal((hm.N, nrhs), seed=600 + i))).to("cuda", dt) for i in range(3)]
    eager = []
    for Wi in Ws:
        Y = torch.empty_like(Wi)
        plan.eval(Wi, Y, nrhs, precision=prec)
        torch.cuda.synchronize()
        eager.append(Y.clone())
    other = hgemm.Plan(hm, precisions=(prec,))       # eager work between replays
    plan.set_graph(True)
    W = Ws[0].clone()
    Yg = torch.full_like(W, float("nan"))
    for rep in range(5):
        i = rep % 3
        W.copy_(Ws[i])                                # new contents, same buffer
        plan.eval(W, Yg, nrhs, precision=prec)        # captured on first use, replayed after
        torch.cuda.synchronize()
        assert torch.equal(Yg, eager[i]), (prec, rep)
        Ye = torch.empty_like(W)
        other.eval(Ws[(i + 1) % 3], Ye, nrhs, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(Ye, eager[(i + 1) % 3])
