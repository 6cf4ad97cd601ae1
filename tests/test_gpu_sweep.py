"""The dependency-driven sweep (MX_SCHED_SWEEP): one persistent BF16x3
launch per evaluation, work items started as their producers finish.

* against the oracle (1e-5, the BF16x3 bar) and against the level launches
  of the same split products, forward and adjoint, for the four tiles and
  column counts that are not multiples of them.  The same products summed
  in the same order up to the place of the u^ segment agree within 1e-6;
  the split-K tile sums its partial tiles in another order, and the
  re-split of every skeleton panel to 16 bits turns such last-bit
  differences into differences at the split's 2^-17 resolution: within
  5e-6;
* deterministic: every output element has one writer and a fixed sum
  order, so repeated evaluations are bit-identical;
* the skeletons it leaves (w^, u^ panels) equal the level launches', and
  the backward pass, graph replays and the allocation-free entry point use
  it with the same results;
* no wait ever ran out of time (mx_plan_sweep_status).
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = {
    "h2b_small": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0),
    "hss_small": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.0, h=1.0),
    "h2b_ragged": dict(n=3001, d=5, leaf=100, rank=70, tol=1e-7, budget=0.3, h=1.2),
    # leaves of 300 points and ranks up to 150: several 128-row tiles per task
    "h2b_tall": dict(n=4800, d=6, leaf=300, rank=150, tol=1e-9, budget=0.3, h=1.0),
}
WIDTHS = (1, 5, 6, 12)   # MX_X3_128x128_S2, MX_X3_128x64_S2, MX_X3_128x32_S2, MX_X3_128x32_K8


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _eval(plan, direction, W, nrhs):
    import torch
    Y = torch.full_like(W, float("nan"))
    if direction == "forward":
        plan.eval(W, Y, nrhs, "bf16x3")
    else:
        plan.eval_adjoint(W, Y, nrhs, "bf16x3")
    return Y


@pytest.mark.parametrize("case", sorted(CASES))
def test_sweep_matches_oracle_and_levels(gpu, case):
    import torch
    X, hm = make_problem(**CASES[case])
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    for nrhs in (45, 200):
        TW = host.normal((hm.N * nrhs,), seed=41 + nrhs)
        ref = {"forward": oracle.hgemm(hm, TW, nrhs)[0], "adjoint": oracle.hgemm(hm.transposed(), TW, nrhs)[0]}
        W = torch.from_numpy(TW).cuda()
        for direction in ("forward", "adjoint"):
            plan.set_schedule(direction, "bf16x3", "fused", 1)
            Yl = _eval(plan, direction, W, nrhs)
            for width in WIDTHS:
                _say(f"[sweep] case={case} nrhs={nrhs} {direction} width={width}")
                plan.set_schedule(direction, "bf16x3", "sweep", width)
                assert plan.get_schedule(direction, "bf16x3") == ("sweep", width)
                Y1 = _eval(plan, direction, W, nrhs)
                Y2 = _eval(plan, direction, W, nrhs)
                torch.cuda.synchronize()
                assert plan.sweep_status() == 0
                assert torch.isfinite(Y1).all()
                assert torch.equal(Y1, Y2), "the sweep is not deterministic"
                e_or = rel(Y1.cpu().numpy(), ref[direction])
                e_lv = rel(Y1.cpu().numpy(), Yl.cpu().numpy())
                _say(f"    vs oracle {e_or:.2e}, vs level launches {e_lv:.2e}")
                assert e_or <= 1e-5 and e_lv <= (5e-6 if width == 12 else 1e-6)


def test_sweep_skeletons_and_backward(gpu):
    import torch
    X, hm = make_problem(**CASES["h2b_tall"])
    nrhs = 96
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    ref = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=51)).cuda()
    dY = torch.from_numpy(host.normal((hm.N * nrhs,), seed=52)).cuda()
    for d in ("forward", "adjoint"):
        plan.set_schedule(d, "bf16x3", "sweep", 1)
        ref.set_schedule(d, "bf16x3", "fused", 1)
    _say("[sweep] skeletons")
    Y, Yr = torch.empty_like(W), torch.empty_like(W)
    plan.eval(W, Y, nrhs, "bf16x3")
    ref.eval(W, Yr, nrhs, "bf16x3")
    ws, us = plan.skeletons(nrhs, torch.float32)
    wr, ur = ref.skeletons(nrhs, torch.float32)
    torch.cuda.synchronize()
    assert plan.sweep_status() == 0
    # w^ comes from the same sums in the same order; u^ only moves the parent's transfer
    assert torch.equal(ws, wr)
    assert rel(us.cpu().numpy(), ur.cpu().numpy()) <= 1e-6
    _say("[sweep] backward")
    g = plan.backward(W, dY, nrhs, "bf16x3")
    gr = ref.backward(W, dY, nrhs, "bf16x3")
    torch.cuda.synchronize()
    assert plan.sweep_status() == 0
    for a, b in zip(g, gr):
        assert torch.isfinite(a).all()
        assert rel(a.cpu().numpy(), b.cpu().numpy()) <= 1e-6


def test_sweep_graph_replay(gpu):
    """Captured sweeps (the counter reset is a kernel, not a memset) replay
    bit for bit, with eager work between the replays, at two column counts."""
    import torch
    X, hm = make_problem(**CASES["h2b_ragged"])
    plan = hgemm.Plan(hm, precisions=("bf16x3",), adjoint=True)
    eager = hgemm.Plan(hm, precisions=("bf16x3",), adjoint=True)
    for p in (plan, eager):
        for d in ("forward", "adjoint"):
            p.set_schedule(d, "bf16x3", "sweep", 5)
    plan.set_graph(True)
    for rep in range(3):
        if rep == 2:   # the split-K tile in the sweep, captured too
            for p in (plan, eager):
                for d in ("forward", "adjoint"):
                    p.set_schedule(d, "bf16x3", "sweep", 12)
        for nrhs in (64, 130):
            _say(f"[sweep graph] rep={rep} nrhs={nrhs}")
            W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=60 + rep)).cuda()
            for d in ("forward", "adjoint"):
                Y = _eval(plan, d, W, nrhs)
                Ye = _eval(eager, d, W, nrhs)
                torch.cuda.synchronize()
                assert plan.sweep_status() == 0 and eager.sweep_status() == 0
                assert torch.equal(Y, Ye), (rep, nrhs, d)


def test_sweep_noalloc(gpu):
    """mx_eval_noalloc runs the sweep too (control block in the static
    scratch) and equals mx_eval bit for bit."""
    import torch
    X, hm = make_problem(**CASES["h2b_small"])
    plan = hgemm.Plan(hm, precisions=("bf16x3",))
    plan.set_schedule("forward", "bf16x3", "sweep", 1)
    for nrhs in (96, 40):
        _say(f"[sweep noalloc] nrhs={nrhs}")
        W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=70 + nrhs)).cuda()
        Yn = torch.full_like(W, float("nan"))
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = hgemm.lib().mx_eval_noalloc(plan._h, hgemm.BF16X3, W.data_ptr(), Yn.data_ptr(), nrhs, hgemm.IO_F64,
                                         stream)
        assert rc == 0, hgemm.lib().mx_last_error()
        Ye = _eval(plan, "forward", W, nrhs)
        torch.cuda.synchronize()
        assert torch.equal(Yn, Ye)


def test_sweep_schedule_errors(gpu):
    X, hm = make_problem(**CASES["h2b_small"])
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"))
    with pytest.raises(hgemm.MxError):
        plan.set_schedule("forward", "fp64", "sweep", 0)
    with pytest.raises(hgemm.MxError):
        plan.set_schedule("forward", "bf16x3", "sweep", 3)   # not a sweep tile width
    plan.set_schedule("forward", "bf16x3", "sweep", 6)
    plan.set_schedule("forward", "bf16x3", "split", 0)      # any level schedule turns it off
    assert plan.get_schedule("forward", "bf16x3")[0] == "split"
