"""The coupling/downward schedules (fused per level, split: all couplings
in one launch then per-level transfers accumulated with beta = 1, their
balanced forms, and the coarse lists: runs of BF16x3 launches merged over
composite blocks), both directions and every precision, against the
oracle; the FP64/FP32 kernel variants; and the autotuner."""
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
    # leaves of 300 points and ranks up to 150: three 128-row tiles per near
    # block, so the 256-row tiles of variant 9 cover a full and a half pair
    "h2b_tall": dict(n=4800, d=6, leaf=300, rank=150, tol=1e-9, budget=0.3, h=1.0),
}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("schedule", sorted(s for s in hgemm.SCHEDULES if s != "sweep"))   # sweep: test_gpu_sweep.py
def test_schedules_match_oracle(gpu, case, schedule):
    import torch
    _say(f"[schedule] case={case} schedule={schedule}")
    X, hm = make_problem(**CASES[case])
    nrhs = 45
    TW = host.normal((hm.N * nrhs,), seed=17)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    Go, _, _ = oracle.hgemm(hm.transposed(), TW, nrhs)
    plan = hgemm.Plan(hm, precisions=("fp64", "fp32", "bf16x3"), adjoint=True)
    W = torch.from_numpy(TW).cuda()
    coarse = schedule.startswith("coarse")
    precs = [("fp64", 0), ("fp32", 0)] if not coarse else []   # coarse lists: BF16x3 only
    if coarse:
        for prec in ("fp64", "fp32"):
            with pytest.raises(hgemm.MxError):
                plan.set_schedule("forward", prec, schedule)
    for prec, stages in precs + [("bf16x3", v) for v in sorted(hgemm.X3_VARIANTS)]:
        for direction in ("forward", "adjoint"):
            plan.set_schedule(direction, prec, schedule, stages)
            assert plan.get_schedule(direction, prec)[0] == schedule
        Y = torch.full_like(W, float("nan"))
        G = torch.full_like(W, float("nan"))
        plan.eval(W, Y, nrhs, prec)
        plan.eval_adjoint(W, G, nrhs, prec)
        tol = 1e-12 if prec == "fp64" else 1e-5
        ey, eg = rel(Y.cpu().numpy(), Yo), rel(G.cpu().numpy(), Go)
        _say(f"  {prec} variant={stages}: forward {ey:.2e} adjoint {eg:.2e}")
        assert ey <= tol and eg <= tol, (prec, stages)


def test_autotune_keeps_results(gpu):
    import torch
    _say("[autotune]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 64
    TW = host.normal((hm.N * nrhs,), seed=18)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    W = torch.from_numpy(TW).cuda()
    for prec in ("fp64", "bf16x3"):
        ms = plan.autotune(prec, nrhs, "f64", reps=2)
        assert len(ms) == 2 and all(m > 0 for m in ms)
        for d in ("forward", "adjoint"):
            sc, st = plan.get_schedule(d, prec)
            assert sc in hgemm.SCHEDULES and st in hgemm.X3_VARIANTS
        if prec == "bf16x3":
            lv = plan.launch_variants("forward")
            # (the coarse lists, which the autotuner may pick, can hold two launches)
            assert len(lv) >= 1 and all(v in hgemm.X3_VARIANTS for v in lv)
        Y = torch.empty_like(W)
        plan.eval(W, Y, nrhs, prec)
        assert rel(Y.cpu().numpy(), Yo) <= (1e-12 if prec == "fp64" else 1e-5)
    with pytest.raises(hgemm.MxError):
        plan.autotune("fp32", nrhs)


def test_pinned_launch_variants(gpu):
    """mx_plan_set_launch_variants: every launch on its own variant gives the
    same result as the plan-wide default, in both directions."""
    import torch
    _say("[pinned launch variants]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 64
    TW = host.normal((hm.N * nrhs,), seed=19)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    plan = hgemm.Plan(hm, precisions=("bf16x3",), adjoint=True)
    W = torch.from_numpy(TW).cuda()
    Y0, Y1 = torch.empty_like(W), torch.empty_like(W)
    for d in ("forward", "adjoint"):
        run = (lambda Y: plan.eval(W, Y, nrhs, "bf16x3")) if d == "forward" else \
              (lambda Y: plan.eval_adjoint(W, Y, nrhs, "bf16x3"))
        run(Y0)
        n = len(plan.launch_variants(d))
        pinned = [1 + (i * 3) % 8 for i in range(n)]
        plan.set_launch_variants(pinned, d)
        assert plan.launch_variants(d) == pinned
        run(Y1)
        torch.cuda.synchronize()
        e = rel(Y1.cpu().numpy(), Y0.cpu().numpy())
        _say(f"  {d}: {n} launches, pinned vs default {e:.2e}")
        assert e <= 1e-6
        with pytest.raises(hgemm.MxError):
            plan.set_launch_variants(pinned[:-1], d)
        with pytest.raises(hgemm.MxError):
            plan.set_launch_variants([max(hgemm.X3_VARIANTS) + 1] * n, d)
        plan.set_launch_variants([], d)
        assert plan.launch_variants(d) == [plan.get_schedule(d, "bf16x3")[1]] * n
    Yf = torch.empty_like(W)
    plan.eval(W, Yf, nrhs, "bf16x3")
    torch.cuda.synchronize()
    assert rel(Yf.cpu().numpy(), Yo) <= 1e-5


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp64_fp32_kernel_variants(gpu, case):
    """Every FP64/FP32 kernel variant (64x64 tiles with one or two chunks of
    loads in flight, 128x128 tiles for FP32, 128x128 and 128x256 tiles on
    packed A operands with two or three LDS stages for FP64, two or four for
    FP32) on every launch, both directions, against the oracle; ragged
    column counts (150: a partial 128- and 256-column tile)."""
    import torch
    _say(f"[seg variants] {case}")
    X, hm = make_problem(**CASES[case])
    nrhs = 150
    TW = host.normal((hm.N * nrhs,), seed=29)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    Go, _, _ = oracle.hgemm(hm.transposed(), TW, nrhs)
    plan = hgemm.Plan(hm, precisions=("fp64", "fp32"), adjoint=True)
    W = torch.from_numpy(TW).cuda()
    for prec, variants in (("fp64", (1, 2, 4, 5, 6, 7)), ("fp32", (1, 2, 3, 4, 5, 6, 7))):
        for v in variants:
            for d in ("forward", "adjoint"):
                n = len(plan.seg_variants(prec, d))
                plan.set_seg_variants([v] * n, prec, d)
                assert plan.seg_variants(prec, d) == [v] * n
            Y = torch.full_like(W, float("nan"))
            G = torch.full_like(W, float("nan"))
            plan.eval(W, Y, nrhs, prec)
            plan.eval_adjoint(W, G, nrhs, prec)
            torch.cuda.synchronize()
            tol = 1e-12 if prec == "fp64" else 1e-5
            ey, eg = rel(Y.cpu().numpy(), Yo), rel(G.cpu().numpy(), Go)
            _say(f"  {prec} variant {v}: forward {ey:.2e} adjoint {eg:.2e}")
            assert ey <= tol and eg <= tol, (prec, v)
    with pytest.raises(hgemm.MxError):
        plan.set_seg_variants([3] * len(plan.seg_variants("fp64")), "fp64")
    with pytest.raises(hgemm.MxError):
        plan.set_seg_variants([8] * len(plan.seg_variants("fp32")), "fp32")
    with pytest.raises(hgemm.MxError):
        plan.set_seg_variants([1], "fp32")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_packed_operands_follow_new_blocks(gpu, prec):
    """The packed FP64 and FP32 A operands are repacked by mx_plan_set_blocks:
    after new blocks, the packed-operand kernel (every launch) agrees with
    the 64x64 kernel reading the stored blocks, forward and adjoint, at nrhs
    1 and 300 (FP64 to 1e-13; FP32 to 2e-6: the two kernels sum the same
    products in another order).  The packed variants (128 or 256 columns,
    any stage count) sum every output in the same order: bit-identical."""
    import torch
    _say(f"[packed operands after set_blocks] {prec}")
    X, hm = make_problem(**CASES["h2b_tall"])
    plan = hgemm.Plan(hm, precisions=(prec,), adjoint=True)
    D, B, VT = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (hm.D, hm.B, hm.VT))
    plan.set_blocks(D * 1.5, B - 0.25 * B.abs(), VT * -0.75)
    dt = torch.float64 if prec == "fp64" else torch.float32
    tol = 1e-13 if prec == "fp64" else 2e-6
    for nrhs in (1, 300):
        W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=31 + nrhs)).cuda().to(dt)
        out = {}
        for v in (1, 4, 5, 6, 7):
            for d in ("forward", "adjoint"):
                plan.set_seg_variants([v] * len(plan.seg_variants(prec, d)), prec, d)
            Y = torch.full_like(W, float("nan"))
            G = torch.full_like(W, float("nan"))
            plan.eval(W, Y, nrhs, prec)
            plan.eval_adjoint(W, G, nrhs, prec)
            torch.cuda.synchronize()
            out[v] = (Y.double().cpu().numpy(), G.double().cpu().numpy())
        for v in (4, 5, 6, 7):
            ey, eg = rel(out[v][0], out[1][0]), rel(out[v][1], out[1][1])
            _say(f"  nrhs={nrhs} variant {v} vs 1: forward {ey:.2e} adjoint {eg:.2e}")
            assert ey <= tol and eg <= tol
            assert np.array_equal(out[v][0], out[4][0]) and np.array_equal(out[v][1], out[4][1]), v


def test_balanced_schedule_moves_near_work(gpu):
    """The balanced schedules move the leading near blocks of leaves with long
    near lists into the UP launches: on the transposed structure (whose row
    lists are the original's column lists, up to twice the mean here) the
    final NEAR+LEAF launch loses flops to earlier launches, the total stays,
    and the result still matches the oracle."""
    import torch
    _say("[balanced schedule]")
    X, hm = make_problem(**CASES["h2b_small"])
    hmT = hm.transposed()
    nrhs = 40
    TW = host.normal((hmT.N * nrhs,), seed=23)
    Yo, _, _ = oracle.hgemm(hmT, TW, nrhs)
    plan = hgemm.Plan(hmT, precisions=("fp64", "fp32", "bf16x3"))
    W = torch.from_numpy(TW).cuda()
    for prec in ("fp64", "fp32", "bf16x3"):
        flops = {}
        for sc in ("split", "split_balanced", "fused", "fused_balanced"):
            plan.set_schedule("forward", prec, sc)
            flops[sc] = plan.launch_flops(prec, nrhs)
            Y = torch.full_like(W, float("nan"))
            plan.eval(W, Y, nrhs, prec)
            torch.cuda.synchronize()
            assert rel(Y.cpu().numpy(), Yo) <= (1e-12 if prec == "fp64" else 1e-5), (prec, sc)
        for base in ("split", "fused"):
            f0, f1 = flops[base], flops[base + "_balanced"]
            _say(f"  {prec} {base}: NEAR+LEAF {f0[-1] / 1e6:.1f} -> {f1[-1] / 1e6:.1f} MFLOP")
            assert len(f0) == len(f1)
            assert sum(f1) == pytest.approx(sum(f0), rel=1e-12)
            assert f1[-1] < f0[-1]


def test_long_interaction_lists_are_split(gpu):
    """Blocks with more interactions than the BF16x3 kernel's segment window
    (up to 459 near blocks per row leaf, 512 per column leaf here) run as a
    task plus continuation launches; forward and adjoint still match."""
    import torch
    _say("[long lists]")
    X, hm = make_problem(n=2600, d=3, ncomp=4, leaf=10, rank=6, tol=1e-6, budget=0.9, h=0.8, seed=5)
    nx = np.bincount(hm.npairx, minlength=hm.nleaf)
    assert nx.max() > 256
    nrhs = 20
    TW = host.normal((hm.N * nrhs,), seed=31)
    Yo, _, _ = oracle.hgemm(hm, TW, nrhs)
    Go, _, _ = oracle.hgemm(hm.transposed(), TW, nrhs)
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    W = torch.from_numpy(TW).cuda()
    for prec, tol in (("fp64", 1e-12), ("bf16x3", 1e-5)):
        Y = torch.full_like(W, float("nan"))
        G = torch.full_like(W, float("nan"))
        plan.eval(W, Y, nrhs, prec)
        plan.eval_adjoint(W, G, nrhs, prec)
        ey, eg = rel(Y.cpu().numpy(), Yo), rel(G.cpu().numpy(), Go)
        _say(f"  {prec}: forward {ey:.2e} adjoint {eg:.2e}")
        assert ey <= tol and eg <= tol, prec


@pytest.mark.parametrize("case", ["h2b_small", "hss_small", "h2b_tall"])
def test_coarse_schedule_backward_and_graphs(gpu, case):
    """The coarse schedule writes every skeleton block: the block gradients
    after a coarse forward evaluation are within 5e-5 of the FP64 ones, as
    after a fused one (the bound of the split-BF16x3 gradients: operands
    split to 16 significant bits); its graph replays equal its eager
    evaluations bit for bit."""
    import torch
    _say(f"[coarse backward] {case}")
    X, hm = make_problem(**CASES[case])
    nrhs = 40
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=61)).cuda()
    dY = torch.from_numpy(host.normal((hm.N * nrhs,), seed=62)).cuda()
    Y = torch.empty_like(W)
    plan.eval(W, Y, nrhs, "fp64")
    ref = [g.clone() for g in plan.backward(W, dY, nrhs, "fp64")]
    res = {}
    for sc in ("fused", "coarse"):
        for d in ("forward", "adjoint"):
            plan.set_schedule(d, "bf16x3", sc)
        plan.eval(W, Y, nrhs, "bf16x3")
        res[sc] = [g.clone() for g in plan.backward(W, dY, nrhs, "bf16x3")]
        torch.cuda.synchronize()
    for i, nm in enumerate(["dW", "dD", "dB", "dVT"]):
        r = ref[i].double().cpu().numpy()
        ef = rel(res["fused"][i].double().cpu().numpy(), r)
        ec = rel(res["coarse"][i].double().cpu().numpy(), r)
        _say(f"  {nm}: vs FP64 fused {ef:.2e} coarse {ec:.2e}")
        assert ec <= 5e-5 and ef <= 5e-5, nm
    eager = torch.empty_like(W)
    plan.set_graph(False)
    plan.eval(W, eager, nrhs, "bf16x3")
    plan.set_graph(True)
    for _ in range(3):
        Yg = torch.full_like(W, float("nan"))
        plan.eval(W, Yg, nrhs, "bf16x3")
        torch.cuda.synchronize()
        assert torch.equal(Yg, eager)
    plan.set_graph(False)


def test_coarse_composites_follow_new_blocks(gpu):
    """mx_plan_set_blocks forms the coarse schedule's composites again, from
    the FP64 B and VT the plan keeps: after all new blocks, and after new B
    alone, the coarse BF16x3 evaluation agrees with the FP64 one to 1e-5."""
    import torch
    _say("[coarse after set_blocks]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 33
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"))
    plan.set_schedule("forward", "bf16x3", "coarse")
    W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=63)).cuda()
    D, B, VT = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (hm.D, hm.B, hm.VT))
    for blocks in ((D * 1.5, B - 0.25 * B.abs(), VT * -0.75), (None, B * 0.5, None)):
        plan.set_blocks(*blocks)
        Y64 = torch.empty_like(W)
        Yc = torch.empty_like(W)
        plan.eval(W, Y64, nrhs, "fp64")
        plan.eval(W, Yc, nrhs, "bf16x3")
        torch.cuda.synchronize()
        e = rel(Yc.cpu().numpy(), Y64.cpu().numpy())
        _say(f"  coarse vs fp64: {e:.2e}")
        assert e <= 1e-5
