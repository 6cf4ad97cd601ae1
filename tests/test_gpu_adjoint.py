"""GPU adjoint and gradients through the C-ABI.

* dW = K~^T dY (mx_eval_adjoint) against the CPU oracle applied to the
  transposed structure (pair lists swapped, D and B transposed, same VT):
  <= 1e-12 relative Frobenius error in FP64; FP32 and split-BF16x3 within
  1e-5 of the FP64 adjoint; the identity <K~x, y> = <x, K~^T y> to 1e-13.
* Gradients of L = 1/2 |K~W - T|^2 with respect to W, D, B and VT
  (mx_backward behind torch.autograd) against FP64 central finite
  differences with eps = 1e-6 on 64 random directions: 16 in W only, 16 in
  D only, 16 in B only, 16 in VT only (a mistake in one block type cannot
  hide behind the others).  L is quadratic in W, D and B, so the central
  difference is exact up to rounding there; in VT it is quartic and the
  truncation term is O(eps^2).
"""
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
    # blocks above 128 rows: several ragged 128 x 128 gradient tiles per block
    "h2b_wide": dict(n=3200, d=6, leaf=200, rank=150, tol=1e-9, budget=0.3, h=1.0),
}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("nrhs", [1, 37, 130])
def test_adjoint_fp64_matches_oracle_on_transposed_structure(gpu, case, nrhs):
    import torch
    _say(f"[adjoint fp64] case={case} nrhs={nrhs}")
    X, hm = make_problem(**CASES[case])
    dY = host.normal((hm.N * nrhs,), seed=300 + nrhs)
    Wo, _, _ = oracle.hgemm(hm.transposed(), dY, nrhs)
    plan = hgemm.Plan(hm, precisions=("fp64",), adjoint=True)
    dW = torch.full((hm.N * nrhs,), float("nan"), dtype=torch.float64, device="cuda")
    plan.eval_adjoint(_dev(dY), dW, nrhs, "fp64")
    g = dW.cpu().numpy()
    assert np.isfinite(g).all()
    assert rel(g, Wo) <= 1e-12


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_identity(gpu, case):
    import torch
    _say(f"[adjoint identity] case={case}")
    X, hm = make_problem(**CASES[case])
    nrhs = 9
    x = _dev(host.normal((hm.N * nrhs,), seed=1))
    y = _dev(host.normal((hm.N * nrhs,), seed=2))
    plan = hgemm.Plan(hm, precisions=("fp64",), adjoint=True)
    Kx = torch.empty_like(x)
    Kty = torch.empty_like(y)
    plan.eval(x, Kx, nrhs, "fp64")
    plan.eval_adjoint(y, Kty, nrhs, "fp64")
    a = float((Kx * y).sum())
    b = float((x * Kty).sum())
    scale = float(Kx.norm() * y.norm())
    assert abs(a - b) <= 1e-13 * scale


@pytest.mark.parametrize("prec,io", [("fp32", "f64"), ("fp32", "f32"), ("bf16x3", "f64"), ("bf16x3", "f32")])
@pytest.mark.parametrize("case", ["h2b_small", "hss_small"])
def test_adjoint_mixed_within_1e5(gpu, case, prec, io):
    import torch
    _say(f"[adjoint mixed] case={case} prec={prec} io={io}")
    X, hm = make_problem(**CASES[case])
    nrhs = 96
    dY = _dev(host.normal((hm.N * nrhs,), seed=5))
    plan = hgemm.Plan(hm, precisions=("fp64", prec), adjoint=True)
    ref = torch.empty_like(dY)
    plan.eval_adjoint(dY, ref, nrhs, "fp64")
    dt = torch.float64 if io == "f64" else torch.float32
    dYm = dY.to(dt)
    out = torch.full_like(dYm, float("nan"))
    plan.eval_adjoint(dYm, out, nrhs, prec)
    ref_in = ref
    if io == "f32":   # compare with the FP64 adjoint of the rounded input
        plan.eval_adjoint(dYm.double(), ref_in, nrhs, "fp64")
    assert rel(out.double().cpu().numpy(), ref_in.cpu().numpy()) <= 1e-5


def test_adjoint_and_forward_interleave_and_graph_replay(gpu):
    import torch
    _say("[adjoint graph]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 40
    W = _dev(host.normal((hm.N * nrhs,), seed=21))
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    for prec in ("fp64", "bf16x3"):
        Y0 = torch.empty_like(W)
        G0 = torch.empty_like(W)
        plan.set_graph(False)
        plan.eval(W, Y0, nrhs, prec)
        plan.eval_adjoint(W, G0, nrhs, prec)
        plan.set_graph(True)
        Y1 = torch.empty_like(W)
        G1 = torch.empty_like(W)
        for _ in range(3):
            plan.eval(W, Y1, nrhs, prec)
            plan.eval_adjoint(W, G1, nrhs, prec)
            torch.cuda.synchronize()
            assert torch.equal(Y1, Y0) and torch.equal(G1, G0)
        plan.set_graph(False)


def _loss_fd(plan, W, T, D, B, VT, nrhs, v, eps):
    """Central difference (L(x + eps v) - L(x - eps v)) / (2 eps) of
    L = 1/2|K~W - T|^2 along v = (vW, vD, vB, vVT), FP64.  The difference of
    the two losses is summed as 1/2 sum (Y+ - Y-)(Y+ + Y- - 2T), which is the
    same number without the cancellation of two large sums."""
    import torch
    Ys = []
    for s in (1.0, -1.0):
        plan.set_blocks(D + s * eps * v[1], B + s * eps * v[2], VT + s * eps * v[3])
        Wp = W + s * eps * v[0]
        Y = torch.empty_like(Wp)
        plan.eval(Wp, Y, nrhs, "fp64")
        Ys.append(Y)
    plan.set_blocks(D, B, VT)
    return 0.5 * float(((Ys[0] - Ys[1]) * (Ys[0] + Ys[1] - 2 * T)).sum()) / (2 * eps)


def _directions(hm, nrhs, seed, n_each=16):
    """16 random directions in each of W, D, B, VT (unit-norm pieces)."""
    import torch
    lens = [hm.N * nrhs, len(hm.D), len(hm.B), len(hm.VT)]
    dirs = []
    for g in range(4):
        for i in range(n_each):
            v = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in lens]
            if lens[g] == 0:
                continue
            r = _dev(host.normal((lens[g],), seed=seed + 1000 * g + i))
            v[g] = r / r.norm()
            dirs.append((g, v))
    return dirs


@pytest.mark.parametrize("case", ["h2b_small", "hss_small", "h2b_ragged"])
def test_gradients_match_fp64_finite_differences(gpu, case):
    import torch
    from matrox_ru_amd.autograd import hgemm_apply
    _say(f"[grad fd] case={case}")
    X, hm = make_problem(**CASES[case])
    nrhs = 16
    plan = hgemm.Plan(hm, precisions=("fp64",), adjoint=True)
    W0 = _dev(host.normal((hm.N * nrhs,), seed=41))
    T = _dev(host.normal((hm.N * nrhs,), seed=42))
    D0, B0, VT0 = _dev(hm.D), _dev(hm.B), _dev(hm.VT)
    W = W0.clone().requires_grad_()
    D = D0.clone().requires_grad_()
    B = B0.clone().requires_grad_()
    VT = VT0.clone().requires_grad_()
    Y = hgemm_apply(plan, W, D, B, VT, nrhs=nrhs)
    L = 0.5 * ((Y - T) ** 2).sum()
    L.backward()
    grads = [W.grad, D.grad, B.grad, VT.grad]
    # dW against the oracle on the transposed structure
    dY = (Y.detach() - T).cpu().numpy()
    Wo, _, _ = oracle.hgemm(hm.transposed(), dY, nrhs)
    assert rel(W.grad.cpu().numpy(), Wo) <= 1e-12
    eps = 1e-6
    by_group = {0: ([], []), 1: ([], []), 2: ([], []), 3: ([], [])}
    for g, v in _directions(hm, nrhs, seed=500):
        an = float(sum((gr * vi).sum() for gr, vi in zip(grads, v)))
        fd = _loss_fd(plan, W0, T, D0, B0, VT0, nrhs, v, eps)
        by_group[g][0].append(an)
        by_group[g][1].append(fd)
    names = ["W", "D", "B", "VT"]
    for g, (an, fd) in by_group.items():
        if not an:
            continue
        e = rel(an, fd)
        _say(f"  {names[g]}: rel err {e:.2e} over {len(an)} directions")
        assert e <= 1e-6, names[g]


@pytest.mark.parametrize("io", ["f64", "f32"])
def test_fp32_gradients_within_1e5_of_fp64(gpu, io):
    import torch
    _say(f"[grad fp32] io={io}")
    X, hm = make_problem(**CASES["h2b_ragged"])
    nrhs = 64
    plan = hgemm.Plan(hm, precisions=("fp64", "fp32"), adjoint=True)
    dt = torch.float64 if io == "f64" else torch.float32
    W = _dev(host.normal((hm.N * nrhs,), seed=61)).to(dt)
    dY = _dev(host.normal((hm.N * nrhs,), seed=62)).to(dt)
    Y = torch.empty_like(W.double())
    plan.eval(W.double(), Y, nrhs, "fp64")
    ref = plan.backward(W.double(), dY.double(), nrhs, "fp64")
    Yf = torch.empty_like(W)
    plan.eval(W, Yf, nrhs, "fp32")
    got = plan.backward(W, dY, nrhs, "fp32")
    for a, b, nm in zip(got, ref, ["dW", "dD", "dB", "dVT"]):
        assert a.dtype == dt
        e = rel(a.double().cpu().numpy(), b.cpu().numpy())
        _say(f"  {nm}: {e:.2e}")
        assert e <= 1e-5, nm


def test_backward_explicit_skeletons_equal_plan_skeletons(gpu):
    import torch
    _say("[grad skeletons]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 24
    plan = hgemm.Plan(hm, precisions=("fp64",), adjoint=True)
    W = _dev(host.normal((hm.N * nrhs,), seed=71))
    dY = _dev(host.normal((hm.N * nrhs,), seed=72))
    Y = torch.empty_like(W)
    plan.eval(W, Y, nrhs, "fp64")
    ws, us = plan.skeletons(nrhs, torch.float64)
    a = plan.backward(W, dY, nrhs, "fp64")
    # another forward evaluation clobbers the plan's skeletons; explicit ones still hold
    plan.eval(dY, Y, nrhs, "fp64")
    b = plan.backward(W, dY, nrhs, "fp64", skeletons=(ws, us))
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    only_w = plan.backward(W, dY, nrhs, "fp64", skeletons=(ws, us), blocks=(False, False, False))
    assert torch.equal(only_w[0], a[0]) and only_w[1:] == (None, None, None)


def test_set_blocks_rederives_every_precision(gpu):
    import torch
    _say("[set blocks]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 33
    plan = hgemm.Plan(hm, precisions=("fp64", "fp32", "bf16x3"), adjoint=True)
    scale = np.array([0.5, -2.0, 1.5])
    hm2 = host.HMatrix(scalars=dict(hm.scalars), arrays=dict(hm.arrays), config=dict(hm.config))
    hm2.arrays["D"] = hm.D * scale[0]
    hm2.arrays["B"] = hm.B * scale[1]
    hm2.arrays["VT"] = hm.VT * scale[2]
    TW = host.normal((hm.N * nrhs,), seed=81)
    Yo, _, _ = oracle.hgemm(hm2, TW, nrhs)
    Go, _, _ = oracle.hgemm(hm2.transposed(), TW, nrhs)
    plan.set_blocks(_dev(hm2.D), _dev(hm2.B), _dev(hm2.VT))
    fresh = hgemm.Plan(hm2, precisions=("fp64", "fp32", "bf16x3"), adjoint=True)
    W = _dev(TW)
    for prec in ("fp64", "fp32", "bf16x3"):
        Y, G, Yf, Gf = (torch.empty_like(W) for _ in range(4))
        plan.eval(W, Y, nrhs, prec)
        plan.eval_adjoint(W, G, nrhs, prec)
        fresh.eval(W, Yf, nrhs, prec)
        fresh.eval_adjoint(W, Gf, nrhs, prec)
        # the same derived blocks as a plan built from the new arrays: same bits
        assert torch.equal(Y, Yf) and torch.equal(G, Gf), prec
        if prec == "fp64":
            assert rel(Y.cpu().numpy(), Yo) <= 1e-12
            assert rel(G.cpu().numpy(), Go) <= 1e-12


def test_backward_errors(gpu):
    import torch
    _say("[grad errors]")
    X, hm = make_problem(**CASES["h2b_small"])
    nrhs = 4
    W = _dev(host.normal((hm.N * nrhs,), seed=91))
    plain = hgemm.Plan(hm, precisions=("fp64",))
    with pytest.raises(hgemm.MxError) as e:
        plain.eval_adjoint(W, torch.empty_like(W), nrhs, "fp64")
    assert e.value.code == -4
    with pytest.raises(hgemm.MxError) as e:
        plain.backward(W, W, nrhs, "fp64")
    assert e.value.code == -4
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    with pytest.raises(hgemm.MxError) as e:   # no forward evaluation to take skeletons from
        plan.backward(W, W, nrhs, "bf16x3")
    assert e.value.code == -3
    with pytest.raises(hgemm.MxError) as e:
        plan.backward(W, W, nrhs, "fp64")
    assert e.value.code == -3
    with pytest.raises(hgemm.MxError) as e:   # not built into the plan
        plan.backward(W, W, nrhs, "fp32")
    assert e.value.code == -4
    # dW aliasing W or dY with block gradients requested: the products would
    # read the overwritten operand
    dD = torch.empty(plan.block_lens[0], dtype=W.dtype, device=W.device)
    dY = W.clone()
    for dw in (W, dY):
        rc = hgemm.lib().mx_backward(plan._h, hgemm.PRECISIONS["fp64"], W.data_ptr(), dY.data_ptr(),
                                     dw.data_ptr(), nrhs, hgemm.IO_F64, None, None, dD.data_ptr(), None, None, None)
        assert rc == -1
    D_host = np.ascontiguousarray(hm.D)
    import ctypes
    rc = hgemm.lib().mx_plan_set_blocks(plan._h, D_host.ctypes.data, None, None, None)
    assert rc == -1


@pytest.mark.parametrize("case,nrhs,io", [("h2b_ragged", 48, "f64"), ("h2b_ragged", 48, "f32"),
                                          ("h2b_ragged", 1, "f32"), ("h2b_wide", 160, "f64"),
                                          ("h2b_wide", 77, "f32")])
def test_bf16x3_backward_close_to_fp64(gpu, case, nrhs, io):
    """Split-BF16x3 forward, adjoint and block products: gradients within
    5e-5 of the FP64 ones (operands split to 16 significant bits carry ~1e-5
    relative error), and the FP32 skeleton export of a BF16x3 evaluation
    against the FP64 skeletons.  nrhs not a multiple of the 32-column chunk
    and blocks of more than 128 rows exercise the ragged tile edges."""
    import torch
    _say(f"[grad bf16x3] {case} nrhs={nrhs} io={io}")
    X, hm = make_problem(**CASES[case])
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    dt = torch.float64 if io == "f64" else torch.float32
    W = _dev(host.normal((hm.N * nrhs,), seed=161)).to(dt)
    dY = _dev(host.normal((hm.N * nrhs,), seed=162)).to(dt)
    Y64 = torch.empty_like(W.double())
    plan.eval(W.double(), Y64, nrhs, "fp64")
    ws64, us64 = plan.skeletons(nrhs, torch.float64)
    ref = plan.backward(W.double(), dY.double(), nrhs, "fp64")
    Y = torch.empty_like(W)
    plan.eval(W, Y, nrhs, "bf16x3")
    ws, us = plan.skeletons(nrhs, torch.float32)
    assert rel(ws.double().cpu().numpy(), ws64.cpu().numpy()) <= 2e-5
    assert rel(us.double().cpu().numpy(), us64.cpu().numpy()) <= 2e-5
    got = plan.backward(W, dY, nrhs, "bf16x3")
    # the same products from caller FP32 skeletons (packed from the column
    # layout instead of the plan's panels)
    got2 = plan.backward(W, dY, nrhs, "bf16x3", skeletons=(ws, us))
    for a, a2, b, nm in zip(got, got2, ref, ["dW", "dD", "dB", "dVT"]):
        assert a.dtype == dt
        e = rel(a.double().cpu().numpy(), b.cpu().numpy())
        e2 = rel(a2.double().cpu().numpy(), b.cpu().numpy())
        _say(f"  {nm}: {e:.2e} (explicit skeletons {e2:.2e})")
        assert e <= 5e-5 and e2 <= 5e-5, nm
    # one block gradient at a time packs only its operands; same products
    for k in range(3):
        flags = tuple(i == k for i in range(3))
        sub = plan.backward(W, dY, nrhs, "bf16x3", skeletons=(ws, us), blocks=flags)
        assert all(sub[1 + i] is None for i in range(3) if i != k)
        assert torch.equal(sub[1 + k], got2[1 + k]), k


@pytest.mark.parametrize("case,nrhs,io", [("h2b_ragged", 48, "f32"), ("h2b_wide", 77, "f64"),
                                          ("h2b_wide", 160, "f32")])
def test_bf16x3_gradient_tiles_agree(gpu, case, nrhs, io):
    """The 128 x 128 and 256 x 256 tilings of the split-BF16x3 block
    gradients give bit-identical results (each output element takes the same
    products in the same order); 0 picks the tile by padded work: 128 for
    blocks of about 100 rows, 256 for blocks of 150-200 rows, whose 128 x 128
    tiles pad to the same 256 x 256 area."""
    import torch
    _say(f"[grad tiles] {case} nrhs={nrhs} io={io}")
    X, hm = make_problem(**CASES[case])
    plan = hgemm.Plan(hm, precisions=("bf16x3",), adjoint=True)
    assert plan.grad_tile() == (128 if case == "h2b_ragged" else 256)
    dt = torch.float64 if io == "f64" else torch.float32
    W = _dev(host.normal((hm.N * nrhs,), seed=171)).to(dt)
    dY = _dev(host.normal((hm.N * nrhs,), seed=172)).to(dt)
    Y = torch.empty_like(W)
    res = {}
    for tile in (128, 256):
        plan.eval(W, Y, nrhs, "bf16x3")   # the skeletons the block gradients read
        plan.set_grad_tile(tile)
        assert plan.grad_tile() == tile
        res[tile] = plan.backward(W, dY, nrhs, "bf16x3")
        torch.cuda.synchronize()
    for a, b, nm in zip(res[128], res[256], ["dW", "dD", "dB", "dVT"]):
        assert torch.isfinite(b).all(), nm
        assert torch.equal(a, b), nm
    plan.set_grad_tile(0)
    with pytest.raises(hgemm.MxError):
        plan.set_grad_tile(64)
