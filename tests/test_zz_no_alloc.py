"""mx_eval_noalloc: the allocation-free evaluation equals mx_eval bit for
bit (same kernels, work buffers in the library's static scratch array), on
consecutive calls whose second input is smaller, and refuses evaluations
larger than the scratch array without touching the output."""
import ctypes
import sys

import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

pytestmark = pytest.mark.gpu

ARG = -1   # MX_ERR_ARG
BIG = dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0)
SMALL = dict(n=700, d=3, leaf=40, rank=20, tol=1e-6, budget=0.0, h=1.1, seed=4)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(plan, prec, W, Y, nrhs):
    """mx_eval_noalloc called through the C-ABI."""
    import torch
    io = hgemm.IO_F64 if W.dtype == torch.float64 else hgemm.IO_F32
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return hgemm.lib().mx_eval_noalloc(plan._h, hgemm.PRECISIONS[prec], W.data_ptr(), Y.data_ptr(), nrhs, io, stream)


@pytest.mark.parametrize("prec,io", [("fp64", "f64"), ("fp32", "f64"), ("fp32", "f32"), ("bf16x3", "f64"),
                                     ("bf16x3", "f32")])
def test_noalloc_matches_eval(gpu, prec, io):
    import torch
    dt = torch.float64 if io == "f64" else torch.float32
    cases = [(BIG, 96), (BIG, 40), (SMALL, 24)]   # each call smaller than the one before
    plans = {}
    for step, (kw, nrhs) in enumerate(cases):
        _say(f"[noalloc] {prec} io={io} case n={kw['n']} nrhs={nrhs}")
        X, hm = make_problem(**kw)
        key = kw["n"]
        if key not in plans:
            plans[key] = hgemm.Plan(hm, precisions=("fp64", "fp32", "bf16x3"))
        plan = plans[key]
        TW = hm.transform(host.normal((hm.N, nrhs), seed=7300 + step))
        W = torch.from_numpy(TW).to("cuda", dt)
        Yn = torch.full_like(W, float("nan"))
        Ye = torch.full_like(W, float("nan"))
        assert _noalloc(plan, prec, W, Yn, nrhs) == 0, hgemm.lib().mx_last_error()
        plan.eval(W, Ye, nrhs, precision=prec)
        torch.cuda.synchronize()
        assert torch.isfinite(Yn).all()
        assert torch.equal(Yn, Ye), (prec, io, step)
        Yo = oracle.hgemm(hm, TW, nrhs)[0]
        bar = 1e-12 if prec == "fp64" else 1e-5
        assert rel(Yn.double().cpu().numpy(), Yo) <= bar


def test_noalloc_rejects_oversized_evaluations(gpu):
    import torch
    X, hm = make_problem(**BIG)
    plan = hgemm.Plan(hm, precisions=("fp64",))
    # FP64 needs two skeleton buffers of skel_rows * nrhs doubles
    nrhs = hgemm.NOALLOC_SCRATCH_BYTES // (16 * plan.skel_rows) + 8
    _say(f"[noalloc] oversized fp64 nrhs={nrhs} skel_rows={plan.skel_rows}")
    W = torch.zeros(plan.rows * nrhs, dtype=torch.float64, device="cuda")
    Y = torch.full_like(W, 7.0)
    rc = _noalloc(plan, "fp64", W, Y, nrhs)
    torch.cuda.synchronize()
    assert rc == ARG
    assert b"work space" in hgemm.lib().mx_last_error()
    assert bool((Y == 7.0).all())
    # the largest evaluation that fits still runs, and equals mx_eval
    nrhs_fit = hgemm.NOALLOC_SCRATCH_BYTES // (16 * plan.skel_rows) - 16
    W = torch.from_numpy(host.normal((plan.rows * nrhs_fit,), seed=7400)).cuda()
    Yn = torch.empty_like(W)
    Ye = torch.empty_like(W)
    assert _noalloc(plan, "fp64", W, Yn, nrhs_fit) == 0, hgemm.lib().mx_last_error()
    plan.eval(W, Ye, nrhs_fit)
    torch.cuda.synchronize()
    assert torch.equal(Yn, Ye)
    assert np.isfinite(Yn[:1000].cpu().numpy()).all()
