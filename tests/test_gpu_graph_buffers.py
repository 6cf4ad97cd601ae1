"""Graph replay across work-buffer reallocation, and the skeleton state that
graph replays leave behind.

A plan's work buffers (skeletons, FP32 staging, split panels) grow when a
call needs more room than they hold; graphs captured before such a growth
would write through freed addresses, so a plan must re-capture instead of
replaying them.  Every result below is compared bit for bit with an eager
evaluation on a second plan of the same matrix (the kernels are
deterministic: every output block has one writer).
"""
import sys

import numpy as np
import pytest

from conftest import make_problem
from matrox_ru_amd import hgemm, host

pytestmark = pytest.mark.gpu

CASE = dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0)
STATE = -3   # MX_ERR_STATE

# (precision, direction, nrhs, I/O dtype name): precision changes and nrhs
# growth both reallocate the work buffers between replays of one graph
SEQUENCE = [
    ("fp32", "forward", 64, "f64"),
    ("fp64", "forward", 64, "f64"),     # skeletons grow from 4 to 8 bytes
    ("fp32", "forward", 64, "f64"),
    ("fp32", "forward", 256, "f64"),    # nrhs growth
    ("fp32", "forward", 64, "f64"),
    ("fp32", "forward", 256, "f64"),
    ("fp32", "adjoint", 64, "f64"),
    ("fp64", "adjoint", 256, "f64"),    # adjoint skeletons grow
    ("fp32", "adjoint", 64, "f64"),
    ("fp64", "forward", 64, "f64"),
    ("bf16x3", "forward", 64, "f64"),
    ("bf16x3", "forward", 256, "f64"),  # split panels grow
    ("bf16x3", "forward", 64, "f64"),
    ("bf16x3", "adjoint", 256, "f64"),
    ("bf16x3", "adjoint", 64, "f64"),
    ("fp32", "forward", 64, "f32"),
    ("fp32", "forward", 512, "f32"),    # staging-free growth
    ("fp32", "forward", 64, "f32"),
    ("bf16x3", "forward", 256, "f64"),
    ("fp64", "forward", 256, "f64"),
    ("fp32", "adjoint", 64, "f64"),
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(plan, prec, direction, W, Y, nrhs):
    if direction == "forward":
        plan.eval(W, Y, nrhs, precision=prec)
    else:
        plan.eval_adjoint(W, Y, nrhs, precision=prec)


def test_graph_replay_across_buffer_growth(gpu):
    import torch
    X, hm = make_problem(**CASE)
    precs = ("fp64", "fp32", "bf16x3")
    plan = hgemm.Plan(hm, precisions=precs, adjoint=True)
    ref = hgemm.Plan(hm, precisions=precs, adjoint=True)
    plan.set_graph(True)
    bufs = {}
    for rep in range(2):   # the second pass replays every graph that is still valid
        for step, (prec, direction, nrhs, io) in enumerate(SEQUENCE):
            _say(f"[graph buffers] pass={rep} step={step} {prec} {direction} nrhs={nrhs} io={io}")
            dt = torch.float64 if io == "f64" else torch.float32
            key = (nrhs, io)
            if key not in bufs:   # one W and one Y per shape: captured pointers repeat
                bufs[key] = (torch.empty(hm.N * nrhs, dtype=dt, device="cuda"),
                             torch.empty(hm.N * nrhs, dtype=dt, device="cuda"))
            W, Y = bufs[key]
            W.copy_(torch.from_numpy(host.normal((hm.N * nrhs,), seed=7000 + 100 * rep + step)).to(dt))
            Y.fill_(float("nan"))
            _run(plan, prec, direction, W, Y, nrhs)
            Yr = torch.full_like(Y, float("nan"))
            _run(ref, prec, direction, W, Yr, nrhs)
            torch.cuda.synchronize()
            assert torch.isfinite(Yr).all()
            assert torch.equal(Y, Yr), (rep, step, prec, direction, nrhs, io)


def test_skeleton_state_follows_replays(gpu):
    """Forward and adjoint BF16x3 evaluations share the split panels: after
    an adjoint replay the forward skeletons are gone, after a forward replay
    they are back, whichever order the graphs were captured in."""
    import torch
    X, hm = make_problem(**CASE)
    plan = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    ref = hgemm.Plan(hm, precisions=("fp64", "bf16x3"), adjoint=True)
    plan.set_graph(True)
    nrhs = 64
    W = torch.from_numpy(host.normal((hm.N * nrhs,), seed=71)).cuda()
    dY = torch.from_numpy(host.normal((hm.N * nrhs,), seed=72)).cuda()
    Y = torch.empty_like(W)
    dW = torch.empty_like(W)
    _say("[skeleton state] bf16x3 adjoint capture, forward capture, adjoint replay")
    plan.eval_adjoint(dY, dW, nrhs, "bf16x3")
    plan.eval(W, Y, nrhs, "bf16x3")
    plan.eval_adjoint(dY, dW, nrhs, "bf16x3")
    torch.cuda.synchronize()
    with pytest.raises(hgemm.MxError) as e:
        plan.skeletons(nrhs, torch.float32)
    assert e.value.code == STATE
    with pytest.raises(hgemm.MxError) as e:
        plan.backward(W, dY, nrhs, "bf16x3")
    assert e.value.code == STATE
    _say("[skeleton state] bf16x3 forward replay, then skeletons and backward")
    plan.eval(W, Y, nrhs, "bf16x3")
    ws, us = plan.skeletons(nrhs, torch.float32)
    Yr = torch.empty_like(W)
    ref.eval(W, Yr, nrhs, "bf16x3")
    wr, ur = ref.skeletons(nrhs, torch.float32)
    torch.cuda.synchronize()
    assert torch.equal(Y, Yr) and torch.equal(ws, wr) and torch.equal(us, ur)
    g = plan.backward(W, dY, nrhs, "bf16x3")
    gr = ref.backward(W, dY, nrhs, "bf16x3")
    torch.cuda.synchronize()
    for a, b in zip(g, gr):
        assert torch.equal(a, b)

    _say("[skeleton state] fp64 forward replay, then a request at another nrhs")
    plan.eval(W, Y, nrhs, "fp64")
    plan.eval(W, Y, nrhs, "fp64")   # replay
    ref.eval(W, Yr, nrhs, "fp64")
    torch.cuda.synchronize()
    assert torch.equal(Y, Yr)
    ws, us = plan.skeletons(nrhs, torch.float64)
    wr, ur = ref.skeletons(nrhs, torch.float64)
    assert torch.equal(ws, wr) and torch.equal(us, ur)
    with pytest.raises(hgemm.MxError) as e:
        plan.skeletons(nrhs // 2, torch.float64)
    assert e.value.code == STATE
    half = nrhs // 2
    with pytest.raises(hgemm.MxError) as e:
        plan.backward(W[: hm.N * half], dY[: hm.N * half], half, "fp64")
    assert e.value.code == STATE
    g = plan.backward(W, dY, nrhs, "fp64")
    gr = ref.backward(W, dY, nrhs, "fp64")
    torch.cuda.synchronize()
    for a, b in zip(g, gr):
        assert torch.equal(a, b)


def test_batch_graph_across_buffer_growth(gpu):
    """A batch graph is re-captured when any group plan's buffers grow,
    including growth caused by evaluating a group plan on its own."""
    import torch
    specs = [(dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0), 32),
             (dict(n=1500, d=6, leaf=50, rank=30, tol=1e-6, budget=0.1, h=0.9, seed=3), 8),
             (dict(n=700, d=3, leaf=40, rank=20, tol=1e-6, budget=0.0, h=1.1, seed=4), 8)]
    hms = [make_problem(**kw)[1] for kw, _ in specs]
    nrhs = [n for _, n in specs]
    precs = ("fp64", "fp32", "bf16x3")
    b = hgemm.Batch(hms, nrhs, precisions=precs, adjoint=True)
    r = hgemm.Batch(hms, nrhs, precisions=precs, adjoint=True)
    b.set_graph(True)
    W64, Y64 = b.buffers(torch.float64), b.buffers(torch.float64)
    seq = ["fp32", "fp64", "fp32", "grow", "fp32", "bf16x3", "grow_x3", "bf16x3", "fp32-adjoint", "fp64-adjoint",
           "fp32-adjoint", "fp32"]
    for step, what in enumerate(seq):
        _say(f"[batch graph buffers] step={step} {what}")
        if what.startswith("grow"):
            g0 = b.plan(0)
            nr, rows = b.groups[0]
            Wg = torch.from_numpy(host.normal((rows * 4 * nr,), seed=80 + step)).cuda()
            Yg = torch.empty_like(Wg)
            g0.eval(Wg, Yg, 4 * nr, precision="bf16x3" if what == "grow_x3" else "fp64")
            continue
        prec, _, adj = what.partition("-")
        for i, w in enumerate(W64):
            w.copy_(torch.from_numpy(host.normal((w.numel(),), seed=8100 + 10 * step + i)))
        for y in Y64:
            y.fill_(float("nan"))
        Yr = [torch.full_like(y, float("nan")) for y in Y64]
        if adj:
            b.eval_adjoint(W64, Y64, prec)
            r.eval_adjoint(W64, Yr, prec)
        else:
            b.eval(W64, Y64, prec)
            r.eval(W64, Yr, prec)
        torch.cuda.synchronize()
        for y, yr in zip(Y64, Yr):
            assert torch.equal(y, yr), (step, what)
    # a batch backward after a replayed forward takes that forward's skeletons
    b.eval(W64, Y64, "fp32")
    r.eval(W64, [torch.empty_like(y) for y in Y64], "fp32")
    dY = [torch.from_numpy(host.normal((y.numel(),), seed=8900 + i)).cuda() for i, y in enumerate(Y64)]
    got = b.backward(W64, dY, "fp32")
    exp = r.backward(W64, dY, "fp32")
    torch.cuda.synchronize()
    for a, e in zip(got, exp):
        for x, y in zip(a, e):
            assert torch.equal(x, y)
    assert np.isfinite(got[0][0].cpu().numpy()).all()
