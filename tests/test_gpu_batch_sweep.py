"""Batched sweep (mx_batch_set_sweep): every BF16x3 group of a batch, whatever
its nrhs, evaluated by one dependency-driven launch.

Each instance against the oracle (1e-5), every group against the batch's
per-group launches of the same split products (1e-6; only the place of the
u^ segment in each sum differs), forward and adjoint, the three tile
widths; repeated evaluations and graph replays bit for bit; a batch whose
accuracy guard routes some instances to FP32 (those groups run beside the
sweep); the batch autotuner's choice.
"""
import sys

import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

pytestmark = pytest.mark.gpu

INSTANCES = [
    # (builder arguments, nrhs): three nrhs groups, H2-b and HSS, ragged sizes
    (dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0), 32),
    (dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.0, h=1.0), 32),
    (dict(n=3001, d=5, leaf=100, rank=70, tol=1e-7, budget=0.3, h=1.2), 8),
    (dict(n=1500, d=6, leaf=50, rank=30, tol=1e-6, budget=0.1, h=0.9, seed=3), 200),
    (dict(n=1500, d=6, leaf=50, rank=30, tol=1e-6, budget=0.0, h=0.9, seed=5), 8),
    (dict(n=4800, d=6, leaf=300, rank=150, tol=1e-9, budget=0.3, h=1.0), 200),
]
# an instance whose BF16x3 error (1.1e-5, with either launch form) exceeds
# the bound: the guarded batch routes it to an FP32 group
GUARDED = INSTANCES + [(dict(n=700, d=3, leaf=40, rank=20, tol=1e-6, budget=0.0, h=1.1, seed=4), 8)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case():
    hms = [make_problem(**kw)[1] for kw, _ in INSTANCES]
    nrhs = [n for _, n in INSTANCES]
    TW = [host.normal((hm.N * n,), seed=1900 + i) for i, (hm, n) in enumerate(zip(hms, nrhs))]
    ref = {"forward": [oracle.hgemm(hm, tw, n)[0] for hm, tw, n in zip(hms, TW, nrhs)],
           "adjoint": [oracle.hgemm(hm.transposed(), tw, n)[0] for hm, tw, n in zip(hms, TW, nrhs)]}
    return hms, nrhs, TW, ref


def _inputs(b, TW):
    import torch
    W = b.buffers(torch.float64, fill=float("nan"))
    for i, tw in enumerate(TW):
        b.instance_view(W, i).copy_(torch.from_numpy(tw))
    return W


def _run(b, direction, W):
    import torch
    Y = b.buffers(torch.float64, fill=float("nan"))
    if direction == "forward":
        b.eval(W, Y, "bf16x3")
    else:
        b.eval_adjoint(W, Y, "bf16x3")
    return Y


def test_batched_sweep_matches(gpu, case):
    import torch
    hms, nrhs, TW, ref = case
    b = hgemm.Batch(hms, nrhs, precisions=("fp64", "bf16x3"), adjoint=True)
    assert sorted({nr for nr, _ in b.groups}) == [8, 32, 200]
    W = _inputs(b, TW)
    for direction in ("forward", "adjoint"):
        # the reference: each group's fused level launches (the sums the
        # sweep runs; the split schedule rounds u^ to split panels between
        # its launches)
        for g in range(len(b.groups)):
            b.plan(g).set_schedule(direction, "bf16x3", "fused", 1)
        b.set_sweep(direction, 0)
        Yl = _run(b, direction, W)
        for width in (1, 5, 6, 12):
            _say(f"[batched sweep] {direction} width={width}")
            b.set_sweep(direction, width)
            assert b.get_sweep(direction) == width
            Y1 = _run(b, direction, W)
            Y2 = _run(b, direction, W)
            torch.cuda.synchronize()
            assert b.sweep_status() == 0
            for g in range(len(b.groups)):
                assert torch.isfinite(Y1[g]).all()
                assert torch.equal(Y1[g], Y2[g]), "not deterministic"
                # the split-K tile sums its partials in another order (see test_gpu_sweep.py)
                assert rel(Y1[g].cpu().numpy(), Yl[g].cpu().numpy()) <= (5e-6 if width == 12 else 1e-6)
            errs = [rel(b.instance_view(Y1, i).cpu().numpy(), ref[direction][i]) for i in range(len(hms))]
            _say(f"    worst instance vs oracle {max(errs):.2e}")
            assert max(errs) <= 1e-5
        b.set_sweep(direction, 0)


def test_batched_sweep_graph_replay(gpu, case):
    import torch
    hms, nrhs, TW, ref = case
    b = hgemm.Batch(hms, nrhs, precisions=("bf16x3",), adjoint=True)
    e = hgemm.Batch(hms, nrhs, precisions=("bf16x3",), adjoint=True)
    for x in (b, e):
        x.set_sweep("forward", 5)
        x.set_sweep("adjoint", 1)
    b.set_graph(True)
    W = _inputs(b, TW)
    for rep in range(3):
        for direction in ("forward", "adjoint"):
            _say(f"[batched sweep graph] rep={rep} {direction}")
            Y = _run(b, direction, W)
            Ye = _run(e, direction, W)
            torch.cuda.synchronize()
            assert b.sweep_status() == 0
            for g in range(len(b.groups)):
                assert torch.equal(Y[g], Ye[g])


def test_batched_sweep_with_fp32_groups(gpu, case):
    """Guarded batch: instances whose estimated BF16x3 error exceeds the
    guard run in FP32 groups beside the one sweep launch."""
    import torch
    hms = [make_problem(**kw)[1] for kw, _ in GUARDED]
    nrhs = [n for _, n in GUARDED]
    TW = [host.normal((hm.N * n,), seed=2900 + i) for i, (hm, n) in enumerate(zip(hms, nrhs))]
    ref = {"forward": [oracle.hgemm(hm, tw, n)[0] for hm, tw, n in zip(hms, TW, nrhs)]}
    b = hgemm.Batch(hms, nrhs, precisions=("fp32", "bf16x3"))
    assert len(b.fp32_instances) >= 1
    W = _inputs(b, TW)
    Yl = _run(b, "forward", W)
    b.set_sweep("forward", 1)
    Y = _run(b, "forward", W)
    torch.cuda.synchronize()
    assert b.sweep_status() == 0
    _say(f"[batched sweep + fp32 groups] routed {b.fp32_instances}")
    for g in range(len(b.groups)):
        assert rel(Y[g].cpu().numpy(), Yl[g].cpu().numpy()) <= 1e-6
    for i in range(len(hms)):
        assert rel(b.instance_view(Y, i).cpu().numpy(), ref["forward"][i]) <= 1e-5


def test_batch_autotune_picks_launch_form(gpu, case):
    import torch
    hms, nrhs, TW, ref = case
    b = hgemm.Batch(hms, nrhs, precisions=("bf16x3",), adjoint=True)
    b.autotune("bf16x3", io="f64", reps=2)
    for d in ("forward", "adjoint"):
        assert b.get_sweep(d) in (0, 1, 5, 6, 12)
    _say(f"[batch autotune] sweep forward={b.get_sweep('forward')} adjoint={b.get_sweep('adjoint')}")
    W = _inputs(b, TW)
    Y = _run(b, "forward", W)
    torch.cuda.synchronize()
    for i in range(len(hms)):
        assert rel(b.instance_view(Y, i).cpu().numpy(), ref["forward"][i]) <= 1e-5
