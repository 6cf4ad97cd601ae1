"""Near and coupling blocks generated on the device from the points
(mx_plan_create_generated) against the same matrix with host-computed
blocks, and against the oracle."""
import sys

import numpy as np
import pytest

from conftest import rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = {
    "h2b": dict(n=2500, d=7, ncomp=5, leaf=80, rank=40, tol=1e-6, budget=0.25, h=1.1),
    "hss": dict(n=2048, d=8, ncomp=6, leaf=64, rank=48, tol=1e-6, budget=0.0, h=1.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_generated_blocks_match_host_blocks(gpu, case):
    import torch
    c = CASES[case]
    print(f"[generated] case={case}", flush=True)
    X = host.gaussian_mixture(c["n"], c["d"], c["ncomp"], seed=21)
    kw = dict(leaf_size=c["leaf"], max_rank=c["rank"], tol=c["tol"], budget=c["budget"], h=c["h"])
    full = host.build_hmatrix(X, **kw)
    lay = host.build_hmatrix(X, skip_blocks=1, **kw)
    nrhs = 24
    TW = host.normal((full.N * nrhs,), seed=22)
    Yo, _, _ = oracle.hgemm(full, TW, nrhs)
    pf = hgemm.Plan(full, precisions=("fp64", "bf16x3"), adjoint=True)
    pg = hgemm.Plan(lay, precisions=("fp64", "bf16x3"), adjoint=True, generate=(X, c["h"]))
    W = torch.from_numpy(TW).cuda()
    for prec, tol in (("fp64", 1e-13), ("bf16x3", 1e-6)):
        Yf, Yg, Gf, Gg = (torch.empty_like(W) for _ in range(4))
        pf.eval(W, Yf, nrhs, prec)
        pg.eval(W, Yg, nrhs, prec)
        pf.eval_adjoint(W, Gf, nrhs, prec)
        pg.eval_adjoint(W, Gg, nrhs, prec)
        assert rel(Yg.cpu().numpy(), Yf.cpu().numpy()) <= tol, prec
        assert rel(Gg.cpu().numpy(), Gf.cpu().numpy()) <= tol, prec
        if prec == "fp64":
            assert rel(Yg.cpu().numpy(), Yo) <= 1e-12
    with pytest.raises(hgemm.MxError):
        hgemm.Plan(lay, precisions=("fp64",))   # no blocks and nothing to generate them from
