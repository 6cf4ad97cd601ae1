"""Pins the CPU oracle (oracle/hgemm_oracle.c) before it is trusted.

The reference ships no golden vectors, known-answer tests or fixtures for the
evaluation path, and its evaluator cannot be built here (it needs Intel MKL's
mkl.h).  The oracle is therefore pinned by what the mathematics fixes:

* closed form: with ID tolerance 0 and no rank cap every projection is a
  permuted identity, so K~ = K exactly and HGEMM must reproduce dense K W;
  this fixes the point permutation, transform, and every index/offset;
* a second route: K~ assembled densely through explicit nested bases
  (orc_assemble) must equal the sweep (orc_hgemm_h2 / orc_hgemm_hss) for a
  genuinely low-rank structure;
* invariants: linearity/column independence, HSS symmetry K~ = K~^T,
  sweep-stage checks against numpy GEMMs;
* planted mistakes (corrupted structure arrays) must break these checks.
"""
import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import host
from oracle import oracle


def _perm(hm):
    return hm.arrays["lids"][: hm.N]


@pytest.mark.parametrize("budget", [0.25, 0.0])
def test_closed_form_full_rank_equals_dense(budget):
    X, hm = make_problem(n=1024, d=5, leaf=64, rank=100000, tol=0.0, budget=budget, h=1.3)
    W = host.normal((1024, 7), seed=3)
    Y, _, _ = oracle.hgemm(hm, hm.transform(W), 7)
    K = oracle.dense_rows(X, 1.3, np.arange(1024), W)
    assert rel(hm.untransform(Y, 7), K) < 1e-12


@pytest.mark.parametrize("budget", [0.25, 0.0])
def test_sweep_equals_nested_basis_assembly(budget):
    X, hm = make_problem(n=1536, d=6, leaf=48, rank=24, tol=1e-7, budget=budget, h=1.0)
    assert hm.slen[1:].max() <= 24
    W = host.normal((1536, 5), seed=4)
    Y, _, _ = oracle.hgemm(hm, hm.transform(W), 5)
    Kt = oracle.assemble(hm)
    p = _perm(hm)
    assert rel(hm.untransform(Y, 5)[p], Kt @ W[p]) < 1e-12


def test_hss_variant_matches_h2_form_and_is_symmetric():
    X, hm = make_problem(n=1024, d=4, leaf=64, rank=20, tol=1e-8, budget=0.0, h=1.0)
    assert hm.is_hss == 1
    W = host.normal((1024, 3), seed=5)
    TW = hm.transform(W)
    Yh, wh, uh = oracle.hgemm(hm, TW, 3, variant="hss")
    Y2, w2, u2 = oracle.hgemm(hm, TW, 3, variant="h2")
    assert rel(Yh, Y2) < 1e-14 and rel(wh, w2) < 1e-14 and rel(uh, u2) < 1e-14
    Kt = oracle.assemble(hm)
    assert rel(Kt, Kt.T) < 1e-14


def test_linearity_and_column_independence():
    X, hm = make_problem(n=1024, d=6, leaf=64, rank=32, tol=1e-6, budget=0.25, h=1.0)
    W = host.normal((1024, 6), seed=6)
    Y, _, _ = oracle.hgemm(hm, hm.transform(W), 6)
    Yw = hm.untransform(Y, 6)
    for j in (0, 3, 5):
        y1, _, _ = oracle.hgemm(hm, hm.transform(W[:, j:j + 1]), 1)
        assert rel(hm.untransform(y1, 1)[:, 0], Yw[:, j]) < 1e-14
    Y2, _, _ = oracle.hgemm(hm, hm.transform(2.5 * W[:, :2] - W[:, 2:4]), 2)
    assert rel(hm.untransform(Y2, 2), 2.5 * Yw[:, :2] - Yw[:, 2:4]) < 1e-13


def test_accumulation_rules():
    """H2-b accumulates into apres and uskel (beta = 1); HSS overwrites."""
    X, hm = make_problem(n=1024, d=6, leaf=64, rank=32, tol=1e-6, budget=0.25, h=1.0)
    W = host.normal((1024, 4), seed=8)
    TW = hm.transform(W)
    Y0, w0, u0 = oracle.hgemm(hm, TW, 4)
    A = host.normal((Y0.size,), seed=9)
    U = host.normal((u0.size,), seed=10)
    Y1, w1, u1 = oracle.hgemm(hm, TW, 4, apres=A)
    assert rel(w1, w0) < 1e-15
    assert rel(Y1 - A, Y0) < 1e-12 and rel(u1, u0) < 1e-15
    # a non-zero initial uskel is propagated linearly by the down pass
    Y2, _, u2 = oracle.hgemm(hm, TW, 4, uskel=U)
    Y3, _, u3 = oracle.hgemm(hm, TW, 4, uskel=2 * U)
    assert rel(Y3 - Y0, 2 * (Y2 - Y0)) < 1e-12 and rel(u3 - u0, 2 * (u2 - u0)) < 1e-12
    assert rel(Y2, Y0) > 1e-3
    X2, hs = make_problem(n=1024, d=4, leaf=64, rank=20, tol=1e-8, budget=0.0, h=1.0)
    TW2 = hs.transform(W)
    Ya, _, ua = oracle.hgemm(hs, TW2, 4)
    A2 = host.normal((Ya.size,), seed=11)
    Yb, _, ub = oracle.hgemm(hs, TW2, 4, apres=A2, uskel=host.normal((ua.size,), seed=12))
    assert rel(Yb, Ya) < 1e-14 and rel(ub, ua) < 1e-14


def test_upward_stage_against_numpy():
    X, hm = make_problem(n=1024, d=6, leaf=64, rank=32, tol=1e-6, budget=0.25, h=1.0)
    nrhs = 3
    W = host.normal((1024, nrhs), seed=13)
    TW = hm.transform(W)
    _, ws, _ = oracle.hgemm(hm, TW, nrhs)
    a = hm.arrays
    f = hm.nleaf - 1
    for v in (f, f + 7, 3, 1):
        s = a["slen"][v]
        blk = ws[a["skelrow"][v] * nrhs: a["skelrow"][v] * nrhs + s * nrhs].reshape(nrhs, s).T
        if v >= f:
            l = a["lm"][v]
            dim = a["Ddim"][l]
            VT = a["VT"][a["VTptr"][v]: a["VTptr"][v] + s * dim].reshape(dim, s).T
            Wl = TW[a["wrow"][l] * nrhs: a["wrow"][l] * nrhs + dim * nrhs].reshape(nrhs, dim).T
            ref = VT @ Wl
        else:
            lc, rc = 2 * v + 1, 2 * v + 2
            sl, sr = a["slen"][lc], a["slen"][rc]
            VT = a["VT"][a["VTptr"][v]: a["VTptr"][v] + s * (sl + sr)].reshape(sl + sr, s).T
            wl = ws[a["skelrow"][lc] * nrhs: a["skelrow"][lc] * nrhs + sl * nrhs].reshape(nrhs, sl).T
            wr = ws[a["skelrow"][rc] * nrhs: a["skelrow"][rc] * nrhs + sr * nrhs].reshape(nrhs, sr).T
            ref = VT[:, :sl] @ wl + VT[:, sl:] @ wr
        assert rel(blk, ref) < 1e-14


def test_gemm_restatement_matches_numpy():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((37, 21))
    Bm = rng.standard_normal((21, 9))
    assert rel(oracle.gemm(A, Bm), A @ Bm) < 1e-15
    At = rng.standard_normal((21, 37))
    C = rng.standard_normal((37, 9))
    assert rel(oracle.gemm(At, Bm, transA=True, C=C, beta=1.0), C + At.T @ Bm) < 1e-15


def test_mkl_route_of_the_cpu_baseline_matches():
    """The CPU baseline's MKL route (cblas_dgemm inside the same sweeps) gives
    the built-in GEMM's results up to rounding, and switching back restores
    the built-in GEMM bit for bit."""
    X, hm = make_problem(n=2048, d=6, leaf=64, rank=40, tol=1e-6, budget=0.2, h=1.0)
    TW = host.normal((hm.N * 9,), seed=77)
    Y0, w0, u0 = oracle.hgemm(hm, TW, 9)
    path = oracle.use_mkl()
    if path is None:
        pytest.skip("no MKL runtime in this image")
    try:
        Y1, w1, u1 = oracle.hgemm(hm, TW, 9)
        G1 = oracle.gemm(np.eye(5)[:, :3], np.arange(6.0).reshape(3, 2))
    finally:
        oracle.use_blas(None)
    assert rel(Y1, Y0) < 1e-13 and rel(w1, w0) < 1e-13 and rel(u1, u0) < 1e-13
    assert np.array_equal(G1, np.eye(5)[:, :3] @ np.arange(6.0).reshape(3, 2))
    Y2, _, _ = oracle.hgemm(hm, TW, 9)
    assert np.array_equal(Y2, Y0)


def test_acc_and_flops_definitions():
    X, hm = make_problem(n=1024, d=5, leaf=64, rank=100000, tol=0.0, budget=0.25, h=1.3)
    W = host.normal((1024, 4), seed=14)
    Y, _, _ = oracle.hgemm(hm, hm.transform(W), 4)
    # exact K~ = K: acc (computeError) must be at rounding level
    assert oracle.acc(hm, X, 1.3, W, Y) < 1e-13
    # perturb the output rows of leaf 0 by a known relative amount
    Yp = Y.copy()
    dim0 = hm.Ddim[0]
    for j in range(4):
        Yp[j * dim0: j * dim0 + dim0] *= 1.01
    assert abs(oracle.acc(hm, X, 1.3, W, Yp) - 0.01) < 1e-9
    # flops: direct count from the pair lists and ranks
    a = hm.arrays
    nn = hm.numnodes
    tree = 0
    for v in range(1, nn):
        if a["lchildren"][v] == -1:
            tree += 2 * a["slen"][v] * 4 * a["Ddim"][a["lm"][v]]
        else:
            tree += 2 * a["slen"][v] * (a["slen"][a["lchildren"][v]] + a["slen"][a["rchildren"][v]]) * 4
    near = sum(2 * a["Ddim"][x] * a["Ddim"][y] * 4 for x, y in zip(a["npairx"], a["npairy"]))
    far = sum(2 * a["slen"][x] * a["slen"][y] * 4 for x, y in zip(a["fpairx"], a["fpairy"]))
    assert oracle.flops(hm, 4) == pytest.approx(2 * tree + near + far, rel=1e-15)


def test_planted_mistakes_are_caught():
    """Corrupting the structure the oracle reads must break the closed form."""
    X, hm = make_problem(n=1024, d=5, leaf=64, rank=100000, tol=0.0, budget=0.25, h=1.3)
    W = host.normal((1024, 3), seed=15)
    K = oracle.dense_rows(X, 1.3, np.arange(1024), W)
    TW = hm.transform(W)
    import copy
    # (1) a near pair pointing at the wrong column leaf
    bad = copy.copy(hm)
    bad.arrays = dict(hm.arrays)
    py = hm.arrays["npairy"].copy()
    i = int(np.nonzero(hm.arrays["npairx"] != hm.arrays["npairy"])[0][0])
    py[i] = hm.arrays["npairx"][i]
    bad.arrays["npairy"] = py
    Y, _, _ = oracle.hgemm(bad, TW, 3)
    assert rel(hm.untransform(Y, 3), K) > 1e-3
    # (2) VT blocks read transposed (wrong leading dimension in the sweep)
    bad2 = copy.copy(hm)
    bad2.arrays = dict(hm.arrays)
    VT = hm.arrays["VT"].copy()
    v = hm.nleaf - 1
    s = hm.arrays["slen"][v]
    c = hm.arrays["Ddim"][hm.arrays["lm"][v]]
    o = hm.arrays["VTptr"][v]
    VT[o:o + s * c] = VT[o:o + s * c].reshape(c, s).T.ravel() if s == c else VT[o:o + s * c][::-1]
    bad2.arrays["VT"] = VT
    Y2, _, _ = oracle.hgemm(bad2, TW, 3)
    assert rel(hm.untransform(Y2, 3), K) > 1e-6
