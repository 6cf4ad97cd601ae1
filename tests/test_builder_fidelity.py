"""Accuracy of the host builder's compression (the interpolative
decomposition of matroxTest-V1/compression.h:446-523 on rows sampled as in
compression.h:362-443, 597-797).

* With every point as a row (all_rows = 1, no sampling), each node's ID
  obeys the bound of column-pivoted QR: A P = Q [R11 R12; 0 R22], the ID
  error A - A(:, skels) proj is Q2 R22, and every column of R22 is no longer
  than the first left-out pivot |R_ss|, so
      ||A - A(:, skels) proj||_F <= sqrt(ncols - s) |R_ss|,
  with |R_ss| = id_ratio * |R_00| and |R_00| the largest column norm of A.
  The rank stops at the first |R_ss / R_00| < tol unless the cap stops it.
* The tolerance controls the error: on a smooth kernel (d = 3, h = 5) the
  result follows tol within a small factor, sampled or not.
* At the 65,536-point workload's parameters (d = 28, h = 1.5, leaves of 256,
  rank cap 256) the cap binds at every internal node with the first left-out
  pivot near 10% of the first: the compression error is set by the cap, not
  by the tolerance or the sampling.

Run as a script for the per-level account at a given size:
    python tests/test_builder_fidelity.py [N] [all_rows] [max_rank]
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from matrox_ru_amd import host  # noqa: E402
from oracle import oracle  # noqa: E402


def kernel(X, a, b, h):
    A, B = X[a], X[b]
    d2 = (A * A).sum(1)[:, None] + (B * B).sum(1)[None, :] - 2.0 * A @ B.T
    return np.exp(-np.maximum(d2, 0.0) / (2.0 * h * h))


def node_columns(hm, v):
    """The points whose kernel columns node v's ID compresses: a leaf's own
    points (lids order), else its children's skeletons, left first."""
    a = hm.arrays
    if a["lchildren"][v] < 0:
        o, n = a["lidsoffset"][v], a["lidslen"][v]
        return a["lids"][o:o + n]
    so, sl = a["skelsoffset"], a["slen"]
    lc, rc = a["lchildren"][v], a["rchildren"][v]
    return np.concatenate([a["skels"][so[lc]:so[lc] + sl[lc]], a["skels"][so[rc]:so[rc] + sl[rc]]])


def node_id(hm, v):
    """(columns, skeleton point ids, proj (s x ncols))."""
    a = hm.arrays
    cols = node_columns(hm, v)
    s = int(a["slen"][v])
    sk = a["skels"][a["skelsoffset"][v]:a["skelsoffset"][v] + s]
    P = a["VT"][a["VTptr"][v]:a["VTptr"][v] + s * len(cols)].reshape(len(cols), s).T
    return cols, sk, P


def node_points(hm, v):
    a = hm.arrays
    o, n = a["lidsoffset"][v], a["lidslen"][v]
    return a["lids"][o:o + n]


def id_error(hm, X, h, v, rows):
    cols, sk, P = node_id(hm, v)
    A = kernel(X, rows, cols, h)
    return np.linalg.norm(A - kernel(X, rows, sk, h) @ P) / np.linalg.norm(A)


def depths(hm):
    nn = 2 * int(hm.nleaf) - 1
    d = np.zeros(nn, dtype=int)
    for v in range(1, nn):
        d[v] = d[(v - 1) // 2] + 1
    return d


@pytest.mark.parametrize("tol,rank", [(1e-4, 10000), (1e-7, 10000), (1e-7, 12)])
def test_id_error_bound_with_all_rows(tol, rank):
    """Every node of an all-rows build: the pivoted-QR bound, the stopping
    rule, and (with a binding cap) the cap."""
    h = 1.5
    X = host.gaussian_mixture(1024, 4, 5, seed=21)
    hm = host.build_hmatrix(X, leaf_size=64, max_rank=rank, tol=tol, budget=0.1, h=h, all_rows=1)
    n = hm.N
    checked = capped = by_tol = 0
    tightness = []
    for v in range(1, 2 * int(hm.nleaf) - 1):
        cols, sk, P = node_id(hm, v)
        s, nc = len(sk), len(cols)
        assert hm.id_rows[v] == n
        A = kernel(X, np.arange(n), cols, h)
        E = A - kernel(X, np.arange(n), sk, h) @ P
        r00 = np.linalg.norm(A, axis=0).max()
        ratio = float(hm.id_ratio[v])
        if s == nc:                      # every column kept
            assert ratio == 0.0 and np.linalg.norm(E) <= 1e-12 * np.linalg.norm(A)
        elif s == rank:                  # the cap stopped it
            capped += 1
            assert ratio > 0.0
        else:                            # the tolerance stopped it
            by_tol += 1
            assert 0.0 < ratio < tol
        bound = np.sqrt(nc - s) * ratio * r00
        assert np.linalg.norm(E) <= 1.05 * bound + 1e-12 * np.linalg.norm(A), (v, np.linalg.norm(E), bound)
        if bound > 0:
            tightness.append(np.linalg.norm(E) / bound)
        checked += 1
    assert checked == 2 * int(hm.nleaf) - 2
    if rank < 64:
        assert capped > 0
    else:
        assert by_tol > 0
    # the recorded pivot is the real one: the bound is within a few times the error
    assert np.median(tightness) > 0.3


def _acc(hm, X, h, nrhs=6):
    W = host.normal((hm.N, nrhs), seed=22)
    Y = oracle.hgemm(hm, hm.transform(W), nrhs)[0]
    return oracle.acc(hm, X, h, W, Y)


def test_tolerance_controls_the_error():
    """Smooth kernel (d = 3, h = 5, leaves of 128, no binding cap): acc
    follows the tolerance, within 2 tol when every ID sees all rows and
    within 20 tol with the sampled rows (the nested levels compound the
    sampled IDs' error on rows they never saw)."""
    X = host.gaussian_mixture(4096, 3, 6, seed=1)
    accs = {}
    for tol in (1e-3, 1e-5, 1e-7):
        for all_rows in (0, 1):
            hm = host.build_hmatrix(X, leaf_size=128, max_rank=256, tol=tol, budget=0.03, h=5.0,
                                    all_rows=all_rows)
            assert int(hm.slen[1:].max()) < 256
            accs[tol, all_rows] = a = _acc(hm, X, 5.0)
            assert a <= (2 * tol if all_rows else 20 * tol), (tol, all_rows, a)
    assert accs[1e-3, 0] / accs[1e-7, 0] > 100
    assert accs[1e-3, 1] / accs[1e-7, 1] > 1000


def test_rank_cap_binds_at_the_headline_parameters():
    """d = 28, h = 1.5, leaves of 256, rank cap 256, tol 1e-5 (the 65,536-point
    workload's parameters, here at 8,192 points): a leaf keeps all its 256
    columns (exact); every internal node compresses 512 columns to the cap
    with a left-out pivot far above the tolerance, and its ID misses a large
    share of the rows outside it."""
    X = host.gaussian_mixture(8192, 28, 16, sig_lo=0.5, sig_hi=2.0, diag=True, seed=1002)
    hm = host.build_hmatrix(X, leaf_size=256, max_rank=256, tol=1e-5, budget=0.03, h=1.5)
    dep = depths(hm)
    leaf = dep == dep.max()
    internal = (dep > 0) & ~leaf
    assert (hm.slen[internal] == 256).all()
    assert np.median(hm.id_ratio[internal]) > 1e-2
    assert (hm.id_ratio[leaf] < 1e-5).all()
    rng = np.random.default_rng(5)
    v = int(np.flatnonzero(internal)[-1])
    outside = np.setdiff1d(np.arange(hm.N), node_points(hm, v))
    assert id_error(hm, X, 1.5, v, rng.choice(outside, 3000, replace=False)) > 0.05


def account(n=65536, all_rows=0, max_rank=256, nodes=6, rows=6000):
    """Per-level account of the compression at the 65,536-point workload's
    parameters: ranks, left-out pivots, ID error on rows outside the node,
    and the dense error of the evaluation."""
    X = host.gaussian_mixture(n, 28, 16, sig_lo=0.5, sig_hi=2.0, diag=True, seed=1002)
    hm = host.build_hmatrix(X, leaf_size=256, max_rank=max_rank, tol=1e-5, budget=0.03, h=1.5, all_rows=all_rows)
    nrhs = 8
    W = host.normal((n, nrhs), seed=2)
    oracle.use_mkl()
    Y = oracle.hgemm(hm, hm.transform(W), nrhs)[0]
    acc = oracle.acc(hm, X, 1.5, W, Y)
    check = np.random.default_rng(3).choice(n, 500, replace=False)
    dense = oracle.dense_rows(X, 1.5, check, W)
    derr = np.linalg.norm(hm.untransform(Y, nrhs)[check] - dense) / np.linalg.norm(dense)
    print(f"N={n} all_rows={all_rows} rank cap {max_rank}: acc {acc:.3e}, dense error {derr:.3e}")
    dep = depths(hm)
    rng = np.random.default_rng(5)
    for d in range(1, dep.max() + 1):
        vs = np.flatnonzero(dep == d)
        pick = rng.choice(vs, min(nodes, len(vs)), replace=False)
        errs = []
        for v in pick:
            out = np.setdiff1d(np.arange(n), node_points(hm, v))
            errs.append(id_error(hm, X, 1.5, int(v), rng.choice(out, min(rows, len(out)), replace=False)))
        sl, rr = hm.slen[vs], hm.id_ratio[vs]
        print(f"  level {d}: {len(vs)} nodes, rank {sl.min()}-{sl.max()} ({np.mean(sl == max_rank):.0%} at the cap), "
              f"left-out pivot median {np.median(rr):.2e}, ID error outside ({len(pick)} nodes) "
              f"mean {np.mean(errs):.2e} max {np.max(errs):.2e}")


if __name__ == "__main__":
    a = sys.argv[1:]
    account(int(a[0]) if a else 65536, int(a[1]) if len(a) > 1 else 0, int(a[2]) if len(a) > 2 else 256)
