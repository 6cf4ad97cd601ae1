"""The drop-in HGEMM / HGEMM_HSS on host arrays across repeated calls: the
kept device plan (mx_hgemm_cache) gives the oracle's results when W, apres,
uskel and nrhs change between calls, rebuilds itself when the blocks change
in place (the default mode hashes them every call), and every mode keeps
the reference's accumulation rules (HGEMM_gen.h:125-236: apres and uskel
accumulate; HGEMM_v1.h:131,269: the HSS near and far products overwrite).
"""
import sys

import numpy as np
import pytest

from conftest import rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(budget, seed):
    """A matrix of this test only (its blocks are modified in place)."""
    X = host.gaussian_mixture(3001, 5, 6, seed=seed)
    return host.build_hmatrix(X, leaf_size=100, max_rank=70, tol=1e-7, budget=budget, h=1.2)


def _call(hm, TW, nrhs, A, U):
    fn = hgemm.hgemm_hss if hm.is_hss else hgemm.hgemm
    return fn(hm, TW, nrhs, apres=A, uskel=U)


def _check(hm, TW, nrhs, A, U, tag):
    Yo, wo, uo = oracle.hgemm(hm, TW, nrhs, apres=A, uskel=U, variant="hss" if hm.is_hss else "h2")
    Yg, wg, ug = _call(hm, TW, nrhs, A, U)
    e = (rel(Yg, Yo), rel(wg, wo), rel(ug, uo))
    _say(f"  {tag}: apres {e[0]:.1e} wskel {e[1]:.1e} uskel {e[2]:.1e}")
    assert max(e) <= 1e-12, (tag, e)


@pytest.mark.parametrize("mode", ["verify", "trust", "off"])
@pytest.mark.parametrize("budget", [0.3, 0.0])
def test_repeated_calls_match_oracle(gpu, mode, budget):
    _say(f"[drop-in cache] mode={mode} budget={budget}")
    hm = _problem(budget, seed=41)
    hgemm.set_hgemm_cache(mode)
    try:
        sk = int(hm.skel_rows)
        for step, nrhs in enumerate([5, 70, 70, 5, 33]):
            TW = hm.transform(host.normal((hm.N, nrhs), seed=4100 + step))
            A = host.normal((hm.N * nrhs,), seed=4200 + step) if step % 2 else np.zeros(hm.N * nrhs)
            U = host.normal((sk * nrhs,), seed=4300 + step) if step in (2, 3) else np.zeros(sk * nrhs)
            _check(hm, TW, nrhs, A, U, f"step {step} nrhs={nrhs} apres={'rand' if step % 2 else 0} "
                                       f"uskel={'rand' if step in (2, 3) else 0}")
    finally:
        hgemm.set_hgemm_cache("verify")


@pytest.mark.parametrize("budget", [0.3, 0.0])
def test_blocks_changed_in_place(gpu, budget):
    """The default mode sees a block changed at the same address and
    rebuilds; trust mode is told by a new mx_hgemm_cache call."""
    _say(f"[drop-in cache] blocks changed in place, budget={budget}")
    hm = _problem(budget, seed=43)
    nrhs = 12
    TW = hm.transform(host.normal((hm.N, nrhs), seed=4400))
    A = np.zeros(hm.N * nrhs)
    U = np.zeros(int(hm.skel_rows) * nrhs)
    hgemm.set_hgemm_cache("verify")
    try:
        _check(hm, TW, nrhs, A, U, "before")
        for name, scale in (("D", 1.5), ("B", 0.5), ("VT", 1.25)):
            arr = hm.arrays[name]
            ptr = arr.ctypes.data
            k = len(arr) // 3
            arr[k:k + 7] *= scale            # a few entries of one block
            assert arr.ctypes.data == ptr
            _check(hm, TW, nrhs, A, U, f"verify after changing {name}")
        hgemm.set_hgemm_cache("trust")
        _check(hm, TW, nrhs, A, U, "trust")
        hm.arrays["D"][:5] *= 2.0
        hgemm.set_hgemm_cache("trust")      # drops the kept plan
        _check(hm, TW, nrhs, A, U, "trust after a new mx_hgemm_cache call")
    finally:
        hgemm.set_hgemm_cache("verify")


def test_two_matrices_alternating(gpu):
    """Calls alternating between two matrices each rebuild the plan of the
    one they pass (one plan is kept)."""
    _say("[drop-in cache] two matrices alternating")
    h1, h2 = _problem(0.3, seed=45), _problem(0.2, seed=46)
    for step in range(4):
        hm = (h1, h2)[step % 2]
        nrhs = 9
        TW = hm.transform(host.normal((hm.N, nrhs), seed=4500 + step))
        _check(hm, TW, nrhs, np.zeros(hm.N * nrhs), np.zeros(int(hm.skel_rows) * nrhs), f"step {step}")
