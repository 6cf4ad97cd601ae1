import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    from matrox_ru_amd import _native
    if not (os.path.exists(_native.HOST_LIB) and os.path.exists(_native.HGEMM_LIB)):
        from matrox_ru_amd import build
        build.build_all()
    from oracle import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


_cache = {}


def make_problem(n=2048, d=8, ncomp=6, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0, seed=1, **kw):
    """Points + compressed matrix (cached per argument set)."""
    from matrox_ru_amd import host
    key = (n, d, ncomp, leaf, rank, tol, budget, h, seed, tuple(sorted(kw.items())))
    if key not in _cache:
        X = host.gaussian_mixture(n, d, ncomp, seed=seed)
        hm = host.build_hmatrix(X, leaf_size=leaf, max_rank=rank, tol=tol, budget=budget, h=h, **kw)
        _cache[key] = (X, hm)
    return _cache[key]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
