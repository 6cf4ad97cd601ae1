"""Host structure builder: the arrays keep the reference's index semantics
(sympiler/HTree.cpp:299-542, ClusterTree.cpp:421-1107, hcds.h:289-485)."""
import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import host


@pytest.fixture(scope="module", params=[0.25, 0.0], ids=["h2b", "hss"])
def prob(request):
    return make_problem(n=3000, d=6, leaf=100, rank=40, tol=1e-6, budget=request.param, h=1.0)


def test_tree_shape_and_permutation(prob):
    X, hm = prob
    a = hm.arrays
    n, nl, nn = hm.N, hm.nleaf, hm.numnodes
    assert nn == 2 * nl - 1
    assert a["Ddim"].max() <= 100 and a["Ddim"].min() >= 1
    assert a["Ddim"].max() - a["Ddim"].min() <= 1     # median splits keep sizes within 1
    leaf_ids = a["lids"][:n]
    assert np.array_equal(np.sort(leaf_ids), np.arange(n))  # lids is a permutation (leaves first)
    assert a["wrow"][0] == 0 and np.all(np.diff(a["wrow"]) == a["Ddim"][:-1])
    f = nl - 1
    for v in range(nn):
        if v < f:
            assert a["lchildren"][v] == 2 * v + 1 and a["rchildren"][v] == 2 * v + 2 and a["lm"][v] == -1
        else:
            assert a["lchildren"][v] == -1 and a["lm"][v] == v - f
    # every internal node's points are its children's points
    lo, ll = a["lidsoffset"], a["lidslen"]
    pts = lambda v: set(a["lids"][lo[v]:lo[v] + ll[v]].tolist())
    for v in range(1, f):
        assert pts(v) == pts(2 * v + 1) | pts(2 * v + 2)


def test_near_far_cover_every_pair_once(prob):
    X, hm = prob
    a = hm.arrays
    nl, nn, f = hm.nleaf, hm.numnodes, hm.nleaf - 1
    desc = {}
    for v in range(nn - 1, -1, -1):
        desc[v] = {v - f} if v >= f else desc[2 * v + 1] | desc[2 * v + 2]
    cover = np.zeros((nl, nl), dtype=int)
    for x, y in zip(a["npairx"], a["npairy"]):
        cover[x, y] += 1
    for x, y in zip(a["fpairx"], a["fpairy"]):
        for lx in desc[x]:
            for ly in desc[y]:
                cover[lx, ly] += 1
    assert np.all(cover == 1)
    # near lists: self included, size (int)(nleaf*budget - 1) or 1
    cnt = np.bincount(a["npairx"], minlength=nl)
    budget = hm.config["budget"]
    assert np.all(cnt == max(1, int(nl * budget - 1)))
    assert np.all(a["npairx"][a["npairx"] == a["npairy"]] >= 0)
    for x in range(nl):
        assert x in set(a["npairy"][a["npairx"] == x].tolist())


def test_hss_far_is_sibling():
    X, hm = make_problem(n=3000, d=6, leaf=100, rank=40, tol=1e-6, budget=0.0, h=1.0)
    a = hm.arrays
    assert hm.is_hss == 1
    assert np.array_equal(a["fpairx"], np.arange(1, hm.numnodes))
    sib = np.where(a["fpairx"] % 2 == 0, a["fpairx"] - 1, a["fpairx"] + 1)
    assert np.array_equal(a["fpairy"], sib)
    assert np.array_equal(a["npairx"], np.arange(hm.nleaf)) and np.array_equal(a["npairy"], np.arange(hm.nleaf))


def test_blocking_is_consistent(prob):
    X, hm = prob
    a = hm.arrays
    nl_rows, fl_rows = hm.nl, hm.fl
    ks = []
    for i in range(nl_rows):
        for j in range(a["nblockSet"][i], a["nblockSet"][i + 1]):
            ks.extend(range(a["nblocks"][j], a["nblocks"][j + 1]))
    assert ks == list(range(len(a["npairx"])))
    ks = []
    for i in range(fl_rows):
        for j in range(a["fblockSet"][i], a["fblockSet"][i + 1]):
            ks.extend(range(a["fblocks"][j], a["fblocks"][j + 1]))
    assert ks == list(range(len(a["fpairx"])))
    if not hm.is_hss:
        # near row-block i holds destination leaves 2i, 2i+1 (NearBlock(2))
        for i in range(nl_rows):
            xs = a["npairx"][a["nblocks"][a["nblockSet"][i]]:a["nblocks"][a["nblockSet"][i + 1]]]
            assert set(xs.tolist()) <= {2 * i, 2 * i + 1}


def test_coarsening_partitions(prob):
    X, hm = prob
    a = hm.arrays
    idx, wp, cl = a["idx"], a["wpart"], a["clevelset"]
    assert sorted(idx.tolist()) == list(range(1, hm.numnodes))
    assert cl[0] == 0 and cl[-1] == len(wp) - 1 and wp[-1] == len(idx)
    pos = {v: i for i, v in enumerate(idx.tolist())}
    for p in range(len(wp) - 1):
        part = idx[wp[p]:wp[p + 1]].tolist()
        ps = set(part)
        for v in part:  # post-order inside a partition: children before parents
            for c in (a["lchildren"][v], a["rchildren"][v]):
                if c in ps:
                    assert pos[c] < pos[v]
    # groups: children are in the same or an earlier group
    grp = {}
    for g in range(len(cl) - 1):
        for p in range(cl[g], cl[g + 1]):
            for v in idx[wp[p]:wp[p + 1]]:
                grp[int(v)] = g
    for v in range(1, hm.numnodes):
        if a["lchildren"][v] != -1:
            assert grp[int(a["lchildren"][v])] <= grp[v]


def test_cds_offsets(prob):
    X, hm = prob
    a = hm.arrays
    # VT blocks: s x cols, disjoint, covering VT
    spans = []
    for v in range(1, hm.numnodes):
        s = a["slen"][v]
        cols = a["Ddim"][a["lm"][v]] if a["lchildren"][v] == -1 else \
            a["slen"][a["lchildren"][v]] + a["slen"][a["rchildren"][v]]
        assert a["proj_column"][v] == cols
        spans.append((a["VTptr"][v], a["VTptr"][v] + s * cols))
    spans.sort()
    assert spans[0][0] == 0 and spans[-1][1] == len(a["VT"])
    assert all(spans[i][1] == spans[i + 1][0] for i in range(len(spans) - 1))
    assert a["slen"][0] == 0
    # D blocks in pair order, exact kernel values
    k = 3
    x, y = a["npairx"][k], a["npairy"][k]
    f = hm.nleaf - 1
    px = a["lids"][a["lidsoffset"][f + x]:a["lidsoffset"][f + x] + a["Ddim"][x]]
    py = a["lids"][a["lidsoffset"][f + y]:a["lidsoffset"][f + y] + a["Ddim"][y]]
    h = hm.config["h"]
    Kd = np.exp(-((X[px][:, None, :] - X[py][None, :, :]) ** 2).sum(-1) / (2 * h * h))
    Dk = a["D"][a["Dptr"][k]:a["Dptr"][k] + len(px) * len(py)].reshape(len(py), len(px)).T
    assert np.abs(Dk - Kd).max() < 1e-13
    # skeletons are points of the node; projections reproduce skeleton columns
    v = f + 2
    s = a["slen"][v]
    sk = a["skels"][a["skelsoffset"][v]:a["skelsoffset"][v] + s]
    assert set(sk.tolist()) <= set(a["lids"][a["lidsoffset"][v]:a["lidsoffset"][v] + a["lidslen"][v]].tolist())
    cols = a["lids"][a["lidsoffset"][v]:a["lidsoffset"][v] + a["lidslen"][v]]
    VT = a["VT"][a["VTptr"][v]:a["VTptr"][v] + s * len(cols)].reshape(len(cols), s).T
    for q, p in enumerate(sk):
        c = int(np.nonzero(cols == p)[0][0])
        e = np.zeros(s)
        e[q] = 1
        assert np.abs(VT[:, c] - e).max() < 1e-12


def test_transform_roundtrip(prob):
    X, hm = prob
    W = host.normal((hm.N, 5), seed=21)
    TW = hm.transform(W)
    assert np.array_equal(hm.untransform(TW, 5), W)
    l = 1
    dim = hm.Ddim[l]
    blk = TW[hm.wrow[l] * 5: hm.wrow[l] * 5 + dim * 5].reshape(5, dim).T
    f = hm.nleaf - 1
    ids = hm.lids[hm.lidsoffset[f + l]: hm.lidsoffset[f + l] + dim]
    assert np.array_equal(blk, W[ids])


def test_deterministic_build():
    X = host.gaussian_mixture(1500, 5, 3, seed=99)
    a = host.build_hmatrix(X, leaf_size=64, max_rank=32, tol=1e-6, budget=0.2, h=1.0, nthreads=1)
    b = host.build_hmatrix(X, leaf_size=64, max_rank=32, tol=1e-6, budget=0.2, h=1.0, nthreads=4)
    for k in ("lids", "slen", "npairx", "npairy", "fpairx", "fpairy", "VTptr", "skels"):
        assert np.array_equal(a.arrays[k], b.arrays[k]), k
    assert np.array_equal(a.arrays["VT"], b.arrays["VT"])


def test_builder_errors():
    X = host.gaussian_mixture(100, 3, 2, seed=1)
    with pytest.raises(ValueError):
        host.build_hmatrix(X, leaf_size=200)          # one leaf only
    with pytest.raises(ValueError):
        host.build_hmatrix(X, leaf_size=16, h=0.0)    # bad bandwidth
    with pytest.raises(ValueError):
        host.build_hmatrix(X, leaf_size=16, budget=2.0)


def test_synthetic_generators_are_deterministic():
    a = host.gaussian_mixture(500, 7, 4, sig_lo=0.5, sig_hi=2.0, diag=True, seed=3)
    b = host.gaussian_mixture(500, 7, 4, sig_lo=0.5, sig_hi=2.0, diag=True, seed=3)
    assert np.array_equal(a, b)
    assert np.allclose(a.mean(0), 0, atol=1e-12) and np.allclose(a.std(0), 1, atol=1e-12)
    z = host.normal((200000,), seed=5)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


@pytest.mark.parametrize("budget", [0.0, 0.25])
def test_transposed_structure_is_the_transpose(budget):
    """The transposed CDS (pair lists swapped, D/B transposed, same VT)
    assembles to K~^T, and the oracle on it satisfies <K~x, y> = <x, K~^T y>."""
    from oracle import oracle
    X, hm = make_problem(n=900, d=4, leaf=50, rank=30, tol=1e-7, budget=budget, h=1.1)
    T = hm.transposed()
    K = oracle.assemble(hm)
    assert rel(oracle.assemble(T), K.T) <= 1e-15
    nrhs = 3
    x = host.normal((hm.N * nrhs,), seed=3)
    y = host.normal((hm.N * nrhs,), seed=4)
    Kx, _, _ = oracle.hgemm(hm, x, nrhs)
    Kty, _, _ = oracle.hgemm(T, y, nrhs)
    assert abs(float(Kx @ y) - float(x @ Kty)) <= 1e-13 * np.linalg.norm(Kx) * np.linalg.norm(y)


@pytest.mark.parametrize("budget", [0.0, 0.25])
def test_skip_blocks_keeps_the_layout(budget):
    """skip_blocks lays out D and B (Dptr, Bptr, D_len, B_len) exactly as a
    full build, without computing them; everything else is identical."""
    X = host.gaussian_mixture(1800, 5, 4, seed=11)
    kw = dict(leaf_size=60, max_rank=30, tol=1e-6, budget=budget, h=1.2)
    full = host.build_hmatrix(X, **kw)
    lay = host.build_hmatrix(X, skip_blocks=1, **kw)
    assert len(lay.D) == 0 and len(lay.B) == 0
    assert lay.scalars["D_len"] == len(full.D) and lay.scalars["B_len"] == len(full.B)
    for k, v in full.arrays.items():
        if k in ("D", "B"):
            continue
        assert np.array_equal(v, lay.arrays[k]), k
