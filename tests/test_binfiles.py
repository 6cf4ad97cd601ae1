"""The reference's on-disk structure (.bin file set): round trip and
rejection of damaged sets (sympiler/HTree.cpp:469-542 <-> readHtree,
matroxTest-V1/HTree.h:76-260)."""
import os
import shutil

import numpy as np
import pytest

from matrox_ru_amd import host

FILES_H2 = ["levelset.bin", "idx.bin", "clevelset.bin", "wpart.bin", "cidx.bin", "lchildren.bin",
            "rchildren.bin", "nblockset.bin", "nblocks.bin", "nxval.bin", "nyval.bin", "fblockset.bin",
            "fblocks.bin", "fxval.bin", "fyval.bin", "dim.bin", "leafmap.bin", "lids.bin", "lidsoffset.bin",
            "lidslen.bin", "NN.bin", "distNN.bin"]


@pytest.mark.parametrize("budget", [0.25, 0.0], ids=["h2b", "hss"])
def test_roundtrip_reproduces_every_array(tmp_path, budget):
    X = host.gaussian_mixture(2000, 5, 4, seed=31)
    kw = dict(leaf_size=64, max_rank=40, tol=1e-6, budget=budget, h=1.1)
    a = host.build_hmatrix(X, **kw)
    host.save_bin(a, str(tmp_path))
    present = sorted(os.listdir(tmp_path))
    expected = FILES_H2 if budget > 0 else [f for f in FILES_H2 if "block" not in f]
    assert present == sorted(expected)
    # raw int32 files with the reference's element counts
    n_nodes = a.numnodes
    assert os.path.getsize(tmp_path / "lchildren.bin") == 4 * n_nodes
    assert os.path.getsize(tmp_path / "dim.bin") == 4 * a.nleaf
    assert os.path.getsize(tmp_path / "idx.bin") == 4 * (n_nodes - 1)
    assert os.path.getsize(tmp_path / "distNN.bin") == 8 * 2000 * 32
    assert np.array_equal(np.fromfile(tmp_path / "cidx.bin", dtype=np.int32), a.idx)
    b = host.load_bin(str(tmp_path), X, **kw)
    for k in ("lids", "lidsoffset", "lidslen", "Ddim", "lm", "idx", "wpart", "clevelset", "npairx", "npairy",
              "fpairx", "fpairy", "nblockSet", "nblocks", "fblockSet", "fblocks", "slen", "VTptr", "skels",
              "skelrow", "wrow"):
        assert np.array_equal(a.arrays[k], b.arrays[k]), k
    for k in ("D", "B", "VT"):
        assert np.array_equal(a.arrays[k], b.arrays[k]), k
    assert a.scalars == b.scalars


def test_damaged_sets_are_rejected(tmp_path):
    X = host.gaussian_mixture(1200, 4, 3, seed=32)
    kw = dict(leaf_size=64, max_rank=30, tol=1e-6, budget=0.3, h=1.0)
    a = host.build_hmatrix(X, **kw)
    good = tmp_path / "good"
    host.save_bin(a, str(good))

    def damaged(name, mutate):
        d = tmp_path / f"bad_{name}_{mutate.__name__}"
        shutil.copytree(good, d)
        mutate(d / name)
        return str(d)

    def truncate(p):
        data = p.read_bytes()
        p.write_bytes(data[: len(data) - 3])

    def drop(p):
        p.unlink()

    def out_of_range(p):
        v = np.fromfile(p, dtype=np.int32)
        v[len(v) // 2] = 10 ** 6
        v.tofile(p)

    for name, mut in (("lids.bin", truncate), ("wpart.bin", drop), ("nyval.bin", out_of_range),
                      ("lchildren.bin", out_of_range), ("lids.bin", out_of_range)):
        with pytest.raises(ValueError):
            host.load_bin(damaged(name, mut), X, **kw)
    # points that do not match the structure
    with pytest.raises(ValueError):
        host.load_bin(str(good), X[:1000], **kw)
