"""Python view of the host structure builder (include/matrox_host.h).

``build_hmatrix`` runs the C++ builder (tree, near/far lists, coarsening,
blocking, kNN sampling, FP64 interpolative decomposition, CDS packing) and
returns an ``HMatrix`` holding numpy copies of every array under the name the
reference's generated ``HGEMM`` gives its argument (symGen/HGEMM_gen.h:121-124),
plus the runtime-structure arrays of matroxTest-V1/HTree.h:17-74.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native

_I32 = ["VTptr", "lchildren", "rchildren", "levelset", "idx_level", "idx", "Ddim", "lm",
        "slen", "nblockSet", "nblocks", "npairx", "npairy", "fblockSet", "fblocks",
        "fpairx", "fpairy", "wpart", "clevelset", "lids", "lidsoffset", "lidslen", "leaf",
        "proj_column", "skels", "skelsoffset", "flat_nx", "flat_ny", "flat_fx", "flat_fy",
        "wrow", "urow", "skelrow", "id_rows"]
_U64 = ["Dptr", "Bptr"]
_F64 = ["D", "B", "VT", "X", "id_ratio"]
_SCALARS = ["n", "d", "nleaf", "numnodes", "depth", "cdepth", "nl", "fl", "ncount",
            "fcount", "is_hss", "N_total", "skel_rows", "D_len", "B_len"]


class MxhConfig(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_int32), ("max_rank", ctypes.c_int32),
                ("tol", ctypes.c_double), ("budget", ctypes.c_double), ("h", ctypes.c_double),
                ("kernel", ctypes.c_int32), ("knn_k", ctypes.c_int32),
                ("near_block", ctypes.c_int32), ("far_block", ctypes.c_int32),
                ("agg_levels", ctypes.c_int32), ("bottom_levels", ctypes.c_int32),
                ("nbins", ctypes.c_int32), ("nthreads", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("skip_blocks", ctypes.c_int32), ("all_rows", ctypes.c_int32)]


_configured = False


def _lib():
    global _configured
    lib = _native.host_lib()
    if not _configured:
        P = ctypes.POINTER
        lib.mxh_config_default.argtypes = [P(MxhConfig)]
        lib.mxh_build.restype = ctypes.c_void_p
        lib.mxh_build.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, P(MxhConfig)]
        lib.mxh_build_structure.restype = ctypes.c_void_p
        lib.mxh_build_structure.argtypes = lib.mxh_build.argtypes
        lib.mxh_free.argtypes = [ctypes.c_void_p]
        lib.mxh_get_scalar.restype = ctypes.c_int64
        lib.mxh_get_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        for nm, ct in (("mxh_get_i32", ctypes.c_int32), ("mxh_get_u64", ctypes.c_uint64),
                       ("mxh_get_f64", ctypes.c_double)):
            f = getattr(lib, nm)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(P(ct)), P(ctypes.c_int64)]
        lib.mxh_transform.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        lib.mxh_untransform.argtypes = lib.mxh_transform.argtypes
        lib.mxh_gaussian_mixture.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_uint64]
        lib.mxh_normal.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64]
        lib.mxh_normal_f32.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64]
        lib.mxh_last_error.restype = ctypes.c_char_p
        lib.mxh_save_bin.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.mxh_save_bin.restype = ctypes.c_int
        lib.mxh_load_bin.restype = ctypes.c_void_p
        lib.mxh_load_bin.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, P(MxhConfig)]
        _configured = True
    return lib


def default_config(**kw):
    """Reference defaults (mxh_config_default) overridden by keyword."""
    cfg = MxhConfig()
    _lib().mxh_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown builder option {k}")
        setattr(cfg, k, v)
    return cfg


@dataclass
class HMatrix:
    """A compressed kernel matrix in the reference's HTree/CDS layout."""
    scalars: dict
    arrays: dict = field(default_factory=dict)
    config: dict = field(default_factory=dict)

    def __getattr__(self, name):
        a = self.__dict__.get("arrays", {})
        if name in a:
            return a[name]
        s = self.__dict__.get("scalars", {})
        if name in s:
            return s[name]
        raise AttributeError(name)

    @property
    def N(self):
        return int(self.scalars["n"])

    def wptr(self, nrhs):
        """Element offsets of the leaf blocks of TW/Y (cachewuoffset, hcds.h:396-413)."""
        return (self.arrays["wrow"].astype(np.int64) * nrhs).astype(np.int32)

    def skeloffset(self, nrhs):
        """Element offsets of the skeleton blocks (coarcacheWUskel, hcds.h:362-393)."""
        return (self.arrays["skelrow"].astype(np.int64) * nrhs).astype(np.int32)

    def transform(self, W):
        """W (N x nrhs, column-major = numpy (nrhs, N) C-order or Fortran (N, nrhs)) -> TW."""
        W = np.asfortranarray(W, dtype=np.float64)
        nrhs = W.shape[1]
        TW = np.empty(self.N * nrhs, dtype=np.float64)
        rc = _lib().mxh_transform(self._handle_for_perm(), W.ctypes.data, nrhs, TW.ctypes.data)
        if rc != 0:
            raise RuntimeError("mxh_transform failed")
        return TW

    def untransform(self, TW, nrhs):
        W = np.empty((self.N, nrhs), dtype=np.float64, order="F")
        TW = np.ascontiguousarray(TW, dtype=np.float64)
        rc = _lib().mxh_untransform(self._handle_for_perm(), TW.ctypes.data, nrhs, W.ctypes.data)
        if rc != 0:
            raise RuntimeError("mxh_untransform failed")
        return W

    def transposed(self):
        """The structure of K~^T in the same layout: pair lists swapped, every
        near block D_k and coupling block B_k transposed in place (same
        offsets, since a transposed block has the same size), the same VT.
        Near and far blocking become one block each (nblockSet = fblockSet =
        [0, 1]) because the row-disjointness the reference's blocking
        guarantees (ClusterTree.cpp:951-1089) does not survive the swap.
        HSS: node i's coupling becomes B_sib(i)^T, the same s_i x s_sib size
        at Bptr[i-1] (HGEMM_v1.h:254-271)."""
        a = {k: np.array(v, copy=True) for k, v in self.arrays.items()}
        s = dict(self.scalars)
        Ddim, slen = a["Ddim"], a["slen"]

        def tr_blocks(src, ptr, rdim, cdim):
            out = src.copy()
            for k in range(len(ptr)):
                r, c, o = int(rdim[k]), int(cdim[k]), int(ptr[k])
                blk = src[o:o + r * c].reshape(c, r)       # column-major r x c
                out[o:o + r * c] = blk.T.reshape(-1)       # column-major c x r
            return out

        if s["is_hss"]:
            nleaf = int(s["nleaf"])
            leafs = np.arange(nleaf)
            a["D"] = tr_blocks(self.arrays["D"], a["Dptr"], Ddim[leafs], Ddim[leafs])
            numnodes = int(s["numnodes"])
            par = np.empty(numnodes, dtype=np.int64)
            for v in range(numnodes):
                if a["lchildren"][v] >= 0:
                    par[a["lchildren"][v]] = v
                    par[a["rchildren"][v]] = v
            B0, B = self.arrays["B"], a["B"]
            for i in range(1, numnodes):
                p = par[i]
                sib = a["rchildren"][p] if a["lchildren"][p] == i else a["lchildren"][p]
                si, ss = int(slen[i]), int(slen[sib])
                o_src, o_dst = int(a["Bptr"][sib - 1]), int(a["Bptr"][i - 1])
                blk = B0[o_src:o_src + ss * si].reshape(si, ss)   # B_sib, ss x si column-major
                B[o_dst:o_dst + si * ss] = blk.T.reshape(-1)
        else:
            nx, ny = a["npairx"], a["npairy"]
            fx, fy = a["fpairx"], a["fpairy"]
            a["D"] = tr_blocks(self.arrays["D"], a["Dptr"], Ddim[nx], Ddim[ny])
            a["B"] = tr_blocks(self.arrays["B"], a["Bptr"], slen[fx], slen[fy])
            a["npairx"], a["npairy"] = ny.copy(), nx.copy()
            a["fpairx"], a["fpairy"] = fy.copy(), fx.copy()
            a["nblockSet"] = np.array([0, 1], dtype=np.int32)
            a["nblocks"] = np.array([0, len(nx)], dtype=np.int32)
            a["fblockSet"] = np.array([0, 1], dtype=np.int32)
            a["fblocks"] = np.array([0, len(fx)], dtype=np.int32)
            s["nl"], s["fl"] = 1, 1
        t = HMatrix(scalars=s, arrays=a, config=dict(self.config))
        if "_owner" in self.__dict__:
            t.__dict__["_owner"] = self.__dict__["_owner"]
        return t

    def _handle_for_perm(self):
        o = self.__dict__.get("_owner")
        if o is None or not o.h:
            raise RuntimeError("HMatrix has no live builder handle")
        return o.h


class _HandleOwner:
    """Owns one mxh_handle; shared by copies of an HMatrix, freed once."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h:
                _lib().mxh_free(self.h)
                self.h = None
        except Exception:
            pass


def _collect(h, keep_handle):
    lib = _lib()
    scal = {k: int(lib.mxh_get_scalar(h, k.encode())) for k in _SCALARS}
    arrays = {}
    for names, ct, getter, dt in ((_I32, ctypes.c_int32, lib.mxh_get_i32, np.int32),
                                  (_U64, ctypes.c_uint64, lib.mxh_get_u64, np.uint64),
                                  (_F64, ctypes.c_double, lib.mxh_get_f64, np.float64)):
        for nm in names:
            p = ctypes.POINTER(ct)()
            n = ctypes.c_int64()
            if getter(h, nm.encode(), ctypes.byref(p), ctypes.byref(n)) != 0:
                continue
            if n.value == 0:
                arrays[nm] = np.zeros(0, dtype=dt)
            else:
                arrays[nm] = np.ctypeslib.as_array(p, shape=(n.value,)).astype(dt, copy=True)
    hm = HMatrix(scalars=scal, arrays=arrays)
    if keep_handle:
        hm.__dict__["_owner"] = _HandleOwner(h)
    else:
        lib.mxh_free(h)
    return hm


def build_hmatrix(X, cfg=None, structure_only=False, **kw):
    """Build a compressed kernel matrix over points X (n x d).

    ``cfg`` is an ``MxhConfig`` (or None for the reference defaults); keyword
    arguments override single fields (leaf_size, max_rank, tol, budget, h, ...).
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    if cfg is None:
        cfg = default_config(**kw)
    else:
        for k, v in kw.items():
            setattr(cfg, k, v)
    lib = _lib()
    fn = lib.mxh_build_structure if structure_only else lib.mxh_build
    h = fn(X.ctypes.data, n, d, ctypes.byref(cfg))
    if not h:
        raise ValueError(f"structure build failed: {lib.mxh_last_error().decode()}")
    hm = _collect(h, keep_handle=True)
    hm.config = {f: getattr(cfg, f) for f, _ in MxhConfig._fields_}
    return hm


def gaussian_mixture(n, d, ncomp, sig_lo=1.0, sig_hi=1.0, diag=False, mean_range=3.0,
                     standardise=True, seed=0):
    X = np.empty((n, d), dtype=np.float64)
    rc = _lib().mxh_gaussian_mixture(X.ctypes.data, n, d, ncomp, mean_range, sig_lo, sig_hi,
                                     int(diag), int(standardise), seed)
    if rc != 0:
        raise ValueError("gaussian_mixture: bad arguments")
    return X


def normal(shape, seed, dtype=np.float64, order="F"):
    """i.i.d. N(0,1) array (deterministic across platforms); Fortran order by default."""
    count = int(np.prod(shape))
    flat = np.empty(count, dtype=dtype)
    if dtype == np.float64:
        _lib().mxh_normal(flat.ctypes.data, count, seed)
    elif dtype == np.float32:
        _lib().mxh_normal_f32(flat.ctypes.data, count, seed)
    else:
        raise TypeError("normal: float64 or float32 only")
    return flat.reshape(shape, order=order)


def save_bin(hm, directory):
    """Write the structure of ``hm`` as the reference's .bin file set
    (savetree2disk / SaveSampling, sympiler/HTree.cpp:234-249, 469-542)."""
    import os
    os.makedirs(directory, exist_ok=True)
    rc = _lib().mxh_save_bin(hm._handle_for_perm(), os.fsencode(directory))
    if rc != 0:
        raise OSError(f"save_bin failed: {_lib().mxh_last_error().decode()}")


def load_bin(directory, X, cfg=None, **kw):
    """Read a .bin structure set (as readHtree, matroxTest-V1/HTree.h:76-260),
    then compress and pack it over the points X."""
    import os
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    if cfg is None:
        cfg = default_config(**kw)
    else:
        for k, v in kw.items():
            setattr(cfg, k, v)
    lib = _lib()
    h = lib.mxh_load_bin(os.fsencode(directory), X.ctypes.data, n, d, ctypes.byref(cfg))
    if not h:
        raise ValueError(f"load_bin failed: {lib.mxh_last_error().decode()}")
    hm = _collect(h, keep_handle=True)
    hm.config = {f: getattr(cfg, f) for f, _ in MxhConfig._fields_}
    return hm
