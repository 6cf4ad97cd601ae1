"""Python binding of the HIP evaluator (include/matrox_hgemm.h).

``Plan`` uploads an ``HMatrix`` (matrox_ru_amd.host) once and evaluates
Y = K~ W on device tensors; ``hgemm``/``hgemm_hss`` call the drop-in entry
points with the reference's exact argument lists on host arrays
(symGen/HGEMM_gen.h:117-124, symGen/HGEMM_v1.h:117-124).

torch is used only for device memory and streams.  There is no fallback:
without the built HIP library every call raises.
"""
import ctypes

import numpy as np

from . import _native

FP64, FP32, BF16X3 = 0, 1, 2
# launch kinds: the FP64 / FP32 kinds cover both seg_gemm_kernel (64 x 64
# tiles on the stored blocks) and f64p_kernel / f32p_kernel (packed operands)
KIND_NAMES = {0: "gemm_f64", 1: "gemm_f32", 2: "bf16x3", 3: "convert", 4: "grad_f64", 5: "grad_f32",
              6: "grad_bf16x3"}
PLAN_ADJOINT = 1 << 8
SCHEDULES = {"fused": 0, "split": 1, "fused_balanced": 2, "split_balanced": 3, "sweep": 4, "coarse": 5,
             "coarse_split": 6}
DIRECTIONS = {"forward": 0, "adjoint": 1}
X3_VARIANTS = {1: "128x128_s2", 2: "128x128_s4", 3: "128x256_s3", 4: "128x128_reg", 5: "128x64_s2", 6: "128x32_s2",
               7: "128x128_w8s3", 8: "128x128_w8s4", 9: "256x128_w8s3", 10: "128x32_s6", 11: "128x64_s5",
               12: "128x32_k8", 13: "256x256_w8s2"}
IO_F64, IO_F32 = 0, 1
NOALLOC_SCRATCH_BYTES = 256 << 20   # MX_NOALLOC_SCRATCH_BYTES
BATCH_X3_GUARD = 8e-6               # MX_BATCH_X3_GUARD
BATCH_NO_X3_GUARD = 1 << 9          # MX_BATCH_NO_X3_GUARD
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64


class MxStructure(ctypes.Structure):
    _fields_ = [("nleaf", I32), ("numnodes", I32), ("is_hss", I32),
                ("ncount", I64), ("fcount", I64),
                ("lchildren", P), ("rchildren", P), ("lm", P), ("Ddim", P), ("slen", P),
                ("npairx", P), ("npairy", P), ("fpairx", P), ("fpairy", P),
                ("Dptr", P), ("Bptr", P), ("VTptr", P), ("wrow", P), ("skelrow", P),
                ("D", P), ("B", P), ("VT", P),
                ("D_len", I64), ("B_len", I64), ("VT_len", I64)]


class MxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mx status {code}: {msg}")
        self.code = code


_configured = False


def lib():
    global _configured
    L = _native.hgemm_lib()
    if not _configured:
        L.mx_last_error.restype = ctypes.c_char_p
        L.mx_hgemm_configure.argtypes = [I32, I32, I32, I32]
        L.mx_hgemm_configure.restype = ctypes.c_int
        L.mx_hgemm_cache.argtypes = [I32]
        L.mx_hgemm_cache.restype = ctypes.c_int
        L.HGEMM.argtypes = [P] * 12 + [I32] + [P] * 19
        L.HGEMM.restype = I32
        L.HGEMM_HSS.argtypes = [P] * 12 + [I32] + [P] * 11
        L.HGEMM_HSS.restype = I32
        L.mx_plan_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(MxStructure), ctypes.c_uint32]
        L.mx_plan_create.restype = ctypes.c_int
        L.mx_plan_create_generated.argtypes = [ctypes.POINTER(P), ctypes.POINTER(MxStructure), ctypes.c_uint32, P,
                                               I64, I32, ctypes.c_double, P, P, P, I64]
        L.mx_plan_create_generated.restype = ctypes.c_int
        L.mx_plan_destroy.argtypes = [P]
        L.mx_plan_destroy.restype = ctypes.c_int
        L.mx_eval.argtypes = [P, ctypes.c_int, P, P, I32, ctypes.c_int, P]
        L.mx_eval.restype = ctypes.c_int
        L.mx_eval_noalloc.argtypes = [P, ctypes.c_int, P, P, I32, ctypes.c_int, P]
        L.mx_eval_noalloc.restype = ctypes.c_int
        L.mx_eval_adjoint.argtypes = [P, ctypes.c_int, P, P, I32, ctypes.c_int, P]
        L.mx_eval_adjoint.restype = ctypes.c_int
        L.mx_backward.argtypes = [P, ctypes.c_int, P, P, P, I32, ctypes.c_int, P, P, P, P, P, P]
        L.mx_backward.restype = ctypes.c_int
        L.mx_plan_set_blocks.argtypes = [P, P, P, P, P]
        L.mx_plan_set_blocks.restype = ctypes.c_int
        L.mx_plan_grad_flops.argtypes = [P, I32, ctypes.POINTER(ctypes.c_double)]
        L.mx_plan_grad_flops.restype = ctypes.c_int
        L.mx_plan_set_grad_tile.argtypes = [P, I32]
        L.mx_plan_set_grad_tile.restype = ctypes.c_int
        L.mx_plan_get_grad_tile.argtypes = [P, ctypes.POINTER(I32)]
        L.mx_plan_get_grad_tile.restype = ctypes.c_int
        L.mx_plan_set_schedule.argtypes = [P, I32, ctypes.c_int, I32, I32]
        L.mx_plan_set_schedule.restype = ctypes.c_int
        L.mx_plan_get_schedule.argtypes = [P, I32, ctypes.c_int, ctypes.POINTER(I32), ctypes.POINTER(I32)]
        L.mx_plan_get_schedule.restype = ctypes.c_int
        L.mx_sweep_describe.argtypes = [P, I32] + [ctypes.POINTER(I32)] * 3 + [P] * 4
        L.mx_sweep_describe.restype = ctypes.c_int
        L.mx_lists_describe.argtypes = [P, I32, I32] + [ctypes.POINTER(I32)] * 5 + [P] * 5
        L.mx_lists_describe.restype = ctypes.c_int
        L.mx_plan_sweep_status.argtypes = [P, P, ctypes.POINTER(I32)]
        L.mx_plan_sweep_status.restype = ctypes.c_int
        L.mx_plan_autotune.argtypes = [P, ctypes.c_int, I32, ctypes.c_int, I32, P, P]
        L.mx_plan_autotune.restype = ctypes.c_int
        L.mx_plan_launch_variants.argtypes = [P, I32, P, I32, ctypes.POINTER(I32)]
        L.mx_plan_launch_variants.restype = ctypes.c_int
        L.mx_plan_set_launch_variants.argtypes = [P, I32, P, I32]
        L.mx_plan_set_launch_variants.restype = ctypes.c_int
        L.mx_plan_seg_variants.argtypes = [P, I32, ctypes.c_int, P, I32, ctypes.POINTER(I32)]
        L.mx_plan_seg_variants.restype = ctypes.c_int
        L.mx_plan_set_seg_variants.argtypes = [P, I32, ctypes.c_int, P, I32]
        L.mx_plan_set_seg_variants.restype = ctypes.c_int
        L.mx_batch_create.argtypes = [ctypes.POINTER(P), P, P, I32, ctypes.c_uint32]
        L.mx_batch_create.restype = ctypes.c_int
        L.mx_batch_destroy.argtypes = [P]
        L.mx_batch_destroy.restype = ctypes.c_int
        L.mx_batch_groups.argtypes = [P, ctypes.POINTER(I32)]
        L.mx_batch_groups.restype = ctypes.c_int
        L.mx_batch_group_info.argtypes = [P, I32, ctypes.POINTER(I32), ctypes.POINTER(I64)]
        L.mx_batch_group_info.restype = ctypes.c_int
        L.mx_batch_layout.argtypes = [P, I32, ctypes.POINTER(I32), ctypes.POINTER(I64)]
        L.mx_batch_layout.restype = ctypes.c_int
        L.mx_batch_instance_guard.argtypes = [P, I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I32)]
        L.mx_batch_instance_guard.restype = ctypes.c_int
        L.mx_batch_plan.argtypes = [P, I32, ctypes.POINTER(P)]
        L.mx_batch_plan.restype = ctypes.c_int
        L.mx_batch_eval.argtypes = [P, ctypes.c_int, P, P, ctypes.c_int, P]
        L.mx_batch_eval.restype = ctypes.c_int
        L.mx_batch_eval_adjoint.argtypes = [P, ctypes.c_int, P, P, ctypes.c_int, P]
        L.mx_batch_eval_adjoint.restype = ctypes.c_int
        L.mx_batch_backward.argtypes = [P, ctypes.c_int, P, P, P, ctypes.c_int, P, P, P, P]
        L.mx_batch_backward.restype = ctypes.c_int
        L.mx_batch_group_blocks.argtypes = [P, I32, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]
        L.mx_batch_group_blocks.restype = ctypes.c_int
        L.mx_batch_block_layout.argtypes = [P, I32, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]
        L.mx_batch_block_layout.restype = ctypes.c_int
        L.mx_batch_set_graph.argtypes = [P, I32]
        L.mx_batch_set_graph.restype = ctypes.c_int
        L.mx_batch_set_sweep.argtypes = [P, I32, I32]
        L.mx_batch_set_sweep.restype = ctypes.c_int
        L.mx_batch_get_sweep.argtypes = [P, I32, ctypes.POINTER(I32)]
        L.mx_batch_get_sweep.restype = ctypes.c_int
        L.mx_batch_sweep_status.argtypes = [P, P, ctypes.POINTER(I32)]
        L.mx_batch_sweep_status.restype = ctypes.c_int
        L.mx_batch_autotune.argtypes = [P, ctypes.c_int, ctypes.c_int, I32, P]
        L.mx_batch_autotune.restype = ctypes.c_int
        L.mx_batch_flops.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.mx_batch_flops.restype = ctypes.c_int
        L.mx_batch_set_points.argtypes = [P, I32, P, I32, ctypes.c_double]
        L.mx_batch_set_points.restype = ctypes.c_int
        L.mx_batch_dense_rows.argtypes = [P, I32, P, I32, P, P, P]
        L.mx_batch_dense_rows.restype = ctypes.c_int
        L.mx_plan_skeletons.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P)]
        L.mx_plan_skeletons.restype = ctypes.c_int
        L.mx_plan_copy_skeletons.argtypes = [P, P, P, I32, P]
        L.mx_plan_copy_skeletons.restype = ctypes.c_int
        L.mx_plan_dims.argtypes = [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]
        L.mx_plan_dims.restype = ctypes.c_int
        L.mx_plan_set_points.argtypes = [P, P, I32, ctypes.c_double]
        L.mx_plan_set_points.restype = ctypes.c_int
        L.mx_dense_rows.argtypes = [P, P, I32, P, I32, P, P]
        L.mx_dense_rows.restype = ctypes.c_int
        L.mx_plan_set_graph.argtypes = [P, I32]
        L.mx_plan_set_graph.restype = ctypes.c_int
        L.mx_plan_set_profiling.argtypes = [P, I32]
        L.mx_plan_set_profiling.restype = ctypes.c_int
        L.mx_plan_launch_times.argtypes = [P, P, P, I32, ctypes.POINTER(I32)]
        L.mx_plan_launch_times.restype = ctypes.c_int
        L.mx_plan_flops.argtypes = [P, ctypes.c_int, I32, P, ctypes.POINTER(ctypes.c_double)]
        L.mx_plan_flops.restype = ctypes.c_int
        L.mx_plan_launch_flops.argtypes = [P, ctypes.c_int, I32, P, I32, ctypes.POINTER(I32)]
        L.mx_plan_launch_flops.restype = ctypes.c_int
        _configured = True
    return L


def _check(rc):
    if rc != 0:
        raise MxError(rc, lib().mx_last_error().decode())


def _np(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def sweep_describe(hm, direction="forward"):
    """Host-only description of the dependency-driven sweep's work lists
    (mx_sweep_describe): dict of tiles (T x 12), qinfo (8 x 2), tasks
    (n x 8) and waits (W x 2), int32."""
    st, keep = structure_of(hm)
    nt, nk, nw = I32(0), I32(0), I32(0)
    sp = ctypes.byref(st)
    _check(lib().mx_sweep_describe(sp, DIRECTIONS[direction], ctypes.byref(nt), ctypes.byref(nk), ctypes.byref(nw),
                                   None, None, None, None))
    out = dict(tiles=np.zeros((nt.value, 12), np.int32), qinfo=np.zeros((8, 2), np.int32),
               tasks=np.zeros((nk.value, 8), np.int32), waits=np.zeros((max(nw.value, 1), 2), np.int32))
    _check(lib().mx_sweep_describe(sp, DIRECTIONS[direction], ctypes.byref(nt), ctypes.byref(nk), ctypes.byref(nw),
                                   out["tiles"].ctypes.data, out["qinfo"].ctypes.data, out["tasks"].ctypes.data,
                                   out["waits"].ctypes.data))
    out["waits"] = out["waits"][:nw.value]
    del keep
    return out


def lists_describe(hm, direction="forward", schedule="fused"):
    """Host-only description of the BF16x3 task lists of one direction and
    schedule (mx_lists_describe), over the stored blocks and the reference
    layouts: dict of launches (nlaunch + 1 offsets into ltasks), ltasks,
    tasks (n x 6: c_row, m, c_buf, seg_begin, seg_end, beta), segs (n x 8:
    a_off, b_row, a_buf, a_trans, lda, k, b_buf, b_ld; a_buf 3 = composite)
    and comps (n x 11: a1_off, a2_off, a1_buf, a1_trans, a1_lda, a2_buf,
    a2_trans, a2_lda, m, kmid, n), int64."""
    st, keep = structure_of(hm)
    n = [I32(0) for _ in range(5)]
    args = [ctypes.byref(st), DIRECTIONS[direction], SCHEDULES[schedule]] + [ctypes.byref(x) for x in n]
    _check(lib().mx_lists_describe(*args, None, None, None, None, None))
    nl, nlt, nt, ns, nc = (x.value for x in n)
    out = dict(launches=np.zeros(nl + 1, np.int64), ltasks=np.zeros(max(nlt, 1), np.int64),
               tasks=np.zeros((max(nt, 1), 6), np.int64), segs=np.zeros((max(ns, 1), 8), np.int64),
               comps=np.zeros((max(nc, 1), 11), np.int64))
    _check(lib().mx_lists_describe(*args, out["launches"].ctypes.data, out["ltasks"].ctypes.data,
                                   out["tasks"].ctypes.data, out["segs"].ctypes.data, out["comps"].ctypes.data))
    out["ltasks"], out["tasks"] = out["ltasks"][:nlt], out["tasks"][:nt]
    out["segs"], out["comps"] = out["segs"][:ns], out["comps"][:nc]
    del keep
    return out


def structure_of(hm):
    """mx_structure over an HMatrix; returns (struct, keepalive list)."""
    a = hm.arrays
    keep = []

    def ptr(name, dt):
        arr = _np(a[name], dt)
        keep.append(arr)
        return arr.ctypes.data

    s = MxStructure()
    s.nleaf = int(hm.scalars["nleaf"])
    s.numnodes = int(hm.scalars["numnodes"])
    s.is_hss = int(hm.scalars["is_hss"])
    s.ncount = len(a["npairx"])
    s.fcount = len(a["fpairx"])
    for f in ("lchildren", "rchildren", "lm", "Ddim", "slen", "npairx", "npairy", "fpairx", "fpairy", "VTptr",
              "wrow", "skelrow"):
        setattr(s, f, ptr(f, np.int32))
    s.Dptr = ptr("Dptr", np.uint64)
    s.Bptr = ptr("Bptr", np.uint64)
    s.VT = ptr("VT", np.float64)
    s.VT_len = len(a["VT"])
    for f in ("D", "B"):   # blocks laid out but not computed (skip_blocks) stay NULL
        n = len(a[f])
        if n == 0 and hm.scalars.get(f + "_len", -1) > 0:
            setattr(s, f, None)
            setattr(s, f + "_len", int(hm.scalars[f + "_len"]))
        else:
            setattr(s, f, ptr(f, np.float64))
            setattr(s, f + "_len", n)
    return s, keep


def _stream_handle(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class Plan:
    """Device-resident compressed matrix (mx_plan)."""

    def __init__(self, hm, precisions=("fp64",), adjoint=False, generate=None):
        """generate=(X, h): compute D and B on the device from the points X
        (n x d, original order) with bandwidth h (mx_plan_create_generated)
        instead of uploading hm's D and B, which may then be empty."""
        mask = PLAN_ADJOINT if adjoint else 0
        for p in precisions:
            mask |= 1 << PRECISIONS[p]
        s, keep = structure_of(hm)
        h = P()
        if generate is None:
            _check(lib().mx_plan_create(ctypes.byref(h), ctypes.byref(s), mask))
        else:
            X, bw = generate
            X = _np(X, np.float64)
            lids = _np(hm.arrays["lids"][:int(hm.scalars["n"])], np.int32)
            sk = _np(hm.arrays["skels"], np.int32)
            so = _np(hm.arrays["skelsoffset"], np.int64)
            _check(lib().mx_plan_create_generated(ctypes.byref(h), ctypes.byref(s), mask, X.ctypes.data, X.shape[0],
                                                  X.shape[1], float(bw), lids.ctypes.data, sk.ctypes.data,
                                                  so.ctypes.data, len(sk)))
            del X, lids, sk, so
        del keep
        self._h = h
        self.N = int(hm.scalars["n"])
        n, sk = I64(), I64()
        _check(lib().mx_plan_dims(h, ctypes.byref(n), ctypes.byref(sk)))
        self.rows = n.value
        self.skel_rows = sk.value
        self.precisions = tuple(precisions)
        self.adjoint = bool(adjoint)
        self.block_lens = (s.D_len, s.B_len, s.VT_len)
        self._wrow = np.asarray(hm.arrays["wrow"], dtype=np.int64)
        self._ddim = np.asarray(hm.arrays["Ddim"], dtype=np.int64)

    def eval(self, W, Y, nrhs, precision="fp64", stream=None):
        """Y = K~ W; W and Y are contiguous CUDA tensors of N*nrhs elements
        (leaf-blocked layout), float64 or float32 (io type follows W)."""
        import torch
        assert W.is_cuda and Y.is_cuda and W.is_contiguous() and Y.is_contiguous()
        assert W.numel() == self.rows * nrhs and Y.numel() == self.rows * nrhs
        assert W.dtype == Y.dtype and W.dtype in (torch.float64, torch.float32)
        io = IO_F64 if W.dtype == torch.float64 else IO_F32
        _check(lib().mx_eval(self._h, PRECISIONS[precision], W.data_ptr(), Y.data_ptr(), nrhs, io,
                             _stream_handle(stream)))
        return Y

    def eval_noalloc(self, W, Y, nrhs, precision="fp64", stream=None):
        """Y = K~ W through mx_eval_noalloc: the same evaluation as ``eval``
        with its work buffers in the library's static scratch array (no
        device allocation; calls overlapping on one GPU race)."""
        import torch
        assert W.is_cuda and Y.is_cuda and W.is_contiguous() and Y.is_contiguous()
        assert W.numel() == self.rows * nrhs and Y.numel() == self.rows * nrhs
        assert W.dtype == Y.dtype and W.dtype in (torch.float64, torch.float32)
        io = IO_F64 if W.dtype == torch.float64 else IO_F32
        _check(lib().mx_eval_noalloc(self._h, PRECISIONS[precision], W.data_ptr(), Y.data_ptr(), nrhs, io,
                                     _stream_handle(stream)))
        return Y

    def eval_adjoint(self, dY, dW, nrhs, precision="fp64", stream=None):
        """dW = K~^T dY (mx_eval_adjoint); same layouts and dtypes as eval."""
        import torch
        assert dY.is_cuda and dW.is_cuda and dY.is_contiguous() and dW.is_contiguous()
        assert dY.numel() == self.rows * nrhs and dW.numel() == self.rows * nrhs
        assert dY.dtype == dW.dtype and dY.dtype in (torch.float64, torch.float32)
        io = IO_F64 if dY.dtype == torch.float64 else IO_F32
        _check(lib().mx_eval_adjoint(self._h, PRECISIONS[precision], dY.data_ptr(), dW.data_ptr(), nrhs, io,
                                     _stream_handle(stream)))
        return dW

    def backward(self, W, dY, nrhs, precision="fp64", skeletons=None, blocks=(True, True, True), stream=None):
        """mx_backward: returns (dW, dD, dB, dVT); a block gradient is None
        when its flag in ``blocks`` is False.  ``skeletons`` = (wskel, uskel)
        of the forward evaluation of W (Plan.skeletons), or None for the
        plan's own."""
        import torch
        assert W.is_cuda and dY.is_cuda and W.is_contiguous() and dY.is_contiguous()
        assert W.dtype == dY.dtype and W.numel() == self.rows * nrhs and dY.numel() == self.rows * nrhs
        io = IO_F64 if W.dtype == torch.float64 else IO_F32
        dW = torch.empty_like(dY)
        outs = [torch.empty(n, dtype=W.dtype, device=W.device) if f else None
                for f, n in zip(blocks, self.block_lens)]
        ptr = [o.data_ptr() if o is not None else None for o in outs]
        ws = us = None
        if skeletons is not None:
            ws, us = (t.data_ptr() for t in skeletons)
        _check(lib().mx_backward(self._h, PRECISIONS[precision], W.data_ptr(), dY.data_ptr(), dW.data_ptr(), nrhs,
                                 io, ws, us, ptr[0], ptr[1], ptr[2], _stream_handle(stream)))
        return (dW, *outs)

    def set_blocks(self, D=None, B=None, VT=None, stream=None):
        """Replace stored blocks from FP64 CUDA tensors (mx_plan_set_blocks)."""
        import torch
        ps = []
        for t, n in zip((D, B, VT), self.block_lens):
            if t is None:
                ps.append(None)
                continue
            assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == n
            ps.append(t.data_ptr())
        _check(lib().mx_plan_set_blocks(self._h, ps[0], ps[1], ps[2], _stream_handle(stream)))

    def set_schedule(self, direction, precision, schedule, x3_variant=0):
        """Pick the coupling/downward schedule (SCHEDULES) of one direction
        ("forward"/"adjoint") for a precision, and optionally the BF16x3
        kernel variant (X3_VARIANTS; mx_plan_set_schedule).  "sweep" (BF16x3
        only) runs the whole evaluation as one dependency-driven launch; its
        x3_variant is the tile width (1, 5 or 6)."""
        _check(lib().mx_plan_set_schedule(self._h, DIRECTIONS[direction], PRECISIONS[precision],
                                          SCHEDULES[schedule], x3_variant))

    def get_schedule(self, direction, precision):
        sc, st = I32(), I32()
        _check(lib().mx_plan_get_schedule(self._h, DIRECTIONS[direction], PRECISIONS[precision], ctypes.byref(sc),
                                          ctypes.byref(st)))
        names = {v: k for k, v in SCHEDULES.items()}
        return names[sc.value], st.value

    def sweep_status(self, stream=None):
        """Error word of the last sweep (mx_plan_sweep_status): 0 clean, 1 a
        wait timed out, 2 claiming a work item timed out."""
        e = I32()
        _check(lib().mx_plan_sweep_status(self._h, _stream_handle(stream), ctypes.byref(e)))
        return e.value

    def autotune(self, precision, nrhs, io="f64", reps=3, stream=None):
        """mx_plan_autotune; returns the chosen per-direction times in ms."""
        ms = (ctypes.c_float * 2)()
        _check(lib().mx_plan_autotune(self._h, PRECISIONS[precision], nrhs, IO_F64 if io == "f64" else IO_F32, reps,
                                      ms, _stream_handle(stream)))
        return [ms[0], ms[1]] if self.adjoint else [ms[0]]

    def launch_variants(self, direction="forward"):
        """BF16x3 kernel variant of each launch (mx_plan_launch_variants)."""
        cap = 4096
        v = (I32 * cap)()
        cnt = I32()
        _check(lib().mx_plan_launch_variants(self._h, DIRECTIONS[direction], v, cap, ctypes.byref(cnt)))
        return [v[i] for i in range(min(cnt.value, cap))]

    def set_launch_variants(self, variants, direction="forward"):
        """Fix the BF16x3 kernel variant of each launch (mx_plan_set_launch_variants);
        an empty list returns to the plan-wide variant."""
        v = (I32 * max(len(variants), 1))(*variants)
        _check(lib().mx_plan_set_launch_variants(self._h, DIRECTIONS[direction], v, len(variants)))

    def seg_variants(self, precision, direction="forward"):
        """FP64/FP32 kernel variant of each launch (mx_plan_seg_variants):
        1 = 64x64 tiles, one chunk in flight; 2 = two; 3 = 128x128 (FP32);
        4 / 5 = 128x128 on packed A operands, two / three LDS stages (FP64)."""
        cap = 4096
        v = (I32 * cap)()
        cnt = I32()
        _check(lib().mx_plan_seg_variants(self._h, DIRECTIONS[direction], PRECISIONS[precision], v, cap,
                                          ctypes.byref(cnt)))
        return [v[i] for i in range(min(cnt.value, cap))]

    def set_seg_variants(self, variants, precision, direction="forward"):
        """Fix the FP64/FP32 kernel variant of each launch (mx_plan_set_seg_variants)."""
        v = (I32 * max(len(variants), 1))(*variants)
        _check(lib().mx_plan_set_seg_variants(self._h, DIRECTIONS[direction], PRECISIONS[precision], v,
                                              len(variants)))

    def grad_flops(self, nrhs):
        f = ctypes.c_double()
        _check(lib().mx_plan_grad_flops(self._h, nrhs, ctypes.byref(f)))
        return f.value

    def set_grad_tile(self, tile):
        """Tile of the split-BF16x3 block gradients: 128, 256, or 0 (chosen
        by padded work; mx_plan_set_grad_tile)."""
        _check(lib().mx_plan_set_grad_tile(self._h, int(tile)))

    def grad_tile(self):
        t = I32()
        _check(lib().mx_plan_get_grad_tile(self._h, ctypes.byref(t)))
        return t.value

    def skeletons(self, nrhs, dtype, stream=None):
        """Copies of (wskel, uskel) of the last evaluation as CUDA tensors."""
        import torch
        ws = torch.empty(self.skel_rows * nrhs, dtype=dtype, device="cuda")
        us = torch.empty_like(ws)
        _check(lib().mx_plan_copy_skeletons(self._h, ws.data_ptr(), us.data_ptr(), nrhs, _stream_handle(stream)))
        return ws, us

    def set_points(self, X_leaf, h):
        """Points in leaf order (row j = point lids[j]) and bandwidth for dense_rows."""
        X_leaf = _np(X_leaf, np.float64)
        assert X_leaf.shape[0] == self.rows
        _check(lib().mx_plan_set_points(self._h, X_leaf.ctypes.data, X_leaf.shape[1], float(h)))

    def dense_rows(self, rows, W, nrhs, stream=None):
        """Dense FP64 K(rows, :) W on the GPU; rows are leaf-order row ids.
        Returns a CUDA tensor (nrows, nrhs)."""
        import torch
        rows = _np(rows, np.int32)
        assert W.is_cuda and W.dtype == torch.float64 and W.numel() == self.rows * nrhs
        out = torch.empty(nrhs * len(rows), dtype=torch.float64, device=W.device)
        _check(lib().mx_dense_rows(self._h, rows.ctypes.data, len(rows), W.data_ptr(), nrhs, out.data_ptr(),
                                   _stream_handle(stream)))
        return out.view(nrhs, len(rows)).t()

    def rows_of(self, Y, rows, nrhs):
        """Rows (leaf-order ids) of a leaf-blocked N x nrhs device result, as (nrows, nrhs)."""
        return leaf_rows(self._wrow, self._ddim, Y, rows, nrhs)

    def set_graph(self, enable=True):
        """Replay evaluations from captured hipGraphs (mx_plan_set_graph)."""
        _check(lib().mx_plan_set_graph(self._h, int(enable)))

    def set_profiling(self, enable=True):
        _check(lib().mx_plan_set_profiling(self._h, int(enable)))

    def launch_times(self):
        """[(kind_name, ms)] of the launches of the last evaluation."""
        cap = 4096
        ms = (ctypes.c_float * cap)()
        kd = (I32 * cap)()
        cnt = I32()
        _check(lib().mx_plan_launch_times(self._h, ms, kd, cap, ctypes.byref(cnt)))
        return [(KIND_NAMES[kd[i]], float(ms[i])) for i in range(min(cnt.value, cap))]

    def flops(self, precision, nrhs):
        """(total, {kind_name: flops}) algorithmic flops of one evaluation."""
        f = (ctypes.c_double * 4)()
        tot = ctypes.c_double()
        _check(lib().mx_plan_flops(self._h, PRECISIONS[precision], nrhs, f, ctypes.byref(tot)))
        return tot.value, {KIND_NAMES[i]: f[i] for i in range(4) if f[i] > 0}

    def launch_flops(self, precision, nrhs):
        """Algorithmic flops of each kernel launch of one evaluation."""
        cap = 4096
        f = (ctypes.c_double * cap)()
        cnt = I32()
        _check(lib().mx_plan_launch_flops(self._h, PRECISIONS[precision], nrhs, f, cap, ctypes.byref(cnt)))
        return [f[i] for i in range(min(cnt.value, cap))]

    def close(self):
        if getattr(self, "_h", None):
            lib().mx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def leaf_rows(wrow, Ddim, Y, rows, nrhs):
    """Rows (leaf-order ids) of a leaf-blocked N x nrhs tensor Y (leaf l's
    block starts at element wrow[l]*nrhs, leading dimension Ddim[l])."""
    import torch
    wrow = np.asarray(wrow, dtype=np.int64)
    dim = np.asarray(Ddim, dtype=np.int64)
    rows = np.asarray(rows, dtype=np.int64)
    leaf = np.searchsorted(wrow, rows, side="right") - 1
    base = wrow[leaf] * nrhs + (rows - wrow[leaf])
    idx = base[:, None] + np.arange(nrhs)[None, :] * dim[leaf][:, None]
    return Y[torch.from_numpy(idx).to(Y.device)]


class _GroupPlan(Plan):
    """A group plan owned by a Batch (not destroyed by Python)."""

    def __init__(self, h, rows, nrhs):
        self._h = h
        self.rows = rows
        self.nrhs = nrhs
        self.adjoint = False

    def close(self):
        self._h = None


class Batch:
    """Many independent compressed matrices evaluated together (mx_batch).

    Instances with equal nrhs share a group; a group's W and Y are one flat
    CUDA tensor of rows_g * nrhs_g elements in which instance i occupies the
    element range [off_i * nrhs_g, (off_i + N_i) * nrhs_g) in its own
    leaf-blocked layout (``instance_view``)."""

    def __init__(self, hms, nrhs, precisions=("fp32",), adjoint=False, x3_guard=True):
        """x3_guard (with "bf16x3" and "fp32" in precisions): instances whose
        estimated split-BF16x3 error exceeds BATCH_X3_GUARD evaluate in FP32
        when BF16x3 is requested (``x3_err``, ``fp32_instances``)."""
        assert len(hms) == len(nrhs) and len(hms) >= 1
        mask = (PLAN_ADJOINT if adjoint else 0) | (0 if x3_guard else BATCH_NO_X3_GUARD)
        for p in precisions:
            mask |= 1 << PRECISIONS[p]
        structs, keep = [], []
        for hm in hms:
            s, k = structure_of(hm)
            structs.append(s)
            keep.append(k)
        arr = (ctypes.POINTER(MxStructure) * len(hms))(*[ctypes.pointer(s) for s in structs])
        nr = (I32 * len(hms))(*[int(n) for n in nrhs])
        h = P()
        _check(lib().mx_batch_create(ctypes.byref(h), ctypes.cast(arr, P), ctypes.cast(nr, P), len(hms), mask))
        del keep, structs
        self._h = h
        self.precisions = tuple(precisions)
        self.adjoint = bool(adjoint)
        self.N = [int(hm.scalars["n"]) for hm in hms]
        self.nrhs = [int(n) for n in nrhs]
        g = I32()
        _check(lib().mx_batch_groups(h, ctypes.byref(g)))
        self.groups = []
        for i in range(g.value):
            nr_, rows = I32(), I64()
            _check(lib().mx_batch_group_info(h, i, ctypes.byref(nr_), ctypes.byref(rows)))
            self.groups.append((nr_.value, rows.value))
        self.layout = []
        self.x3_err = []
        self.fp32_instances = []
        for i in range(len(hms)):
            gi, off = I32(), I64()
            _check(lib().mx_batch_layout(h, i, ctypes.byref(gi), ctypes.byref(off)))
            self.layout.append((gi.value, off.value))
            e, f = ctypes.c_double(), I32()
            _check(lib().mx_batch_instance_guard(h, i, ctypes.byref(e), ctypes.byref(f)))
            self.x3_err.append(e.value)
            if f.value:
                self.fp32_instances.append(i)

    def buffers(self, dtype, fill=None):
        """One CUDA tensor per group (rows_g * nrhs_g elements)."""
        import torch
        out = []
        for nr, rows in self.groups:
            t = torch.empty(rows * nr, dtype=dtype, device="cuda")
            if fill is not None:
                t.fill_(fill)
            out.append(t)
        return out

    def instance_view(self, bufs, i):
        """Instance i's leaf-blocked N_i * nrhs_i slice of the group buffers."""
        g, off = self.layout[i]
        nr = self.groups[g][0]
        return bufs[g][off * nr:(off + self.N[i]) * nr]

    def _run(self, fn, W, Y, precision, stream):
        import torch
        assert len(W) == len(self.groups) and len(Y) == len(self.groups)
        dt = W[0].dtype
        for w, y, (nr, rows) in zip(W, Y, self.groups):
            assert w.is_cuda and y.is_cuda and w.dtype == dt and y.dtype == dt
            assert w.numel() == rows * nr and y.numel() == rows * nr and w.is_contiguous() and y.is_contiguous()
        io = IO_F64 if dt == torch.float64 else IO_F32
        wp = (P * len(W))(*[w.data_ptr() for w in W])
        yp = (P * len(Y))(*[y.data_ptr() for y in Y])
        _check(fn(self._h, PRECISIONS[precision], ctypes.cast(wp, P), ctypes.cast(yp, P), io, _stream_handle(stream)))
        return Y

    def eval(self, W, Y, precision="fp32", stream=None):
        """Y[g] = K~ W[g], instance by instance (mx_batch_eval)."""
        return self._run(lib().mx_batch_eval, W, Y, precision, stream)

    def eval_adjoint(self, dY, dW, precision="fp32", stream=None):
        """dW[g] = K~^T dY[g], instance by instance (mx_batch_eval_adjoint;
        the batch must be built with adjoint=True)."""
        return self._run(lib().mx_batch_eval_adjoint, dY, dW, precision, stream)

    def group_blocks(self, g):
        """(D_len, B_len, VT_len) of group g's stored blocks."""
        d, b, v = I64(), I64(), I64()
        _check(lib().mx_batch_group_blocks(self._h, g, ctypes.byref(d), ctypes.byref(b), ctypes.byref(v)))
        return d.value, b.value, v.value

    def block_offsets(self, i):
        """First elements of instance i's D, B and VT in its group's blocks."""
        d, b, v = I64(), I64(), I64()
        _check(lib().mx_batch_block_layout(self._h, i, ctypes.byref(d), ctypes.byref(b), ctypes.byref(v)))
        return d.value, b.value, v.value

    def backward(self, W, dY, precision="fp32", blocks=True, stream=None):
        """mx_batch_backward after a forward evaluation of W: (dW, dD, dB,
        dVT) as per-group lists of tensors in W's dtype (the block lists are
        None when blocks is False); block_view slices an instance out."""
        import torch
        assert self.adjoint, "batch was created without adjoint=True"
        dt = W[0].dtype
        dW = [torch.empty_like(w) for w in W]
        io = IO_F64 if dt == torch.float64 else IO_F32
        arr = lambda ts: ctypes.cast((P * len(ts))(*[t.data_ptr() for t in ts]), P)
        grads = [None, None, None]
        if blocks:
            lens = [self.group_blocks(g) for g in range(len(self.groups))]
            grads = [[torch.empty(max(l[w], 1), dtype=dt, device="cuda") for l in lens] for w in range(3)]
        _check(lib().mx_batch_backward(self._h, PRECISIONS[precision], arr(W), arr(dY), arr(dW), io,
                                       arr(grads[0]) if blocks else None, arr(grads[1]) if blocks else None,
                                       arr(grads[2]) if blocks else None, _stream_handle(stream)))
        return dW, grads[0], grads[1], grads[2]

    def block_view(self, bufs, i, which, length):
        """Instance i's `length` elements of its D (which=0), B (1) or VT (2)
        inside the group buffers bufs."""
        g = self.layout[i][0]
        off = self.block_offsets(i)[which]
        return bufs[g][off:off + length]

    def set_graph(self, enable=True):
        _check(lib().mx_batch_set_graph(self._h, int(enable)))

    def set_sweep(self, direction, x3_variant):
        """Evaluate every BF16x3 group in one dependency-driven launch
        (mx_batch_set_sweep): x3_variant = tile width (1, 5 or 6), 0 = off."""
        _check(lib().mx_batch_set_sweep(self._h, DIRECTIONS[direction], int(x3_variant)))

    def get_sweep(self, direction):
        v = I32()
        _check(lib().mx_batch_get_sweep(self._h, DIRECTIONS[direction], ctypes.byref(v)))
        return v.value

    def sweep_status(self, stream=None):
        """Error word of the batched sweeps (mx_batch_sweep_status)."""
        e = I32()
        _check(lib().mx_batch_sweep_status(self._h, _stream_handle(stream), ctypes.byref(e)))
        return e.value

    def autotune(self, precision, io="f32", reps=3, stream=None):
        _check(lib().mx_batch_autotune(self._h, PRECISIONS[precision], IO_F64 if io == "f64" else IO_F32, reps,
                                       _stream_handle(stream)))

    def flops(self, precision):
        f = ctypes.c_double()
        _check(lib().mx_batch_flops(self._h, PRECISIONS[precision], ctypes.byref(f)))
        return f.value

    def plan(self, g):
        """The group's plan (owned by the batch): schedules, profiling."""
        h = P()
        _check(lib().mx_batch_plan(self._h, g, ctypes.byref(h)))
        nr, rows = self.groups[g]
        return _GroupPlan(h, rows, nr)

    def set_points(self, i, X_leaf, h):
        X_leaf = _np(X_leaf, np.float64)
        assert X_leaf.shape[0] == self.N[i]
        _check(lib().mx_batch_set_points(self._h, i, X_leaf.ctypes.data, X_leaf.shape[1], float(h)))

    def dense_rows(self, i, rows, W_inst):
        """Dense FP64 K_i(rows, :) W_i (W_i: the instance's FP64 slice)."""
        import torch
        rows = _np(rows, np.int32)
        nr = self.nrhs[i]
        assert W_inst.dtype == torch.float64 and W_inst.numel() == self.N[i] * nr
        out = torch.empty(nr * len(rows), dtype=torch.float64, device=W_inst.device)
        _check(lib().mx_batch_dense_rows(self._h, i, rows.ctypes.data, len(rows), W_inst.data_ptr(), out.data_ptr(),
                                         None))
        return out.view(nr, len(rows)).t()

    def close(self):
        if getattr(self, "_h", None):
            lib().mx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HGEMM_CACHE = {"off": 0, "verify": 1, "trust": 2}


def set_hgemm_cache(mode):
    """mx_hgemm_cache: whether the drop-in entry points keep their device
    plan between calls ("off", "verify" (default) or "trust")."""
    _check(lib().mx_hgemm_cache(HGEMM_CACHE[mode]))


def configure(nleaf, nl, fl, cdepth):
    """mx_hgemm_configure: the loop bounds the generated HGEMM bakes in."""
    _check(lib().mx_hgemm_configure(nleaf, nl, fl, cdepth))


def _hp(x):
    return x.ctypes.data


class Dropin:
    """One matrix's argument list for the drop-in HGEMM / HGEMM_HSS
    (symGen/HGEMM_gen.h:117-124, symGen/HGEMM_v1.h:117-124) at a given nrhs,
    built once; ``__call__`` runs the entry point on host arrays (TW, apres,
    wskel, uskel: float64, N*nrhs and skel_rows*nrhs elements), updating
    apres, wskel and uskel in place."""

    _INT = ("VTptr", "lchildren", "rchildren", "levelset", "idx", "Ddim", "lm", "slen", "nblockSet", "nblocks",
            "npairx", "npairy", "fblockSet", "fblocks", "fpairx", "fpairy", "wpart", "clevelset")

    def __init__(self, hm, nrhs, hss=None):
        a = hm.arrays
        self.hss = bool(hm.is_hss) if hss is None else bool(hss)
        self.nrhs = nrhs
        self.N = hm.N
        self.sk = int(hm.scalars["skel_rows"]) * nrhs
        self.dims = (int(hm.scalars["nleaf"]), 0 if self.hss else int(hm.scalars["nl"]),
                     0 if self.hss else int(hm.scalars["fl"]), int(hm.scalars["cdepth"]))
        self.arrs = {k: _np(a[k], np.int32) for k in self._INT}
        self.D, self.B, self.VT = (_np(a[k], np.float64) for k in ("D", "B", "VT"))
        self.Dptr, self.Bptr = _np(a["Dptr"], np.uint64), _np(a["Bptr"], np.uint64)
        self.wptr = hm.wptr(nrhs)
        self.so = hm.skeloffset(nrhs)

    def __call__(self, TW, apres, wskel, uskel):
        for x, n in ((TW, self.N * self.nrhs), (apres, self.N * self.nrhs), (wskel, self.sk), (uskel, self.sk)):
            assert x.dtype == np.float64 and x.flags.c_contiguous and x.size >= n
        configure(*self.dims)
        r = self.arrs
        head = [_hp(self.D), _hp(self.B), _hp(self.VT), _hp(self.Dptr), _hp(self.Bptr), _hp(r["VTptr"]),
                _hp(r["lchildren"]), _hp(r["rchildren"]), _hp(r["levelset"]), _hp(r["idx"]), _hp(TW), _hp(apres),
                self.nrhs, _hp(r["Ddim"]), _hp(self.wptr), _hp(self.wptr), _hp(wskel), _hp(self.so), _hp(uskel),
                _hp(self.so), _hp(r["lm"]), _hp(r["slen"])]
        if self.hss:
            rc = lib().HGEMM_HSS(*head, _hp(r["wpart"]), _hp(r["clevelset"]))
        else:
            rc = lib().HGEMM(*head, *(_hp(r[k]) for k in ("nblockSet", "nblocks", "npairx", "npairy", "fblockSet",
                                                           "fblocks", "fpairx", "fpairy", "wpart", "clevelset")))
        _check(rc)


def _dropin(hm, TW, nrhs, apres, uskel, hss):
    call = Dropin(hm, nrhs, hss)
    TW = _np(TW, np.float64)
    Y = np.zeros(hm.N * nrhs) if apres is None else _np(apres, np.float64).copy()
    ws = np.zeros(max(call.sk, 1))
    us = np.zeros(max(call.sk, 1)) if uskel is None else _np(uskel, np.float64).copy()
    call(TW, Y, ws, us)
    return Y, ws[:call.sk], us[:call.sk]


def hgemm(hm, TW, nrhs, apres=None, uskel=None):
    """Drop-in H2-b HGEMM on host arrays; returns (apres, wskel, uskel)."""
    return _dropin(hm, TW, nrhs, apres, uskel, False)


def hgemm_hss(hm, TW, nrhs, apres=None, uskel=None):
    """Drop-in HSS HGEMM (HGEMM_v1.h signature) on host arrays."""
    return _dropin(hm, TW, nrhs, apres, uskel, True)
