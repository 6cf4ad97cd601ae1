"""The C-ABI libraries load on a CPU host and export every function their
headers declare (no GPU calls)."""
import ctypes
import os
import re

import pytest

from matrox_ru_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HEADERS = {
    "matrox_host.h": _native.HOST_LIB,
    "matrox_hgemm.h": _native.HGEMM_LIB,
}


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
        name = m.group(1)
        start = text.rfind("\n", 0, m.start()) + 1
        line = text[start:m.start()]
        if name in ("defined", "sizeof", "if", "while", "for", "return") or line.strip().startswith("#"):
            continue
        # a declaration: return type (and optional pointer) precedes the name on the line
        if re.match(r"^\s*(const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*$", line):
            names.add(name)
    return names


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_header_symbols_exported(header):
    names = declared_functions(os.path.join(ROOT, "include", header))
    assert len(names) >= 5
    lib = ctypes.CDLL(HEADERS[header])
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_expected_entry_points_present():
    names = declared_functions(os.path.join(ROOT, "include", "matrox_hgemm.h"))
    for n in ("HGEMM", "HGEMM_HSS", "mx_hgemm_configure", "mx_plan_create", "mx_eval", "mx_plan_destroy",
              "mx_eval_adjoint", "mx_backward", "mx_plan_set_blocks"):
        assert n in names


def test_configure_rejects_bad_dims_without_gpu():
    lib = ctypes.CDLL(_native.HGEMM_LIB)
    lib.mx_last_error.restype = ctypes.c_char_p
    assert lib.mx_hgemm_configure(-1, 0, 0, 0) == -1
    assert b"bad" in lib.mx_last_error()


def test_hgemm_before_configure_fails_cleanly():
    # a fresh process state cannot be guaranteed here, so only check that a
    # null argument list is rejected before any device work
    lib = ctypes.CDLL(_native.HGEMM_LIB)
    lib.mx_hgemm_configure(4, 2, 2, 1)
    args = [None] * 12 + [4] + [None] * 19
    lib.HGEMM.restype = ctypes.c_int32
    assert lib.HGEMM(*args) == -1


def _enum_values(prefix):
    text = open(os.path.join(ROOT, "include", "matrox_hgemm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"\b(" + prefix + r"[A-Za-z0-9_]*)\s*=\s*(\d+)", text)}


def test_binding_constants_match_header():
    """The Python binding's precision, I/O, schedule, kernel-variant and
    launch-kind numbers are the header's."""
    from matrox_ru_amd import hgemm
    prec = _enum_values("MX_(?:FP64|FP32|BF16X3)")
    assert (prec["MX_FP64"], prec["MX_FP32"], prec["MX_BF16X3"]) == (hgemm.FP64, hgemm.FP32, hgemm.BF16X3)
    io = _enum_values("MX_IO_")
    assert (io["MX_IO_F64"], io["MX_IO_F32"]) == (hgemm.IO_F64, hgemm.IO_F32)
    sched = _enum_values("MX_SCHED_")
    assert sorted(sched.values()) == sorted(hgemm.SCHEDULES.values())
    assert sched["MX_SCHED_FUSED"] == hgemm.SCHEDULES["fused"]
    assert sched["MX_SCHED_SPLIT_BALANCED"] == hgemm.SCHEDULES["split_balanced"]
    variants = _enum_values("MX_X3_")
    assert sorted(variants.values()) == sorted(hgemm.X3_VARIANTS)
    kinds = _enum_values("MX_KIND_")
    assert sorted(kinds.values()) == sorted(hgemm.KIND_NAMES)
    text = open(os.path.join(ROOT, "include", "matrox_hgemm.h")).read()
    m = re.search(r"#define MX_NOALLOC_SCRATCH_BYTES \(\(size_t\)(\d+) << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == hgemm.NOALLOC_SCRATCH_BYTES
    m = re.search(r"#define MX_BATCH_X3_GUARD ([0-9.e-]+)", text)
    assert m and float(m.group(1)) == hgemm.BATCH_X3_GUARD
    m = re.search(r"#define MX_BATCH_NO_X3_GUARD \(1u << (\d+)\)", text)
    assert m and 1 << int(m.group(1)) == hgemm.BATCH_NO_X3_GUARD


def test_hgemm_cache_mode_checked_without_gpu():
    """mx_hgemm_cache rejects an unknown mode before any device work; the
    valid modes only drop what is kept (nothing, here)."""
    lib = ctypes.CDLL(_native.HGEMM_LIB)
    lib.mx_last_error.restype = ctypes.c_char_p
    assert lib.mx_hgemm_cache(7) == -1
    assert b"cache mode" in lib.mx_last_error()
    for mode in (0, 2, 1):
        assert lib.mx_hgemm_cache(mode) == 0
