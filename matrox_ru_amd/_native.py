"""Loading of the package's native libraries.

Two shared libraries are built in-tree next to this file:

* ``libmatrox_host.so`` - the host structure builder (C++/OpenMP),
* ``libmatrox_hgemm.so`` - the gfx950 HIP evaluator and its C-ABI.

Both are required: there is no fallback path.  A missing library raises
``NativeLibraryMissing`` naming the build step.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

HOST_LIB = os.path.join(_HERE, "libmatrox_host.so")
HGEMM_LIB = os.path.join(_HERE, "libmatrox_hgemm.so")


class NativeLibraryMissing(RuntimeError):
    pass


_cache = {}


def load(path):
    """dlopen one of the package libraries (cached)."""
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{os.path.basename(path)} is not built; run "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`python -m matrox_ru_amd.build` from the repository root")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[path] = lib
    return lib


def host_lib():
    return load(HOST_LIB)


def hgemm_lib():
    return load(HGEMM_LIB)
