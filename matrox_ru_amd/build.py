"""Build the package's native libraries in-tree.

* libmatrox_host.so  : host structure builder (g++ -fopenmp)
* libmatrox_hgemm.so : gfx950 HIP kernels + C-ABI (hipcc --offload-arch=gfx950)

Run as ``python -m matrox_ru_amd.build``.  Every call recompiles.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HOST_SRC = [os.path.join(CSRC, "host", "hstruct.cpp")]
HIP_SRC = [os.path.join(CSRC, "hip", f)
           for f in ("mx_plan.hip", "mx_dropin.hip", "mx_lists.hip", "mx_generate.hip", "mx_batch.hip", "mx_bf16x3.hip", "mx_grad.hip", "mx_f64p.hip")]

# Packed FP32 instructions (v_pk_*_f32) are disabled in all device code: some
# of their forms have given wrong results on this hardware without an error.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_host():
    out = os.path.join(HERE, "libmatrox_host.so")
    _run(["g++", *HOST_FLAGS, "-o", out + ".tmp", *HOST_SRC])
    os.replace(out + ".tmp", out)
    return out


def build_hip():
    out = os.path.join(HERE, "libmatrox_hgemm.so")
    _run([hipcc(), *HIP_FLAGS, "-o", out + ".tmp", *HIP_SRC])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_host()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
