#!/bin/bash
# Sanitizer runs of the host-side code over the CPU test suite
# (pytest -m "not gpu"): the host structure builder (libmatrox_host.so,
# OpenMP) and the oracle (liboracle.so, OpenMP) are rebuilt instrumented,
# loaded into the test process, and the normal builds are restored at the
# end (also after a failure).
#   1. AddressSanitizer + UndefinedBehaviorSanitizer (g++ / gcc, libgomp);
#      any UB report aborts the run.
#   2. ThreadSanitizer (clang from ROCm's LLVM, its OpenMP runtime libomp
#      and the Archer tool, which tells TSan about OpenMP's own
#      synchronisation so that only real data races are reported).
# The GPU library (libmatrox_hgemm.so) is not instrumented.
# Usage: tests/sanitize.sh [asan|tsan|all]   (default all)
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
cd "$ROOT"
HOSTSRC=matrox_ru_amd/csrc/host/hstruct.cpp
HOSTLIB=matrox_ru_amd/libmatrox_host.so
ORCSRC=oracle/hgemm_oracle.c
ORCLIB=oracle/liboracle.so
LLVM=/opt/rocm/llvm
WHAT=${1:-all}

restore() {
    python -c 'from matrox_ru_amd import build; build.build_host()' > /dev/null
    make -s -B -C oracle > /dev/null
}
trap restore EXIT

run_tests() {
    # the tests' own child processes inherit the preload
    python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider
}

if [ "$WHAT" = asan ] || [ "$WHAT" = all ]; then
    echo "== AddressSanitizer + UndefinedBehaviorSanitizer"
    FL="-O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=all -fPIC -shared -fopenmp"
    g++ $FL -std=c++17 -o $HOSTLIB $HOSTSRC
    gcc $FL -std=c11 -o $ORCLIB $ORCSRC -lm -ldl
    LD_PRELOAD="$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so)" \
        ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
        run_tests
fi

if [ "$WHAT" = tsan ] || [ "$WHAT" = all ]; then
    echo "== ThreadSanitizer (LLVM OpenMP + Archer)"
    FL="-O1 -g -fno-omit-frame-pointer -fsanitize=thread -fopenmp -fPIC -shared"
    $LLVM/bin/clang++ $FL -std=c++17 -o $HOSTLIB $HOSTSRC -Wl,-rpath,$LLVM/lib -L$LLVM/lib -lomp
    $LLVM/bin/clang $FL -std=c11 -o $ORCLIB $ORCSRC -lm -ldl -Wl,-rpath,$LLVM/lib -L$LLVM/lib -lomp
    TSAN_RT=$(ls $LLVM/lib/clang/*/lib/linux/libclang_rt.tsan-x86_64.so | head -1)
    LD_PRELOAD="$TSAN_RT" OMP_TOOL_LIBRARIES=$LLVM/lib/libarcher.so \
        TSAN_OPTIONS="halt_on_error=1:report_signal_unsafe=0:suppressions=$ROOT/tests/tsan.supp" \
        LD_LIBRARY_PATH=$LLVM/lib:${LD_LIBRARY_PATH:-} \
        run_tests
fi
echo "sanitizer runs clean"
