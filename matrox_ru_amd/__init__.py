"""MI355X-native evaluator for hierarchical-matrix products Y = K~ W.

A drop-in for the evaluation phase of MatRox (bangtianliu/MatRox_RU): the
compressed matrix K~ keeps the reference's HTree/CDS layout, the host builds
and compresses it in FP64 (``matrox_ru_amd.host``), and the evaluation runs in
hand-written gfx950 HIP kernels behind a C-ABI (``matrox_ru_amd.hgemm``).
"""
