// Split-BF16 x3 products on MFMA with FP32 accumulation, for the near-field
// (D W) and coupling (B w) segments.
//
// A value x is split as x_hi = bf16(x), x_lo = bf16(x - x_hi); a product is
// taken as a_hi b_hi + a_hi b_lo + a_lo b_hi (the a_lo b_lo term, ~2^-18
// relative, is dropped), each term on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation.  Per product the relative error is a few 2^-18.
//
// Stored blocks (A operands) are split once, when the plan is built, and
// packed in MFMA fragment order: for a block of M x K (column-major source),
// rows are padded to 16 and K to 32; fragment (mi, ki) holds the 16 x 32
// sub-block as 64 lanes x 8 consecutive k, so a wave loads an A fragment with
// one 16-byte load per lane straight from global memory (L2), no LDS.
// Column operands (W or wskel, FP32) are staged per 32-deep k chunk in LDS,
// split once per workgroup, and read as fragments with 16-byte LDS loads.
//
// Workgroup: 256 threads = 4 waves as 2 x 2, tile 128 x 128, each wave
// 64 x 64 = 4 x 4 fragments (48 MFMA per k chunk).
// MFMA 16x16x32 bf16 operand layout (lane l): A(row l%16, k 8*(l/16)+e),
// B(k 8*(l/16)+e, col l%16), e = 0..7; result register r holds
// C(4*(l/16)+r, l%16).
#pragma once
#include "mx_common.h"

namespace mx {

typedef float x3_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

struct X3Bufs {
    const uint16_t *hi[3];   // packed split A operands from D, B, VT
    const uint16_t *lo[3];
    float *vec[4];           // W, wskel, uskel, Y (FP32)
};

// An A operand to pack: op(A) is M x K; source element (i, k) is
// src[src_off + k*lda + i] (trans = 0) or src[src_off + i*lda + k] (trans = 1),
// FP64; destination offset in packed elements.
struct alignas(16) X3Block {
    int64_t src_off;
    int64_t dst_off;
    int32_t M, K;
    int32_t lda, trans;
};

constexpr int X3_BM = 128, X3_BN = 128, X3_BK = 32;
constexpr int X3_BKP = X3_BK + 8;   // padded k stride of the LDS column chunk (80 B)
constexpr int X3_CP = X3_BM + 1;    // padded column stride of the output staging
constexpr int X3_THREADS = 256;

__host__ __device__ inline int64_t x3_packed_elems(int M, int K) {
    return (int64_t)((M + 15) / 16) * ((K + 31) / 32) * 512;
}

__device__ __forceinline__ uint16_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ x3_f4 mfma_bf16(bf16x8_t a, bf16x8_t b, x3_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hi/lo split of FP64 values (elementwise; used for checks and small buffers)
__global__ void split_bf16_f64(const double *__restrict__ in, uint16_t *__restrict__ hi,
                               uint16_t *__restrict__ lo, int64_t n);

// Pack blocks[blockIdx.y] into fragment order (hi and lo).
__global__ void pack_x3_blocks(const double *__restrict__ src, const X3Block *__restrict__ blocks,
                               uint16_t *__restrict__ hi, uint16_t *__restrict__ lo);

void launch_bf16x3(int ntiles, const Task *tasks, const Seg *segs, const int2 *tiles, int nrhs, const X3Bufs &b,
                   hipStream_t st);

}  // namespace mx
