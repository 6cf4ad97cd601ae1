// FP64 (and FP32) block products on packed A operands: 128 x 128 or
// 128 x 256 tiles, a pipeline of direct global -> LDS loads, v_mfma_f64_16x16x4_f64 (the FP32
// form, v_mfma_f32_16x16x4_f32, is described at F32PArgs).
//
// The same launches as the 64 x 64 FP64 kernel (mx_seg_gemm.h), on the
// program's 128-row tile list; only the A operands differ:
//
// * A operands (stored blocks D, B, VT, and the transposed views the
//   adjoint and the downward pass use) are packed once, when the plan is
//   built or its blocks are replaced: op(A) (M x K) is cut into 16 x 16
//   fragments, rows and K padded to 16 with zeros; fragment (rg, kg) sits at
//   (rg * ceil(K/16) + kg) * 256 doubles, and element (r, k) of a fragment at
//       ((k % 4) / 2) * 128 + (r + 16 * (k / 4)) * 2 + k % 2,
//   so that lane l = r + 16 g of an MFMA reads its operands for k = 4g..4g+3
//   as two 16-byte pieces, 1 KB apart.
// * B operands (W, wskel, uskel) stay in the reference layout (k fastest in
//   each column, leading dimension = rows of the block): the 16 x 128 (or
//   16 x 256) chunk is moved by 4-byte LDS-DMA loads (any alignment), one
//   128-byte row of 16 k per tile column, its 16-byte slots XOR-swizzled by
//   the column so that the MFMA reads hit distinct banks.  Addresses are clamped to the
//   block (k <= K-1, column <= nrhs-1); the A fragments' zero padding makes
//   the clamped k terms vanish, and clamped columns only reach outputs that
//   are not stored.
// * Outputs (Y, wskel, uskel) are written in the reference layout through a
//   per-wave LDS transpose, 64 consecutive rows per column.
//
// MFMA 16x16x4 f64 (lane l): A(row l%16, k l/16), B(k l/16, col l%16);
// result register r holds C(l/16 + 4r, l%16).  Within a chunk, MFMA step
// (h, e) of lane group g takes k = 4g + 2h + e, the same for A and B.
#pragma once
#include "mx_bf16x3.h"   // X3Block: the packing descriptor is shared

namespace mx {

constexpr int F64P_BM = 128, F64P_BN = 128, F64P_BK = 16;
constexpr int F64P_FRAG = 256;   // doubles per 16 x 16 fragment (2 KB)
constexpr int F64P_MAXSEG = 256;  // segments per task the kernel holds in LDS (longer: the 64 x 64 kernel)

__host__ __device__ inline int64_t f64p_packed_elems(int M, int K) {
    return (int64_t)((M + 15) / 16) * ((K + 15) / 16) * F64P_FRAG;
}

struct F64PArgs {
    const Task *tasks;
    const Seg *segs;      // packed segments: a_off in the packed buffer, lda = ceil(k/16), a_trans = 0
    const int2 *tiles;    // (task, 128-row tile)
    int ntiles, nrhs, n_ntiles, pad;
    const double *a[3];   // packed D, B, VT
    double *vec[4];       // W, wskel, uskel, Y (reference layout)
};

// Pack blocks[blockIdx.y] (FP64 source in its stored layout) into fragments.
__global__ void pack_f64_blocks(const double *__restrict__ src, const X3Block *__restrict__ blocks,
                                double *__restrict__ dst);

// Columns per tile: F64P_BN, or F64P_BNW for the wide tile.
constexpr int F64P_BNW = 256;

// One launch: nst = 2 (two stages; two workgroups per CU for 128 columns)
// or 3; wide: 128 x 256 tiles (8 waves of 64 x 64, one workgroup per CU).
void launch_f64p(const F64PArgs &a, int nst, int wide, hipStream_t st);

// FP32 on packed A operands (v_mfma_f32_16x16x4_f32): the same tiles, tile
// lists and packed offsets (a 16 x 16 fragment holds 256 elements in either
// precision).  An FP32 fragment is 1 KB; lane l = r + 16 g of an MFMA reads
// its k = 4g..4g+3 as one 16-byte piece at byte 16 l, so element (r, k)
// sits at (r + 16 (k / 4)) * 4 + k % 4.  B chunks (16 k of a column, 64
// bytes) are moved by 4-byte LDS-DMA loads with their 16-byte slots
// XOR-swizzled by (column / 4) % 4, so that the 16 lanes of a column group
// read distinct banks.
struct F32PArgs {
    const Task *tasks;
    const Seg *segs;      // packed segments (as F64PArgs)
    const int2 *tiles;
    int ntiles, nrhs, n_ntiles, pad;
    const float *a[3];    // packed D, B, VT
    float *vec[4];        // W, wskel, uskel, Y (reference layout)
};

__global__ void pack_f32_blocks(const float *__restrict__ src, const X3Block *__restrict__ blocks,
                                float *__restrict__ dst);

// One launch: nst = 2 or 4 stages (16 KB each; 24 KB for the wide tile).
void launch_f32p(const F32PArgs &a, int nst, int wide, hipStream_t st);

}  // namespace mx
