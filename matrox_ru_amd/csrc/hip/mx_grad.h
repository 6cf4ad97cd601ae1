// Gradients of the stored blocks for the backward pass of Y = K~ W.
//
// Every block gradient is a sum of NT products of two column blocks,
//     dA = sum_s  X_s * Z_s^T        (X_s: m x nrhs, Z_s: n x nrhs),
// contracted over the nrhs columns.  With dY the upstream gradient,
// w^ / u^ the forward skeleton blocks and y^ = dL/du^, g^ = dL/dw^ the
// skeleton blocks of the adjoint evaluation of dY:
//     near pair (x, y):   dD  = dY_x  W_y^T
//     far pair (x, y):    dB  = y^_x  w^_y^T
//     leaf v:             dVT = g^_v  W_v^T  + u^_v dY_v^T
//     internal v (l, r):  dVT[:, :s_l] = g^_v w^_l^T + u^_v y^_l^T,
//                         dVT[:, s_l:] = g^_v w^_r^T + u^_v y^_r^T.
// (The forward reads VT in the upward pass, w^_v = VT_v [W_v or w^_l; w^_r],
// and in the downward pass, [Y_v or u^_l; u^_r] += VT_v^T u^_v,
// symGen/HGEMM_gen.h:142-178, 196-236.)
//
// Column blocks sit in leaf-blocked (N x nrhs) or skeleton (skeleton rows x
// nrhs) buffers: element (i, k) of the block starting at row r with m rows is
// buf[r*nrhs + k*m + i].  Each output block is written once (no
// accumulation), so one launch covers every task without write conflicts.
#pragma once
#include "mx_common.h"

namespace mx {

// operand buffers of the gradient products
enum GradBuf : int32_t { GB_W = 0, GB_DY = 1, GB_WHAT = 2, GB_UHAT = 3, GB_YHAT = 4, GB_GHAT = 5, GB_COUNT = 6 };

struct alignas(16) GTask {
    int64_t c_off;      // element offset of the output block in its gradient buffer
    int32_t m, n;       // output block is m x n
    int32_t ldc;        // its leading dimension
    int32_t c_buf;      // MatBuf of the output (D, B or VT)
    int32_t seg_begin, seg_end;
};

struct alignas(16) GSeg {
    int64_t x_row;      // first row of X_s (m rows) in buffer x_buf
    int64_t z_row;      // first row of Z_s (n rows) in buffer z_buf
    int32_t x_buf, z_buf;
    int32_t x_ld, z_ld; // row counts of the two column blocks (their leading dimensions)
};

template <typename T, typename O>
struct GradArgs {
    const GTask *tasks;
    const GSeg *segs;
    const int4 *tiles;        // (task, row tile, column tile, unused), 64 x 64 tiles
    int ntiles, nrhs;
    const T *src[GB_COUNT];
    O *out[3];                // dD, dB, dVT
};

// One launch over every tile of the gradient task list (FP64 or FP32 MFMA).
template <typename T, typename O>
void launch_grad(const GradArgs<T, O> &a, hipStream_t st);

// Split-BF16x3 products (MX_BF16X3 backward).  Both operands are read as
// K-fragment blocks: an m x nrhs column block, split as x = hi + lo (bf16),
// is stored as 16-row groups by 32-column chunks, fragment (g, c) at
// ((fr + g) * nfk + c) * 512 elements of its buffer (nfk = ceil(nrhs / 32),
// fr the block's first group); lane l of a fragment holds row 16 g + l % 16,
// columns 32 c + 8 (l / 16) .. +7, zero past the block.  A fragment is then
// an MFMA A operand (16 rows x 32 deep) of X and a B operand (32 deep x 16
// columns) of Z alike, and a gradient product runs on the LDS-DMA pipeline
// of the BF16x3 evaluation kernel, 128 x 128 tiles (GRAD_X3_TILE).
constexpr int GRAD_X3_TILE = 128;
constexpr int GRAD_X3_WTILE = 256;  // the wide form: 256 x 256 tiles, 8 waves of 64 x 128
constexpr int GRAD_KMAXSEG = 8;     // segments per gradient task (2 are used)

struct alignas(16) KSeg {
    int64_t x_fr, z_fr;     // first 16-row group of X_s and of Z_s in their K-fragment buffers
    int32_t x_buf, z_buf;   // GradBuf
};

struct alignas(16) KBlock {   // a column block to pack into K-fragments
    int64_t row;    // its first row in the source (column buffer, or split panel row)
    int64_t fr;     // its first 16-row group in the K-fragment buffer
    int32_t m;      // rows
    int32_t pad;
};

template <typename O>
struct GradX3Args {
    const GTask *tasks;
    const KSeg *segs;
    const int4 *tiles;        // (task, row tile, column tile, unused), GRAD_X3_TILE tiles
    int ntiles, nrhs;
    const uint16_t *khi[GB_COUNT];   // K-fragment buffers of the operands
    const uint16_t *klo[GB_COUNT];
    O *out[3];                // dD, dB, dVT
};

template <typename O>
void launch_grad_x3(const GradX3Args<O> &a, hipStream_t st);
// the same products on GRAD_X3_WTILE tiles (a.tiles from the wide list)
template <typename O>
void launch_grad_x3w(const GradX3Args<O> &a, hipStream_t st);

// Column blocks (T = float or double, element (i, k) of a block at
// row * nrhs + k * m + i) into K-fragments.
template <typename T>
void launch_kpack_cols(const T *src, const KBlock *blocks, int nblocks, int nrhs, uint16_t *hi, uint16_t *lo,
                       hipStream_t st);

// Split panels (element (i, k) of a block at ((row + i / 32 * 32) * nrhs +
// k * 32 + i % 32, hi and lo apart) into K-fragments.
void launch_kpack_panels(const uint16_t *phi, const uint16_t *plo, const KBlock *blocks, int nblocks, int nrhs,
                         uint16_t *hi, uint16_t *lo, hipStream_t st);

}  // namespace mx
