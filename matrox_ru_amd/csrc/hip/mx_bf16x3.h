// Split-BF16 x3 products on MFMA with FP32 accumulation: the mixed-precision
// form of every block product of an evaluation (near D W, coupling B w,
// upward VT w, downward VT^T u).
//
// A value x is split as x_hi = bf16(x), x_lo = bf16(x - x_hi); a product is
// taken as a_lo b_hi + a_hi b_lo + a_hi b_hi (the a_lo b_lo term, ~2^-18
// relative, is dropped), each term on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation.
//
// Both operands live in HBM already split and in MFMA fragment order, so the
// main loop only moves 1 KB fragments and issues MFMAs:
//
// * A operands (stored blocks, and transposed views of VT) are packed when
//   the plan is built: op(A) (M x K) is cut into 16 x 32 fragments; fragment
//   (mi, ki) at ((mi * ceil(K/32)) + ki) * 512 holds 64 lanes x 8 consecutive
//   k (rows padded to 16 and K to 32 with zeros).
// * B operands (W, wskel, uskel) are "split panels": a block of R rows and
//   nrhs columns occupies ceil(R/32)*32 padded rows; element (k, j) sits at
//   prow*nrhs + ((k/32)*nrhs + j)*32 + k%32, so the 16 x 32 fragment of
//   columns j..j+15 is 1 KB contiguous.  W is split into panels once per
//   evaluation; wskel/uskel panels are written, split, by the epilogue of the
//   launch that produces them.
//
// Workgroup: 2 x 2 or 2 x 4 waves, tiles of 128 rows by 32 to 256 columns,
// each wave 64 x 64 (4 x 4 fragments, 48 MFMA per 32-deep chunk) or, for
// small nrhs, 64 x 32 / 64 x 16; fragments move global -> LDS by direct
// LDS-DMA loads ahead of the MFMAs (see the kernel).
// MFMA 16x16x32 bf16 operand layout (lane l): A(row l%16, k 8*(l/16)+e),
// B(k 8*(l/16)+e, col l%16), e = 0..7; result register r holds
// C(4*(l/16)+r, l%16).
#pragma once
#include "mx_common.h"
#include "mx_dma.h"

namespace mx {

typedef float x3_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

// An A operand to pack: op(A) is M x K; source element (i, k) is
// src[src_off + k*lda + i] (trans = 0) or src[src_off + i*lda + k] (trans = 1),
// FP64; destination offset in packed elements.
struct alignas(16) X3Block {
    int64_t src_off;
    int64_t dst_off;
    int32_t M, K;
    int32_t lda, trans;
};

// A leaf block of W to split into its panel.
struct alignas(16) X3Rows {
    int64_t src_row;    // first row in the leaf-blocked input
    int64_t prow;       // first padded row of its panel
    int32_t R;          // rows
    int32_t pad;
};

struct X3Args {
    const Task *tasks;
    const Seg *segs;
    const int2 *tiles;
    int ntiles, nrhs, n_ntiles, y_f64;
    const uint16_t *ahi[3];   // packed A operands from D, B, VT
    const uint16_t *alo[3];
    uint16_t *phi[3];         // split panels of W, wskel, uskel (VecBuf 0..2)
    uint16_t *plo[3];
    void *Y;                  // output in the reference layout, FP32 or FP64
};

constexpr int X3_BM = 128, X3_BK = 32;   // tile rows of every variant, chunk depth
constexpr int X3_MAXSEG = 256;      // segments per task the kernel holds in LDS
constexpr int X3_FRAG = 512;        // bf16 per 16 x 32 fragment (1 KB)

__host__ __device__ inline int64_t x3_packed_elems(int M, int K) {
    return (int64_t)((M + 15) / 16) * ((K + 31) / 32) * 512;
}
__host__ __device__ inline int64_t x3_pad_rows(int R) { return ((int64_t)R + 31) / 32 * 32; }

__device__ __forceinline__ uint16_t bf16_rne(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Pack blocks[blockIdx.y] into fragment order (hi and lo).
__global__ void pack_x3_blocks(const double *__restrict__ src, const X3Block *__restrict__ blocks,
                               uint16_t *__restrict__ hi, uint16_t *__restrict__ lo);

// Split the leaf blocks of W (FP64 or FP32, leaf-blocked reference layout)
// into panels.
void launch_x3_split_w(const void *W, int w_f64, const X3Rows *rows, int nrows_blocks, int nrhs, uint16_t *hi,
                       uint16_t *lo, hipStream_t st);

// Skeleton panels back to FP32 in the reference layout (block b: R rows at
// panel row prow, written at src_row*nrhs, leading dimension R).
void launch_x3_panels_to_f32(const uint16_t *hi, const uint16_t *lo, const X3Rows *blocks, int nblocks, int nrhs,
                             float *out, hipStream_t st);

// Kernel variants (tile rows x columns, pipeline stages):
//   128 x 128, 2 stages: 4 waves, 64 KB of stages, two workgroups per CU;
//   128 x 128, 4 stages: 4 waves, 128 KB, one workgroup per CU;
//   128 x 256, 3 stages: 8 waves, 144 KB, one workgroup per CU;
//   128 x 128, fragments staged through VGPRs instead of LDS-DMA;
//   128 x 64 and 128 x 32, 2 stages (narrow waves, for small nrhs);
//   128 x 128 with 8 waves of 64 x 32 and 3 or 4 stages (one workgroup per
//   CU, two waves per SIMD, two or three chunks of prefetch);
//   256 x 128, 3 stages: 8 waves (4 x 2) of 64 x 64, 144 KB, one workgroup
//   per CU; both 128-row halves of a 256-row block in one workgroup, so the
//   B operand (W, skeleton panels) is read once per 256 rows (its launches
//   use the 256-row tile lists of the plan);
//   128 x 32 with 6 stages and 128 x 64 with 5 (120 KB, one workgroup per
//   CU, five or four chunks of loads in flight): for the latency-bound
//   launches of small nrhs and the upper tree levels;
//   128 x 32 with the K chunks split over 8 waves (register operands, the
//   partial tiles summed through LDS in wave order): a task's chunks load
//   in parallel instead of in a pipeline, for the chains of short tasks;
//   256 x 256, 2 stages: 8 waves (4 x 2) of 64 x 128, 128 KB of stages, one
//   workgroup per CU (two waves per SIMD): half the fragment bytes per flop
//   of the 128 x 128 tile, and a quarter fewer LDS reads per MFMA; three
//   quarters of each chunk's MFMAs are issued behind the LDS reads of the
//   next quarter of B fragments.
enum X3Variant {
    X3_V128x128_S2 = 1,
    X3_V128x128_S4 = 2,
    X3_V128x256_S3 = 3,
    X3_V128x128_REG = 4,
    X3_V128x64_S2 = 5,
    X3_V128x32_S2 = 6,
    X3_V128x128_W8S3 = 7,
    X3_V128x128_W8S4 = 8,
    X3_V256x128_W8S3 = 9,
    X3_V128x32_S6 = 10,
    X3_V128x64_S5 = 11,
    X3_V128x32_K8 = 12,
    X3_V256x256_W8S2 = 13
};
constexpr int X3_NVARIANT = 13;
int x3_variant_bn(int variant);   // output columns of a tile
int x3_variant_bm(int variant);   // output rows of a tile (X3_BM or 2 X3_BM)
void launch_x3(const X3Args &a, int variant, hipStream_t st);

// ---- dependency-driven sweep: one persistent launch per evaluation --------
//
// Every tile (task, row tile) of the level launches, times every column
// tile, is a work item.  The work items sit in SWEEP_NQ queues, one per XCD:
// queue z holds the z-th XCD group of every level launch, in launch order,
// so a queue keeps the tile order that shares operands through its XCD's
// L2.  A work item starts as soon as the skeleton blocks it reads are
// written: per (skeleton block, column tile) a counter holds the number of
// finished (tile, wave) writes.  A task waits for the producers of its
// leading segments before its first load ("early": children for UP, far
// sources for the couplings) and for those of its trailing segments, the
// ones reading u^ (the parent's transfer, the leaf's VT^T u), only before
// the first of them ("late"), so coupling and near work runs while the
// top-down chain is still in progress.
//
// Progress does not depend on how many workgroups are resident: a
// workgroup claims the next item of its XCD's queue (fetch-and-add), but
// runs it only once every producer of that item is claimed by some
// workgroup; until then it claims and runs the first untaken item in rank
// order (a topological order of all items), whose producers are all
// claimed.  A held item ranks above everything its holder runs first, so
// waits only point to lower ranks and the lowest unfinished item waits for
// nothing.  Every wait is also bounded in time: past SWEEP_TIMEOUT_TICKS it
// sets the error word, after which every workgroup drains without waiting.
//
// Coherence across the XCDs' L2 caches without whole-cache maintenance: a
// producer writes its skeleton panel tile with device-scope stores (sc1:
// written through its XCD's L2), waits for them to complete, then adds to
// the counter (device-scope atomic); a consumer polls the counter with
// device-scope loads and only then loads the tile.  Within one launch no
// cache line of a panel is read before its one write (blocks, row groups
// and column tiles never share a line, and the caches are invalidated at
// every kernel start), so the consumer's first access misses and reads the
// written data.  The agent-scope fences that would make this formal write
// back and invalidate whole L2s at every item and cost more than the
// evaluation itself.
constexpr int SWEEP_NQ = 8;
constexpr int SWEEP_WAVES = 4;            // waves per workgroup: each signals its own writes
constexpr int SWEEP_HEAD_STRIDE = 32;     // int32 per queue head (one 128-byte line each)
constexpr int SWEEP_ERR = SWEEP_NQ * SWEEP_HEAD_STRIDE;   // error word
constexpr int SWEEP_CNT = SWEEP_ERR + SWEEP_HEAD_STRIDE;  // first counter
constexpr uint64_t SWEEP_TIMEOUT_TICKS = 100000000ull;   // 1 s of the 100 MHz realtime clock
constexpr int SWEEP_MAX_POLLS = 1 << 24;   // and at most this many polls per wait or claim (a second bound)

struct alignas(16) SweepTile {
    int32_t task, rt;     // task and row tile (batched: row tile | column tile << 16)
    int32_t g;            // rank (work item rank = g * ncols + column tile)
    int32_t grp;          // batched: group (index into X3SweepArgs::groups)
    int32_t maxtile[SWEEP_NQ];   // per queue: the last tile index of a producer (-1: none)
};

struct alignas(16) SweepTask {
    int32_t ew_b, ew_e;   // early waits [ew_b, ew_e)
    int32_t lw_b, lw_e;   // late waits, before segment late_seg
    int32_t late_seg;     // index within the task's segments, -1: none
    int32_t sig;          // counter block the task's tiles signal (-1: Y, none)
    int32_t pad[2];
};

struct SweepWait {
    int32_t block;    // counter block (skeleton node; u^ blocks after the w^ ones)
    int32_t target;   // finished (tile, wave) writes of the block per column tile
};

// One group of a batched sweep: its plan's operands, panels and lists and
// its output (x.Y, x.y_f64).
struct SweepGroup {
    X3Args x;
    const SweepTask *st;
    const SweepWait *w;
    int32_t ncols;               // column tiles of the group
    int32_t cbase;               // its first counter in the control block
};
struct X3SweepArgs {
    X3Args x;                    // tasks, segs, operands, panels, Y (tiles unused)
    const SweepTile *q;          // the queues, concatenated
    const int2 *qinfo;           // per queue: first entry, tiles
    const SweepTask *st;         // per task
    const SweepWait *w;
    int32_t *ctl;                // heads, error word, counters (zeroed before the launch)
    int ncols;                   // column tiles (batched: 1, every entry one work item)
    // batched sweep (several plans' evaluations in one launch; null for one
    // plan): per group its lists, buffers and output
    const SweepGroup *groups;
};

int sweep_bn(int variant);               // columns of a tile: X3_V128x128_S2, X3_V128x64_S2, X3_V128x32_S2
bool sweep_variant_ok(int variant);
// Zero the control block, then run the sweep on `grid` workgroups.
void launch_x3_sweep(const X3SweepArgs &a, int variant, int grid, int ctl_ints, hipStream_t st);
// the same for a batched sweep (a.groups set)
void launch_x3_sweep_batch(const X3SweepArgs &a, int variant, int grid, int ctl_ints, hipStream_t st);
int sweep_max_grid(int variant);         // resident workgroups the device holds (occupancy x CUs)

}  // namespace mx
