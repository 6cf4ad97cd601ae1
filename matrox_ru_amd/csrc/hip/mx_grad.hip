// NT block-gradient products on MFMA (FP64 and FP32 accumulation).
//
// A workgroup (256 threads = 4 wave64) computes a 64 x 64 tile of one output
// block as 2 x 2 waves of 32 x 32 (2 x 2 MFMA 16x16x4 blocks each).  The
// contraction runs over the nrhs columns of every segment in 16-deep chunks.
// Both operands are stored with the output index fastest (element (i, k) at
// k*ld + i), so a chunk is 16 runs of 64 consecutive elements for X and for
// Z: fully coalesced loads, staged in LDS with a register double buffer (the
// loads of chunk c+1 are in flight during the MFMAs of chunk c).  Ragged
// edges load zeros; the tile leaves through LDS so that stores are 64
// consecutive rows per column.
#include "mx_grad.h"
#include "mx_bf16x3.h"
#include "mx_seg_gemm.h"

namespace mx {

constexpr int GR_BM = 64, GR_BN = 64, GR_BK = 16, GR_P = 65, GR_THREADS = 256;

template <typename T>
struct GStage {
    T x[4];
    T z[4];
};

template <typename T>
__device__ __forceinline__ void grad_load(GStage<T> &s, const GSeg &sg, int k0, int m0, int M, int n0, int N,
                                          int nrhs, const T *const *src, int tid) {
    const T *X = src[sg.x_buf] + sg.x_row * (int64_t)nrhs;
    const T *Z = src[sg.z_buf] + sg.z_row * (int64_t)nrhs;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int e = r * GR_THREADS + tid;
        int kk = e / GR_BM, ii = e % GR_BM;
        int gk = k0 + kk;
        int gm = m0 + ii, gn = n0 + ii;
        s.x[r] = (gm < M && gk < nrhs) ? X[(int64_t)gk * sg.x_ld + gm] : T(0);
        s.z[r] = (gn < N && gk < nrhs) ? Z[(int64_t)gk * sg.z_ld + gn] : T(0);
    }
}

template <typename T>
__device__ __forceinline__ void grad_store(const GStage<T> &s, T *Xs, T *Zs, int tid) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int e = r * GR_THREADS + tid;
        Xs[(e / GR_BM) * GR_P + (e % GR_BM)] = s.x[r];
        Zs[(e / GR_BM) * GR_P + (e % GR_BM)] = s.z[r];
    }
}

template <typename T, typename O>
__global__ __launch_bounds__(GR_THREADS) void grad_kernel(GradArgs<T, O> a) {
    typedef typename MfmaAcc<T>::type acc_t;
    constexpr int CH = GR_BK * GR_P;            // one operand chunk
    constexpr int STAGE = 2 * CH;
    constexpr int SMEM = (2 * STAGE > GR_BN * GR_P) ? 2 * STAGE : GR_BN * GR_P;
    __shared__ T smem[SMEM];

    const int tid = threadIdx.x;
    const int4 tl = a.tiles[xcd_remap(blockIdx.x, a.ntiles)];
    const GTask task = a.tasks[tl.x];
    if (a.out[task.c_buf] == nullptr) return;   // this block type's gradient was not requested
    const int M = task.m, N = task.n;
    const int m0 = tl.y * GR_BM, n0 = tl.z * GR_BN;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;

    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nch = (a.nrhs + GR_BK - 1) / GR_BK;
    const int total = (task.seg_end - task.seg_begin) * nch;
    GStage<T> st;
    if (total > 0) {
        grad_load(st, a.segs[task.seg_begin], 0, m0, M, n0, N, a.nrhs, a.src, tid);
        grad_store(st, smem, smem + CH, tid);
    }
    __syncthreads();
    for (int c = 0; c < total; ++c) {
        const int cur = c & 1;
        const bool more = c + 1 < total;
        if (more) {
            const int s = task.seg_begin + (c + 1) / nch;
            grad_load(st, a.segs[s], ((c + 1) % nch) * GR_BK, m0, M, n0, N, a.nrhs, a.src, tid);
        }
        const T *Xs = smem + cur * STAGE;
        const T *Zs = Xs + CH;
#pragma unroll
        for (int kk = 0; kk < GR_BK; kk += 4) {
            T x[2], z[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) x[i] = Xs[(kk + lk) * GR_P + wr * 32 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 2; ++j) z[j] = Zs[(kk + lk) * GR_P + wc * 32 + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma16x16x4(x[i], z[j], acc[i][j]);
        }
        if (more) {
            T *Xn = smem + (cur ^ 1) * STAGE;
            grad_store(st, Xn, Xn + CH, tid);
        }
        __syncthreads();
    }

    T *Cs = smem;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = wr * 32 + i * 16 + MfmaAcc<T>::row(lk, r);
                int col = wc * 32 + j * 16 + lr;
                Cs[col * GR_P + row] = acc[i][j][r];
            }
    __syncthreads();
    O *C = a.out[task.c_buf] + task.c_off;
#pragma unroll 4
    for (int e = tid; e < GR_BM * GR_BN; e += GR_THREADS) {
        int col = e / GR_BM, row = e % GR_BM;
        int gm = m0 + row, gn = n0 + col;
        if (gm < M && gn < N) C[(int64_t)gn * task.ldc + gm] = (O)Cs[col * GR_P + row];
    }
}

template <typename T, typename O>
void launch_grad(const GradArgs<T, O> &a, hipStream_t st) {
    if (a.ntiles <= 0) return;
    hipLaunchKernelGGL((grad_kernel<T, O>), dim3((unsigned)a.ntiles), dim3(GR_THREADS), 0, st, a);
}

// ---- split-BF16x3 products (FP32 operands, FP32 accumulation) -------------
//
// The same NT products with every operand split as x = hi + lo (bf16) and
// each product taken as lo*hi + hi*lo + hi*hi on v_mfma_f32_16x16x32_bf16.
// A workgroup (4 waves, 2 x 2) computes a 128 x 128 output tile, each wave
// 64 x 64 (4 x 4 MFMA blocks).  The contraction runs over the nrhs columns in
// 32-deep chunks: every thread loads 2 runs of 8 consecutive columns of one
// row of X and of Z (per load instruction a wave reads 64 consecutive rows,
// 256 B), splits them and writes each run as one 16-byte LDS store, so that
// LDS holds [row][32 columns] in bf16, the MFMA operand order (a lane reads 8
// consecutive columns of one row).  The four 16-byte column groups of a row
// are XOR-swizzled with (row / 4) % 4, which keeps both the stores (16
// consecutive rows, one group) and the fragment reads (16 rows, one group)
// free of bank conflicts.  Loads of chunk c+1 are in flight (registers)
// during the MFMAs of chunk c; two LDS stages of 32 KB.

constexpr int GX_BM = 128, GX_BK = 32, GX_THREADS = 256;
constexpr int GX_OP = GX_BM * GX_BK;          // bf16 elements of one operand half (hi or lo)
constexpr int GX_STAGE = 4 * GX_OP;           // X hi, X lo, Z hi, Z lo
typedef float gx_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int gx_slot(int row, int g) { return row * GX_BK + ((g ^ ((row >> 2) & 3)) << 3); }

// Raw bits of one chunk's loads; the masks that zero out-of-range rows and
// columns are applied when the chunk is split (after the MFMAs of the
// previous chunk), so the loads stay in flight meanwhile.
struct GxRegs {
    uint32_t x[2][8];
    uint32_t z[2][8];
    uint32_t xm, zm;       // row in range (all ones) or not (zero)
    uint32_t kbits;        // bit 8*h + e: column k0 + 16*h + kt + e in range
};

// Out-of-range rows and columns load a valid element (index clamped), so
// the loads need no branches and all 32 of a chunk are in flight together.
__device__ __forceinline__ void gx_load(GxRegs &r, const GSeg &sg, int k0, int m0, int M, int n0, int N, int nrhs,
                                        const float *const *src, int tid) {
    const int row = tid & (GX_BM - 1);
    r.xm = m0 + row < M ? ~0u : 0u;
    r.zm = n0 + row < N ? ~0u : 0u;
    const uint32_t xo = (uint32_t)min(m0 + row, M - 1), zo = (uint32_t)min(n0 + row, N - 1);
    const int kt = __builtin_amdgcn_readfirstlane((tid >> 7) * 8);
    const uint32_t *X = reinterpret_cast<const uint32_t *>(src[sg.x_buf] + sg.x_row * (int64_t)nrhs);
    const uint32_t *Z = reinterpret_cast<const uint32_t *>(src[sg.z_buf] + sg.z_row * (int64_t)nrhs);
    uint32_t kbits = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + h * 16 + kt + e;
            const int kc = min(k, nrhs - 1);
            r.x[h][e] = X[(int64_t)kc * sg.x_ld + xo];
            r.z[h][e] = Z[(int64_t)kc * sg.z_ld + zo];
            kbits |= (k < nrhs ? 1u : 0u) << (8 * h + e);
        }
    }
    r.kbits = kbits;
}

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi), both rounded to nearest
// even (v_cvt_pk_bf16_f32), 8 values to two 16-byte fragments.
__device__ __forceinline__ void gx_split8(const uint32_t (&raw)[8], uint32_t mask, uint32_t kbits, u32x4_t &hi,
                                          u32x4_t &lo) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f2_t __attribute__((ext_vector_type(2)));
    uint32_t h[4], l[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t m0 = (kbits >> (2 * e)) & 1u ? mask : 0u;
        const uint32_t m1 = (kbits >> (2 * e + 1)) & 1u ? mask : 0u;
        const f2_t x = {__uint_as_float(raw[2 * e] & m0), __uint_as_float(raw[2 * e + 1] & m1)};
        const bf16x2_t xh = __builtin_convertvector(x, bf16x2_t);
        const f2_t rest = x - __builtin_convertvector(xh, f2_t);
        h[e] = __builtin_bit_cast(uint32_t, xh);
        l[e] = __builtin_bit_cast(uint32_t, __builtin_convertvector(rest, bf16x2_t));
    }
    hi = u32x4_t{h[0], h[1], h[2], h[3]};
    lo = u32x4_t{l[0], l[1], l[2], l[3]};
}

__device__ __forceinline__ void gx_store(const GxRegs &r, uint16_t *stage, int tid) {
    const int row = tid & (GX_BM - 1);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int slot = gx_slot(row, h * 2 + (tid >> 7));
        u32x4_t hi, lo;
        gx_split8(r.x[h], r.xm, r.kbits >> (8 * h), hi, lo);
        *reinterpret_cast<u32x4_t *>(stage + slot) = hi;
        *reinterpret_cast<u32x4_t *>(stage + GX_OP + slot) = lo;
        gx_split8(r.z[h], r.zm, r.kbits >> (8 * h), hi, lo);
        *reinterpret_cast<u32x4_t *>(stage + 2 * GX_OP + slot) = hi;
        *reinterpret_cast<u32x4_t *>(stage + 3 * GX_OP + slot) = lo;
    }
}

__device__ __forceinline__ bf16x8_t gx_frag(const uint16_t *half, int row, int g) {
    return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t *>(half + gx_slot(row, g)));
}

template <typename O>
__global__ __launch_bounds__(GX_THREADS) void grad_x3_kernel(GradArgs<float, O> a) {
    constexpr int CP = GX_BM + 4;        // padded column stride of the output tile staged in the epilogue
    constexpr int SMEM = (2 * GX_STAGE > CP * GX_BM * 2) ? 2 * GX_STAGE : CP * GX_BM * 2;   // bf16 elements
    __shared__ __attribute__((aligned(16))) uint16_t smem[SMEM];

    const int tid = threadIdx.x;
    const int4 tl = a.tiles[xcd_remap(blockIdx.x, a.ntiles)];
    const GTask task = a.tasks[tl.x];
    if (a.out[task.c_buf] == nullptr) return;   // this block type's gradient was not requested
    const int M = task.m, N = task.n;
    const int m0 = tl.y * GX_BM, n0 = tl.z * GX_BM;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lg = lane >> 4;

    gx_f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = gx_f4{0.f, 0.f, 0.f, 0.f};

    const int nch = (a.nrhs + GX_BK - 1) / GX_BK;
    const int total = (task.seg_end - task.seg_begin) * nch;
    GxRegs r;
    if (total > 0) {
        gx_load(r, a.segs[task.seg_begin], 0, m0, M, n0, N, a.nrhs, a.src, tid);
        gx_store(r, smem, tid);
    }
    __syncthreads();
    for (int c = 0; c < total; ++c) {
        const bool more = c + 1 < total;
        if (more) {
            const int s = task.seg_begin + (c + 1) / nch;
            gx_load(r, a.segs[s], ((c + 1) % nch) * GX_BK, m0, M, n0, N, a.nrhs, a.src, tid);
        }
        const uint16_t *st = smem + (c & 1) * GX_STAGE;
        bf16x8_t xh[4], xl[4], zh[4], zl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = wr * 64 + i * 16 + lr;
            xh[i] = gx_frag(st, row, lg);
            xl[i] = gx_frag(st + GX_OP, row, lg);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = wc * 64 + j * 16 + lr;
            zh[j] = gx_frag(st + 2 * GX_OP, row, lg);
            zl[j] = gx_frag(st + 3 * GX_OP, row, lg);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[i], zh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], zl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], zh[j], acc[i][j], 0, 0, 0);
            }
        if (more) gx_store(r, smem + ((c + 1) & 1) * GX_STAGE, tid);
        __syncthreads();
    }

    // C(row, col) of block (i, j), register q: row 4*(lane/16)+q, column lane%16
    float *Cs = reinterpret_cast<float *>(smem);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = wr * 64 + i * 16 + lg * 4 + q;
                const int col = wc * 64 + j * 16 + lr;
                Cs[col * CP + row] = acc[i][j][q];
            }
    __syncthreads();
    O *C = a.out[task.c_buf] + task.c_off;
#pragma unroll 4
    for (int e = tid; e < GX_BM * GX_BM; e += GX_THREADS) {
        const int col = e / GX_BM, row = e % GX_BM;
        const int gm = m0 + row, gn = n0 + col;
        if (gm < M && gn < N) C[(int64_t)gn * task.ldc + gm] = (O)Cs[col * CP + row];
    }
}

template <typename O>
void launch_grad_x3(const GradArgs<float, O> &a, hipStream_t st) {
    if (a.ntiles <= 0) return;
    hipLaunchKernelGGL((grad_x3_kernel<O>), dim3((unsigned)a.ntiles), dim3(GX_THREADS), 0, st, a);
}

template void launch_grad_x3<float>(const GradArgs<float, float> &, hipStream_t);
template void launch_grad_x3<double>(const GradArgs<float, double> &, hipStream_t);

template void launch_grad<double, double>(const GradArgs<double, double> &, hipStream_t);
template void launch_grad<float, float>(const GradArgs<float, float> &, hipStream_t);
template void launch_grad<float, double>(const GradArgs<float, double> &, hipStream_t);

}  // namespace mx
