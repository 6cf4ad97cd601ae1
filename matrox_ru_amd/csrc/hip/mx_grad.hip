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

template void launch_grad<double, double>(const GradArgs<double, double> &, hipStream_t);
template void launch_grad<float, float>(const GradArgs<float, float> &, hipStream_t);
template void launch_grad<float, double>(const GradArgs<float, double> &, hipStream_t);

}  // namespace mx
