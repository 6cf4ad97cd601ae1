// Split-BF16 x3 kernels (layout and arithmetic described in mx_bf16x3.h).
#include "mx_bf16x3.h"

namespace mx {

__global__ void split_bf16_f64(const double *__restrict__ in, uint16_t *__restrict__ hi,
                               uint16_t *__restrict__ lo, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double x = in[i];
        uint16_t h = bf16_rne((float)x);
        hi[i] = h;
        lo[i] = bf16_rne((float)(x - (double)bf16_f(h)));
    }
}

__global__ void pack_x3_blocks(const double *__restrict__ src, const X3Block *__restrict__ blocks,
                               uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
    const X3Block bk = blocks[blockIdx.y];
    const int nfk = (bk.K + 31) / 32, nfm = (bk.M + 15) / 16;
    const int64_t lanes = (int64_t)nfm * nfk * 64;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < lanes;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t frag = t >> 6;
        int lane = (int)(t & 63);
        int mi = (int)(frag / nfk), ki = (int)(frag % nfk);
        int row = mi * 16 + (lane & 15);
        int kb = ki * 32 + 8 * (lane >> 4);
        int64_t dst = bk.dst_off + frag * 512 + (int64_t)lane * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int k = kb + e;
            double x = 0.0;
            if (row < bk.M && k < bk.K)
                x = bk.trans ? src[bk.src_off + (int64_t)row * bk.lda + k] : src[bk.src_off + (int64_t)k * bk.lda + row];
            uint16_t h = bf16_rne((float)x);
            hi[dst + e] = h;
            lo[dst + e] = bf16_rne((float)(x - (double)bf16_f(h)));
        }
    }
}

__device__ __forceinline__ bf16x8_t ld_frag_global(const uint16_t *p) {
    u32x4_t v = *reinterpret_cast<const u32x4_t *>(p);
    return __builtin_bit_cast(bf16x8_t, v);
}

__device__ __forceinline__ bf16x8_t ld_frag_lds(const uint16_t *p) {
    u32x4_t v = *reinterpret_cast<const u32x4_t *>(p);
    return __builtin_bit_cast(bf16x8_t, v);
}

// One tile of C (128 x 128) += sum over the task's segments of A_seg * B_seg,
// A packed split-BF16, B FP32 column blocks split on the fly.
__global__ __launch_bounds__(X3_THREADS) void bf16x3_kernel(const Task *__restrict__ tasks,
                                                            const Seg *__restrict__ segs,
                                                            const int2 *__restrict__ tiles, int ntiles,
                                                            int nrhs, int n_ntiles, X3Bufs bufs) {
    constexpr int BS_ELEMS = X3_BN * X3_BKP;   // one of hi/lo
    constexpr int LDS_BYTES_B = 2 * BS_ELEMS * 2;
    constexpr int LDS_BYTES_C = X3_BN * X3_CP * 4;
    constexpr int LDS_BYTES = LDS_BYTES_B > LDS_BYTES_C ? LDS_BYTES_B : LDS_BYTES_C;
    __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_BYTES];
    uint16_t *Bhi = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Blo = Bhi + BS_ELEMS;

    const int tid = threadIdx.x;
    const int total = ntiles * n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / n_ntiles, nt = w % n_ntiles;
    const int2 tl = tiles[tile_id];
    const Task task = tasks[tl.x];
    const int M = task.m;
    const int m0 = tl.y * X3_BM, n0 = nt * X3_BN;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int nfm = (M + 15) / 16;

    x3_f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = x3_f4{0.f, 0.f, 0.f, 0.f};

    // column-chunk loader: thread t covers column t/2, k half (t%2)*16
    const int bcol = tid >> 1, bkh = (tid & 1) * 16;
    const int gcol = n0 + bcol;

    for (int si = task.seg_begin; si < task.seg_end; ++si) {
        const Seg sg = segs[si];
        const int K = sg.k;
        const int nfk = (K + 31) / 32;
        const uint16_t *Ahi = bufs.hi[sg.a_buf] + sg.a_off;
        const uint16_t *Alo = bufs.lo[sg.a_buf] + sg.a_off;
        const float *Bv = bufs.vec[sg.b_buf] + sg.b_row * (int64_t)nrhs;
        for (int kt = 0; kt < nfk; ++kt) {
            // stage the 32 x 128 column chunk, split into hi/lo
            {
                float x[16];
                const int kb = kt * 32 + bkh;
                const float *src = Bv + (int64_t)gcol * sg.b_ld + kb;
#pragma unroll
                for (int e = 0; e < 16; ++e) x[e] = (gcol < nrhs && kb + e < K) ? src[e] : 0.f;
                uint16_t h[16], l[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    h[e] = bf16_rne(x[e]);
                    l[e] = bf16_rne(x[e] - bf16_f(h[e]));
                }
                u32x4_t *dh = reinterpret_cast<u32x4_t *>(Bhi + bcol * X3_BKP + bkh);
                u32x4_t *dl = reinterpret_cast<u32x4_t *>(Blo + bcol * X3_BKP + bkh);
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    u32x4_t vh, vl;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        vh[c] = (uint32_t)h[q * 8 + 2 * c] | ((uint32_t)h[q * 8 + 2 * c + 1] << 16);
                        vl[c] = (uint32_t)l[q * 8 + 2 * c] | ((uint32_t)l[q * 8 + 2 * c + 1] << 16);
                    }
                    dh[q] = vh;
                    dl[q] = vl;
                }
            }
            // A fragments of this wave's 4 row groups
            bf16x8_t ahi[4], alo[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int mi = (m0 + wr * 64) / 16 + i;
                if (mi < nfm) {
                    int64_t off = ((int64_t)mi * nfk + kt) * 512 + lane * 8;
                    ahi[i] = ld_frag_global(Ahi + off);
                    alo[i] = ld_frag_global(Alo + off);
                } else {
                    ahi[i] = __builtin_bit_cast(bf16x8_t, u32x4_t{0, 0, 0, 0});
                    alo[i] = ahi[i];
                }
            }
            __syncthreads();
            bf16x8_t bhi[4], blo[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int col = wc * 64 + j * 16 + (lane & 15);
                int kk = 8 * (lane >> 4);
                bhi[j] = ld_frag_lds(Bhi + col * X3_BKP + kk);
                blo[j] = ld_frag_lds(Blo + col * X3_BKP + kk);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = mfma_bf16(alo[i], bhi[j], acc[i][j]);
                    acc[i][j] = mfma_bf16(ahi[i], blo[j], acc[i][j]);
                    acc[i][j] = mfma_bf16(ahi[i], bhi[j], acc[i][j]);
                }
            __syncthreads();
        }
    }

    float *Cs = reinterpret_cast<float *>(smem);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = wr * 64 + i * 16 + 4 * (lane >> 4) + r;
                int col = wc * 64 + j * 16 + (lane & 15);
                Cs[col * X3_CP + row] = acc[i][j][r];
            }
    __syncthreads();
    float *C = bufs.vec[task.c_buf] + task.c_row * (int64_t)nrhs;
    for (int e = tid; e < X3_BM * X3_BN; e += X3_THREADS) {
        int col = e / X3_BM, row = e % X3_BM;
        int gm = m0 + row, gn = n0 + col;
        if (gm < M && gn < nrhs) {
            float v = Cs[col * X3_CP + row];
            float *p = C + (int64_t)gn * M + gm;
            *p = task.beta ? *p + v : v;
        }
    }
}

void launch_bf16x3(int ntiles, const Task *tasks, const Seg *segs, const int2 *tiles, int nrhs, const X3Bufs &b,
                   hipStream_t st) {
    int nnt = (nrhs + X3_BN - 1) / X3_BN;
    int64_t total = (int64_t)ntiles * nnt;
    if (total <= 0) return;
    hipLaunchKernelGGL(bf16x3_kernel, dim3((unsigned)total), dim3(X3_THREADS), 0, st, tasks, segs, tiles, ntiles,
                       nrhs, nnt, b);
}

}  // namespace mx
