// Split-BF16 x3 kernels (layouts and arithmetic described in mx_bf16x3.h).
#include "mx_bf16x3.h"
#include "mx_grad.h"

#include <type_traits>

#include <algorithm>

namespace mx {

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

// Eight columns per workgroup, 32 threads per column, each thread two
// consecutive rows: the loads read 64 consecutive rows of a column
// (coalesced), the stores write 64-byte runs of the panel (a 32-row chunk of
// one column), and no index needs a division.  hi = bf16(x) and
// lo = bf16(x - hi), rounded to nearest even (v_cvt_pk_bf16_f32; FP64 input
// goes through FP32 first); rows past R are written as zeros.
typedef __bf16 x3_bf16x2_t __attribute__((ext_vector_type(2)));
typedef float x3_f2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void split2(double a, double b, uint32_t &h, uint32_t &l) {
    const x3_bf16x2_t xh = __builtin_convertvector((x3_f2_t{(float)a, (float)b}), x3_bf16x2_t);
    const x3_f2_t hf = __builtin_convertvector(xh, x3_f2_t);
    const x3_f2_t r = {(float)(a - (double)hf.x), (float)(b - (double)hf.y)};
    h = __builtin_bit_cast(uint32_t, xh);
    l = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, x3_bf16x2_t));
}
__device__ __forceinline__ void split2(float a, float b, uint32_t &h, uint32_t &l) {
    const x3_f2_t x = {a, b};
    const x3_bf16x2_t xh = __builtin_convertvector(x, x3_bf16x2_t);
    h = __builtin_bit_cast(uint32_t, xh);
    l = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - __builtin_convertvector(xh, x3_f2_t), x3_bf16x2_t));
}

constexpr int X3_SPLIT_COLS = 8;   // columns per workgroup of the panel kernels

template <typename TW>
__global__ __launch_bounds__(256) void x3_split_w_kernel(const TW *__restrict__ W, const X3Rows *__restrict__ blocks,
                                                         int nrhs, uint16_t *__restrict__ hi,
                                                         uint16_t *__restrict__ lo) {
    const X3Rows b = blocks[blockIdx.y];
    const int R = b.R, Rp = (R + 31) & ~31;
    const int lane = threadIdx.x & 31, cw = threadIdx.x >> 5;
    const TW *src = W + b.src_row * nrhs;
    // panel rows start at multiples of 32, so every (even) offset below is 4-byte aligned
    uint32_t *ph = reinterpret_cast<uint32_t *>(hi + b.prow * nrhs);
    uint32_t *pl = reinterpret_cast<uint32_t *>(lo + b.prow * nrhs);
    typedef typename std::conditional<sizeof(TW) == 8, uint64_t, uint32_t>::type bits_t;
    const bits_t *src_bits = reinterpret_cast<const bits_t *>(src);
    for (int col = blockIdx.x * X3_SPLIT_COLS + cw; col < nrhs; col += gridDim.x * X3_SPLIT_COLS) {
        const bits_t *s = src_bits + (int64_t)col * R;
        // four row pairs per lane per pass, all loads issued before any is
        // used; rows past R load row R-1 and are masked to zero (no branches)
        for (int kb = 0; kb < Rp; kb += 256) {
            bits_t v[4][2];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int e = 0; e < 2; ++e) v[q][e] = s[min(kb + 64 * q + 2 * lane + e, R - 1)];
            // keep the eight loads together (not sunk into the guarded stores below)
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int e = 0; e < 2; ++e) asm volatile("" : "+v"(v[q][e]));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = kb + 64 * q + 2 * lane;
                if (k >= Rp) break;
                const bits_t m0 = k < R ? ~bits_t(0) : bits_t(0), m1 = k + 1 < R ? ~bits_t(0) : bits_t(0);
                uint32_t h, l;
                split2(__builtin_bit_cast(TW, bits_t(v[q][0] & m0)), __builtin_bit_cast(TW, bits_t(v[q][1] & m1)), h, l);
                const int64_t o = (((int64_t)(k >> 5) * nrhs + col) * 32 + (k & 31)) >> 1;
                ph[o] = h;
                pl[o] = l;
            }
        }
    }
}

void launch_x3_split_w(const void *W, int w_f64, const X3Rows *rows, int nblocks, int nrhs, uint16_t *hi,
                       uint16_t *lo, hipStream_t st) {
    const int gx = std::min(128, (nrhs + X3_SPLIT_COLS - 1) / X3_SPLIT_COLS);
    for (int b0 = 0; b0 < nblocks; b0 += 65535) {
        const dim3 grid((unsigned)gx, (unsigned)std::min(65535, nblocks - b0));
        if (w_f64)
            hipLaunchKernelGGL(x3_split_w_kernel<double>, grid, dim3(256), 0, st, (const double *)W, rows + b0, nrhs,
                               hi, lo);
        else
            hipLaunchKernelGGL(x3_split_w_kernel<float>, grid, dim3(256), 0, st, (const float *)W, rows + b0, nrhs,
                               hi, lo);
    }
}

// Split panels back to FP32 blocks in the reference skeleton layout: block b
// (R rows at panel row prow) -> out[src_row*nrhs + col*R + k] = hi + lo.
// Same thread layout as the split kernel.
__global__ __launch_bounds__(256) void x3_panels_to_f32_kernel(const uint16_t *__restrict__ hi,
                                                               const uint16_t *__restrict__ lo,
                                                               const X3Rows *__restrict__ blocks, int nrhs,
                                                               float *__restrict__ out) {
    const X3Rows b = blocks[blockIdx.y];
    const int R = b.R;
    const int lane = threadIdx.x & 31, cw = threadIdx.x >> 5;
    const uint16_t *ph = hi + b.prow * nrhs;
    const uint16_t *pl = lo + b.prow * nrhs;
    float *dst = out + b.src_row * nrhs;
    for (int col = blockIdx.x * X3_SPLIT_COLS + cw; col < nrhs; col += gridDim.x * X3_SPLIT_COLS) {
        float *d = dst + (int64_t)col * R;
        for (int k = lane; k < R; k += 32) {
            const int64_t o = ((int64_t)(k >> 5) * nrhs + col) * 32 + (k & 31);
            d[k] = bf16_f(ph[o]) + bf16_f(pl[o]);
        }
    }
}

void launch_x3_panels_to_f32(const uint16_t *hi, const uint16_t *lo, const X3Rows *blocks, int nblocks, int nrhs,
                             float *out, hipStream_t st) {
    const int gx = std::min(128, (nrhs + X3_SPLIT_COLS - 1) / X3_SPLIT_COLS);
    for (int b0 = 0; b0 < nblocks; b0 += 65535) {
        const int nb = std::min(65535, nblocks - b0);
        hipLaunchKernelGGL(x3_panels_to_f32_kernel, dim3((unsigned)gx, (unsigned)nb), dim3(256), 0, st, hi, lo,
                           blocks + b0, nrhs, out);
    }
}

__device__ __forceinline__ bf16x8_t ld16(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t *>(p));
}

__device__ __forceinline__ x3_f4 mfma_bf16(bf16x8_t a, bf16x8_t b, x3_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Main loop: a pipeline of NST stages through LDS.  A workgroup of WM x WN
// waves computes a (64 WM) x (64 WN) tile, each wave 64 x 64.  Per 32-deep
// chunk it needs 8 (WM + WN) fragments of 1 KB (4 WM A row groups and 4 WN B
// column groups, hi and lo), dealt evenly over the waves; each moves its
// share with global_load_lds_dwordx4, the gfx950 direct global->LDS load (one
// instruction per fragment, lane l writing bytes 16l..16l+15, no VGPRs).
// Loads run NST-1 chunks ahead of the MFMAs.  Row groups past the block and
// columns past nrhs load a valid neighbour (clamped) and are masked or
// discarded.  A stage holds A fragments [row group][hi, lo] followed by B
// fragments [column group][hi, lo], so each wave's 8 A and 8 B fragments are
// two contiguous 8 KB runs.

// Tile geometry: BM x BN output tile, NW waves.
template <int BM_, int BN_, int NW_>
struct X3Geo {
    static constexpr int NW = NW_;                     // waves
    static constexpr int THREADS = 64 * NW;
    static constexpr int BM = BM_, BN = BN_;
    static constexpr int FA = BM / 8, FB = BN / 8;     // fragments per chunk (16-row/column groups x hi, lo)
    static constexpr int PER = (FA + FB) / NW;         // DMA loads per wave per chunk
    static constexpr int STAGE = (FA + FB) * X3_FRAG;  // uint16 per stage
    static_assert(PER * NW == FA + FB, "fragments must divide evenly over the waves");
};

// WM x WN waves of 64 x (16 NJ)
template <int WM, int WN, int NJ = 4>
struct X3Shape : X3Geo<64 * WM, 16 * NJ * WN, WM * WN> {};

__device__ __forceinline__ void lds_dma16(const uint16_t *src, uint16_t *dst_wave) {
    __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)dst_wave, 16, 0, 0);
}

// Per-wave DMA source pointers.  Fragment t of the wave (f = wv*PER + t)
// is an A fragment (f < FA: row group f/2, hi or lo) or a B fragment; its
// source advances by a fixed stride per chunk (one packed fragment for A, 32
// panel rows for B), so the per-chunk work is PER pointer increments and PER
// DMA instructions.  The segment descriptor is read (from LDS) once per
// segment, when its chunks start.
template <typename S>
struct X3Loader {
    const uint16_t *p[S::PER];
    int seg, left;
};

template <typename S>
__device__ __forceinline__ void x3_seg_begin(X3Loader<S> &L, const X3Args &a, const Seg &sg, int wv, int lane,
                                             int mtile16, int nfm, int n0) {
    const int nfk = (sg.k + 31) / 32;
    const int64_t pb = sg.b_row * (int64_t)a.nrhs + 8 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < S::PER; ++t) {
        const int f = wv * S::PER + t;   // wave-uniform
        if (f < S::FA) {
            const int mi = min(mtile16 + (f >> 1), nfm - 1);
            L.p[t] = ((f & 1) ? a.alo[sg.a_buf] : a.ahi[sg.a_buf]) + sg.a_off + lane * 8 +
                     (int64_t)mi * nfk * X3_FRAG;
        } else {
            const int g = f - S::FA;
            const int col = min(n0 + (g >> 1) * 16 + (lane & 15), a.nrhs - 1);
            L.p[t] = ((g & 1) ? a.plo[sg.b_buf] : a.phi[sg.b_buf]) + pb + (int64_t)col * 32;
        }
    }
    L.left = nfk;
}

// DMA of the next chunk into a stage (segments with k == 0 are skipped);
// on_seg(si) runs before the first load of segment si
template <typename S, typename OnSeg>
__device__ __forceinline__ void x3_issue(X3Loader<S> &L, const X3Args &a, const Seg *sseg, uint16_t *stage, int wv,
                                         int lane, int mtile16, int nfm, int n0, OnSeg &&on_seg) {
    while (L.left == 0) {
        ++L.seg;
        on_seg(L.seg);
        x3_seg_begin(L, a, sseg[L.seg], wv, lane, mtile16, nfm, n0);
    }
    const int64_t bstep = (int64_t)a.nrhs * 32;
#pragma unroll
    for (int t = 0; t < S::PER; ++t) {
        const int f = wv * S::PER + t;
        lds_dma16(L.p[t], stage + f * X3_FRAG);
        L.p[t] += f < S::FA ? (int64_t)X3_FRAG : bstep;
    }
    --L.left;
}

// The 16 fragments a wave multiplies per chunk (its 4 A row groups and 4 B
// column groups, hi and lo) are 8 KB runs at abase and bbase (LDS byte
// offsets).  They are read in one asm block that waits for its own reads:
// LDS reads the compiler can see would be ordered after every outstanding
// LDS-DMA load, which would drain the pipeline each chunk.  Outputs are
// early-clobber so no result register aliases an address operand.
__device__ __forceinline__ void lds_read_frags(uint32_t abase, uint32_t bbase, u32x4_t (&A)[8], u32x4_t (&B)[8]) {
    asm volatile(
        "ds_read_b128 %0, %16 offset:0\n\t"
        "ds_read_b128 %1, %16 offset:1024\n\t"
        "ds_read_b128 %2, %16 offset:2048\n\t"
        "ds_read_b128 %3, %16 offset:3072\n\t"
        "ds_read_b128 %4, %16 offset:4096\n\t"
        "ds_read_b128 %5, %16 offset:5120\n\t"
        "ds_read_b128 %6, %16 offset:6144\n\t"
        "ds_read_b128 %7, %16 offset:7168\n\t"
        "ds_read_b128 %8, %17 offset:0\n\t"
        "ds_read_b128 %9, %17 offset:1024\n\t"
        "ds_read_b128 %10, %17 offset:2048\n\t"
        "ds_read_b128 %11, %17 offset:3072\n\t"
        "ds_read_b128 %12, %17 offset:4096\n\t"
        "ds_read_b128 %13, %17 offset:5120\n\t"
        "ds_read_b128 %14, %17 offset:6144\n\t"
        "ds_read_b128 %15, %17 offset:7168\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(A[5]), "=&v"(A[6]), "=&v"(A[7]), "=&v"(B[0]), "=&v"(B[1]), "=&v"(B[2]), "=&v"(B[3]), "=&v"(B[4]), "=&v"(B[5]), "=&v"(B[6]), "=&v"(B[7])
        : "v"(abase), "v"(bbase)
        : "memory");
}

// 8 A and 4 B fragments (waves of 64 x 32)
__device__ __forceinline__ void lds_read_frags(uint32_t abase, uint32_t bbase, u32x4_t (&A)[8], u32x4_t (&B)[4]) {
    asm volatile(
        "ds_read_b128 %0, %12 offset:0\n\t"
        "ds_read_b128 %1, %12 offset:1024\n\t"
        "ds_read_b128 %2, %12 offset:2048\n\t"
        "ds_read_b128 %3, %12 offset:3072\n\t"
        "ds_read_b128 %4, %12 offset:4096\n\t"
        "ds_read_b128 %5, %12 offset:5120\n\t"
        "ds_read_b128 %6, %12 offset:6144\n\t"
        "ds_read_b128 %7, %12 offset:7168\n\t"
        "ds_read_b128 %8, %13 offset:0\n\t"
        "ds_read_b128 %9, %13 offset:1024\n\t"
        "ds_read_b128 %10, %13 offset:2048\n\t"
        "ds_read_b128 %11, %13 offset:3072\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(A[5]), "=&v"(A[6]), "=&v"(A[7]),
          "=&v"(B[0]), "=&v"(B[1]), "=&v"(B[2]), "=&v"(B[3])
        : "v"(abase), "v"(bbase)
        : "memory");
}

// 8 A and 2 B fragments (waves of 64 x 16)
__device__ __forceinline__ void lds_read_frags(uint32_t abase, uint32_t bbase, u32x4_t (&A)[8], u32x4_t (&B)[2]) {
    asm volatile(
        "ds_read_b128 %0, %10 offset:0\n\t"
        "ds_read_b128 %1, %10 offset:1024\n\t"
        "ds_read_b128 %2, %10 offset:2048\n\t"
        "ds_read_b128 %3, %10 offset:3072\n\t"
        "ds_read_b128 %4, %10 offset:4096\n\t"
        "ds_read_b128 %5, %10 offset:5120\n\t"
        "ds_read_b128 %6, %10 offset:6144\n\t"
        "ds_read_b128 %7, %10 offset:7168\n\t"
        "ds_read_b128 %8, %11 offset:0\n\t"
        "ds_read_b128 %9, %11 offset:1024\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(A[5]), "=&v"(A[6]), "=&v"(A[7]),
          "=&v"(B[0]), "=&v"(B[1])
        : "v"(abase), "v"(bbase)
        : "memory");
}

// 8 B fragments (the second half of a 64 x 128 wave's columns)
__device__ __forceinline__ void lds_read_frags(uint32_t bbase, u32x4_t (&B)[8]) {
    asm volatile(
        "ds_read_b128 %0, %8 offset:0\n\t"
        "ds_read_b128 %1, %8 offset:1024\n\t"
        "ds_read_b128 %2, %8 offset:2048\n\t"
        "ds_read_b128 %3, %8 offset:3072\n\t"
        "ds_read_b128 %4, %8 offset:4096\n\t"
        "ds_read_b128 %5, %8 offset:5120\n\t"
        "ds_read_b128 %6, %8 offset:6144\n\t"
        "ds_read_b128 %7, %8 offset:7168\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(B[0]), "=&v"(B[1]), "=&v"(B[2]), "=&v"(B[3]), "=&v"(B[4]), "=&v"(B[5]), "=&v"(B[6]), "=&v"(B[7])
        : "v"(bbase)
        : "memory");
}

__device__ __forceinline__ uint32_t lds_offset(const uint16_t *p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) uint16_t *)p;
}

// Fragments of the current chunk from one stage: 4 A row groups by NJ B
// column groups, 12 NJ MFMAs.
// Waves of 64 x 128, pipelined: the 24 MFMAs of column groups J0, J0+1
// (acc += al bh, then ah bl, then ah bh, for 4 row groups) issued in one
// asm block behind the LDS reads of the next four B fragments (N, 1 KB apart
// from byte offset rd), so that the reads proceed while the MFMAs run; the
// block waits for its reads before it ends.  The three products of an
// accumulator keep their order (results are those of the builtins); each
// accumulator is the source C of the MFMA that wrote it, which needs no
// wait states.
template <int J0>
__device__ __forceinline__ void x3_mma24_read4(x3_f4 (&c)[4][8], const u32x4_t (&A)[8], const u32x4_t (&B)[4],
                                               uint32_t rd, u32x4_t (&N)[4]) {
    asm volatile(
        "ds_read_b128 %0, %24 offset:0\n\t"
        "ds_read_b128 %1, %24 offset:1024\n\t"
        "ds_read_b128 %2, %24 offset:2048\n\t"
        "ds_read_b128 %3, %24 offset:3072\n\t"
        // al * bh
        "v_mfma_f32_16x16x32_bf16 %4, %13, %20, %4\n\t"
        "v_mfma_f32_16x16x32_bf16 %5, %13, %22, %5\n\t"
        "v_mfma_f32_16x16x32_bf16 %6, %15, %20, %6\n\t"
        "v_mfma_f32_16x16x32_bf16 %7, %15, %22, %7\n\t"
        "v_mfma_f32_16x16x32_bf16 %8, %17, %20, %8\n\t"
        "v_mfma_f32_16x16x32_bf16 %9, %17, %22, %9\n\t"
        "v_mfma_f32_16x16x32_bf16 %10, %19, %20, %10\n\t"
        "v_mfma_f32_16x16x32_bf16 %11, %19, %22, %11\n\t"
        // ah * bl
        "v_mfma_f32_16x16x32_bf16 %4, %12, %21, %4\n\t"
        "v_mfma_f32_16x16x32_bf16 %5, %12, %23, %5\n\t"
        "v_mfma_f32_16x16x32_bf16 %6, %14, %21, %6\n\t"
        "v_mfma_f32_16x16x32_bf16 %7, %14, %23, %7\n\t"
        "v_mfma_f32_16x16x32_bf16 %8, %16, %21, %8\n\t"
        "v_mfma_f32_16x16x32_bf16 %9, %16, %23, %9\n\t"
        "v_mfma_f32_16x16x32_bf16 %10, %18, %21, %10\n\t"
        "v_mfma_f32_16x16x32_bf16 %11, %18, %23, %11\n\t"
        // ah * bh
        "v_mfma_f32_16x16x32_bf16 %4, %12, %20, %4\n\t"
        "v_mfma_f32_16x16x32_bf16 %5, %12, %22, %5\n\t"
        "v_mfma_f32_16x16x32_bf16 %6, %14, %20, %6\n\t"
        "v_mfma_f32_16x16x32_bf16 %7, %14, %22, %7\n\t"
        "v_mfma_f32_16x16x32_bf16 %8, %16, %20, %8\n\t"
        "v_mfma_f32_16x16x32_bf16 %9, %16, %22, %9\n\t"
        "v_mfma_f32_16x16x32_bf16 %10, %18, %20, %10\n\t"
        "v_mfma_f32_16x16x32_bf16 %11, %18, %22, %11\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(N[0]), "=&v"(N[1]), "=&v"(N[2]), "=&v"(N[3]), "+v"(c[0][J0]), "+v"(c[0][J0 + 1]), "+v"(c[1][J0]),
          "+v"(c[1][J0 + 1]), "+v"(c[2][J0]), "+v"(c[2][J0 + 1]), "+v"(c[3][J0]), "+v"(c[3][J0 + 1])
        : "v"(A[0]), "v"(A[1]), "v"(A[2]), "v"(A[3]), "v"(A[4]), "v"(A[5]), "v"(A[6]), "v"(A[7]), "v"(B[0]),
          "v"(B[1]), "v"(B[2]), "v"(B[3]), "v"(rd)
        : "memory");
}

// MASK: zero the A fragments of row groups past the block (their DMA loads
// read a clamped neighbour).  Without it the caller clears those rows of
// the accumulators once the chunks are done (x3_clear_rows), which gives
// the same result: the products of clamped fragments are finite and only
// reach those rows.
template <int NJ, bool MASK = true>
__device__ __forceinline__ void x3_compute(x3_f4 (&acc)[4][NJ], const uint16_t *stage, int fa, int lane, int wr,
                                           int wc, int mtile16, int nfm) {
    const uint32_t base = lds_offset(stage) + lane * 16;
    if constexpr (NJ == 8) {
        // waves of 64 x 128: A and the first quarter of B (column groups
        // 0-1), then three quarters of the MFMAs each behind the reads of
        // the next quarter, the last quarter by builtins (so the compiler
        // sees the last writes of the chunk)
        static_assert(!MASK, "the pipelined form clears rows past the block afterwards");
        const uint32_t bb = base + fa * 1024 + wc * 16 * 1024;
        u32x4_t A[8], B0[4], B1[4];
        lds_read_frags(base + wr * 8192, bb, A, B0);
        x3_mma24_read4<0>(acc, A, B0, bb + 4096, B1);
        x3_mma24_read4<2>(acc, A, B1, bb + 8192, B0);
        x3_mma24_read4<4>(acc, A, B0, bb + 12288, B1);
        bf16x8_t ah[4], al[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ah[i] = __builtin_bit_cast(bf16x8_t, A[2 * i]);
            al[i] = __builtin_bit_cast(bf16x8_t, A[2 * i + 1]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, B1[2 * j]);
                const bf16x8_t bl = __builtin_bit_cast(bf16x8_t, B1[2 * j + 1]);
                x3_f4 &c = acc[i][6 + j];
                c = mfma_bf16(al[i], bh, c);
                c = mfma_bf16(ah[i], bl, c);
                c = mfma_bf16(ah[i], bh, c);
            }
    } else {
        u32x4_t A[8], B[2 * NJ];
        lds_read_frags(base + wr * 8192, base + fa * 1024 + wc * (2 * NJ) * 1024, A, B);
        bf16x8_t ah[4], al[4], bh[NJ], bl[NJ];
        const u32x4_t z = u32x4_t{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bool ok = !MASK || mtile16 + wr * 4 + i < nfm;
            ah[i] = __builtin_bit_cast(bf16x8_t, ok ? A[2 * i] : z);
            al[i] = __builtin_bit_cast(bf16x8_t, ok ? A[2 * i + 1] : z);
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            bh[j] = __builtin_bit_cast(bf16x8_t, B[2 * j]);
            bl[j] = __builtin_bit_cast(bf16x8_t, B[2 * j + 1]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
                acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
                acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
            }
    }
}

// Epilogue of a wave's (16 NI) x (16 NJ) accumulator block (NI row groups x
// NJ column groups of 16 x 16; the wave's rows start at m0 + wr * 16 NI, its
// columns at n0 + wc * 16 NJ).
//  * Y: each 64 x (16 NJ) part is staged in the wave's LDS (rows
//    XOR-swizzled by column so the 16 lanes of a write hit distinct banks)
//    and stored as 64 consecutive rows per column (beta = 1 adds to Y).
//  * Skeleton buffers: written as split panels straight from registers (rows
//    up to the padded height; padding rows are 0), split with the hardware
//    round-to-nearest-even conversion; beta = 1 adds the panel's current
//    value first.
// 8-byte panel store: plain, or (DEV) a device-scope store (sc1), written
// through to where every XCD's L2 misses find it
template <bool DEV>
__device__ __forceinline__ void st_panel(uint16_t *p, u32x2_t v) {
    if constexpr (DEV)
        __hip_atomic_store(reinterpret_cast<uint64_t *>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    else
        *reinterpret_cast<u32x2_t *>(p) = v;
}

template <int NI, int NJ, bool DEV = false>
__device__ __forceinline__ void x3_epilogue(x3_f4 (&acc)[NI][NJ], uint16_t *stg, const X3Args &a, const Task &task,
                                            int lane, int wv, int wr, int wc, int m0, int n0) {
    constexpr int WC = 16 * NJ;   // columns of the wave
    constexpr int JG = NJ > 4 ? 4 : NJ;   // column groups staged at a time (64 x 16 JG floats per wave)
    constexpr int GC = 16 * JG;
    const int M = task.m;
    if (task.c_buf == VEC_Y) {
        float *Cs = reinterpret_cast<float *>(stg) + wv * 64 * GC;
        const int64_t base = task.c_row * (int64_t)a.nrhs;
#pragma unroll
        for (int hj = 0; hj < (NI / 4) * (NJ / JG); ++hj) {
            const int h = hj / (NJ / JG), j0 = (hj % (NJ / JG)) * JG;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < JG; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int col = j * 16 + (lane & 15), row = i * 16 + 4 * (lane >> 4) + r;
                        Cs[col * 64 + (row ^ ((col & 15) << 2))] = acc[h * 4 + i][j0 + j][r];
                    }
            __builtin_amdgcn_wave_barrier();
            const int rb = m0 + wr * 16 * NI + h * 64, cb = n0 + wc * WC + j0 * 16;
            for (int e = lane; e < 64 * GC; e += 64) {
                const int col = e >> 6, row = e & 63;
                const int gm = rb + row, gn = cb + col;
                if (gm < M && gn < a.nrhs) {
                    const float v = Cs[col * 64 + (row ^ ((col & 15) << 2))];
                    const int64_t o = base + (int64_t)gn * M + gm;
                    if (a.y_f64) {
                        double *y = static_cast<double *>(a.Y) + o;
                        *y = task.beta ? *y + (double)v : (double)v;
                    } else {
                        float *y = static_cast<float *>(a.Y) + o;
                        *y = task.beta ? *y + v : v;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    } else {
        const int Mp = (M + 31) / 32 * 32;
        uint16_t *Ph = a.phi[task.c_buf] + task.c_row * (int64_t)a.nrhs;
        uint16_t *Pl = a.plo[task.c_buf] + task.c_row * (int64_t)a.nrhs;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int r0 = m0 + wr * 16 * NI + i * 16 + 4 * (lane >> 4);
            if (r0 >= Mp) continue;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int col = n0 + wc * WC + j * 16 + (lane & 15);
                if (col >= a.nrhs) continue;
                const int64_t o = ((int64_t)(r0 >> 5) * a.nrhs + col) * 32 + (r0 & 31);
                float v[4] = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
                if (task.beta) {
                    const u32x2_t oh = *reinterpret_cast<const u32x2_t *>(Ph + o);
                    const u32x2_t ol = *reinterpret_cast<const u32x2_t *>(Pl + o);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const uint32_t wh = oh[r >> 1], wl = ol[r >> 1];
                        v[r] += bf16_f((uint16_t)(wh >> (16 * (r & 1)))) + bf16_f((uint16_t)(wl >> (16 * (r & 1))));
                    }
                }
                uint32_t hw[2], lw[2];
                split2(v[0], v[1], hw[0], lw[0]);
                split2(v[2], v[3], hw[1], lw[1]);
                st_panel<DEV>(Ph + o, u32x2_t{hw[0], hw[1]});
                st_panel<DEV>(Pl + o, u32x2_t{lw[0], lw[1]});
            }
        }
    }
}

// Register-staged variant of the 128 x 128 kernel: fragments travel
// global -> VGPR (global_load_dwordx4, 16 B per lane per fragment) -> LDS
// (ds_write_b128) instead of by LDS-DMA.  Two LDS stages; the loads of chunk
// c+2 are issued after chunk c+1 is written to LDS, so they are in flight
// during the MFMAs of chunk c+1.
template <typename S>
__device__ __forceinline__ void x3_load_regs(X3Loader<S> &L, const X3Args &a, const Seg *sseg, u32x4_t (&R)[S::PER],
                                             int wv, int lane, int mtile16, int nfm, int n0) {
    while (L.left == 0) {
        ++L.seg;
        x3_seg_begin(L, a, sseg[L.seg], wv, lane, mtile16, nfm, n0);
    }
    const int64_t bstep = (int64_t)a.nrhs * 32;
#pragma unroll
    for (int t = 0; t < S::PER; ++t) {
        const int f = wv * S::PER + t;
        R[t] = *reinterpret_cast<const u32x4_t *>(L.p[t]);
        L.p[t] += f < S::FA ? (int64_t)X3_FRAG : bstep;
    }
    --L.left;
}

template <typename S>
__device__ __forceinline__ void x3_store_regs(const u32x4_t (&R)[S::PER], uint16_t *stage, int wv, int lane) {
#pragma unroll
    for (int t = 0; t < S::PER; ++t)
        *reinterpret_cast<u32x4_t *>(stage + (wv * S::PER + t) * X3_FRAG + lane * 8) = R[t];
}

__global__ __launch_bounds__(256) void x3r_kernel(X3Args a) {
    typedef X3Shape<2, 2> S;
    constexpr int NST = 2;
    __shared__ __attribute__((aligned(16))) uint16_t stg[NST * S::STAGE];
    __shared__ Seg sseg[X3_MAXSEG];

    const int tid = threadIdx.x;
    const int total = a.ntiles * a.n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / a.n_ntiles, nt = w % a.n_ntiles;
    const int2 tl = a.tiles[tile_id];
    const Task task = a.tasks[tl.x];
    const int M = task.m;
    const int m0 = tl.y * S::BM, n0 = nt * S::BN;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int nfm = (M + 15) / 16;
    const int mtile16 = m0 / 16;

    const int nseg = task.seg_end - task.seg_begin;   // <= X3_MAXSEG (checked by the plan)
    for (int i = tid; i < nseg; i += S::THREADS) sseg[i] = a.segs[task.seg_begin + i];
    __syncthreads();
    int nch = 0;
    for (int si = 0; si < nseg; ++si) nch += (sseg[si].k + 31) / 32;

    x3_f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = x3_f4{0.f, 0.f, 0.f, 0.f};

    X3Loader<S> ld;
    ld.seg = -1;
    ld.left = 0;
    u32x4_t R[S::PER];
    if (nch > 0) {
        x3_load_regs<S>(ld, a, sseg, R, wv, lane, mtile16, nfm, n0);
        x3_store_regs<S>(R, stg, wv, lane);
    }
    if (nch > 1) x3_load_regs<S>(ld, a, sseg, R, wv, lane, mtile16, nfm, n0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        x3_compute<4>(acc, stg + (c % NST) * S::STAGE, S::FA, lane, wr, wc, mtile16, nfm);
        if (c + 1 < nch) {
            x3_store_regs<S>(R, stg + ((c + 1) % NST) * S::STAGE, wv, lane);
            if (c + 2 < nch) x3_load_regs<S>(ld, a, sseg, R, wv, lane, mtile16, nfm, n0);
        }
        __syncthreads();
    }
    x3_epilogue<4, 4>(acc, stg, a, task, lane, wv, wr, wc, m0, n0);
}

// Accumulator rows of row groups past the block (4 per wave from row group
// mtile16 + 4 wr) set to zero.
template <int NJ>
__device__ __forceinline__ void x3_clear_rows(x3_f4 (&acc)[4][NJ], int wr, int mtile16, int nfm) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (mtile16 + wr * 4 + i >= nfm)
#pragma unroll
            for (int j = 0; j < NJ; ++j) acc[i][j] = x3_f4{0.f, 0.f, 0.f, 0.f};
}

// One output tile (rows m0.., columns n0.. of task): the task's segment
// list staged in LDS, then the NST-stage pipeline and the epilogue (DEV:
// device-scope skeleton panel stores).  on_seg as in x3_issue.  Ends with the
// epilogue's stores issued; the caller synchronises before reusing stg or
// sseg.
template <typename S, int NST, int NJ, bool DEV = false, typename OnSeg>
__device__ __forceinline__ void x3_tile(const X3Args &a, const Task &task, int m0, int n0, uint16_t *stg,
                                        Seg *sseg, OnSeg &&on_seg) {
    static_assert(S::NW * 64 * 16 * (NJ > 4 ? 4 : NJ) * 4 <= NST * S::STAGE * 2,
                  "the epilogue staging must fit in the stages");
    constexpr int WN = S::BN / (16 * NJ);
    const int tid = threadIdx.x;
    const int M = task.m;
    // the wave index in a scalar register: every per-wave quantity derived
    // from it (DMA destinations, fragment offsets) stays scalar
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wv / WN, wc = wv % WN;
    const int nfm = (M + 15) / 16;
    const int mtile16 = m0 / 16;

    const int nseg = task.seg_end - task.seg_begin;   // <= X3_MAXSEG (checked by the plan)
    for (int i = tid; i < nseg; i += S::THREADS) sseg[i] = a.segs[task.seg_begin + i];
    __syncthreads();
    int nch = 0;   // 32-deep chunks of the task
    for (int si = 0; si < nseg; ++si) nch += (sseg[si].k + 31) / 32;

    x3_f4 acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = x3_f4{0.f, 0.f, 0.f, 0.f};

    X3Loader<S> ld;
    ld.seg = -1;
    ld.left = 0;
    int issued = 0;
    auto issue = [&]() {
        if (issued < nch) {
            x3_issue<S>(ld, a, sseg, stg + (issued % NST) * S::STAGE, wv, lane, mtile16, nfm, n0, on_seg);
            ++issued;
        }
    };
#pragma unroll
    for (int p = 0; p < NST - 1; ++p) issue();
    for (int c = 0; c < nch; ++c) {
        // chunk c has landed (this wave's part, then everyone's), and every
        // wave is done with the stage chunk c-1 used, which chunk c+NST-1 refills
        wait_dma<S::PER, NST - 1>(issued - c - 1);
        wg_barrier();
        issue();
        x3_compute<NJ, false>(acc, stg + (c % NST) * S::STAGE, S::FA, lane, wr, wc, mtile16, nfm);
    }
    x3_clear_rows<NJ>(acc, wr, mtile16, nfm);
    __syncthreads();
    x3_epilogue<4, NJ, DEV>(acc, stg, a, task, lane, wv, wr, wc, m0, n0);
}

// The pipeline depth NST and the wave grid are template parameters; the
// autotuner picks among the instantiated variants (launch_x3).
template <int WM, int WN, int NST, int NJ = 4>
__global__ __launch_bounds__(64 * WM * WN) void x3_kernel(X3Args a) {
    typedef X3Shape<WM, WN, NJ> S;
    __shared__ __attribute__((aligned(16))) uint16_t stg[NST * S::STAGE];
    // the task's segment list, read from LDS inside the pipeline (a vector
    // load there would wait for every outstanding LDS-DMA load)
    __shared__ Seg sseg[X3_MAXSEG];

    const int total = a.ntiles * a.n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / a.n_ntiles, nt = w % a.n_ntiles;
    const int2 tl = a.tiles[tile_id];
    const Task task = a.tasks[tl.x];
    x3_tile<S, NST, NJ>(a, task, tl.y * S::BM, nt * S::BN, stg, sseg, [](int) {});
}

// ---- split-K variant for latency-bound launches -----------------------------
//
// Tiles of 128 x 32.  The task's 32-deep K chunks are dealt to 8 waves (chunk
// c to wave c % 8).  Each wave loads its chunk's 20 fragments (8 row groups
// and 2 column groups, hi and lo) straight into registers, one chunk ahead,
// and accumulates the whole 128 x 32 tile.  The eight partial tiles are
// summed through LDS in wave order (so results do not depend on timing), and
// two waves run the epilogue.  A task of K <= 256 then waits for one round
// of loads instead of a pipeline of K/32 dependent chunks: the top of the
// tree and small workloads are chains of such tasks.
constexpr int X3SK_WAVES = 8;

struct X3SkFrag {
    u32x4_t a[16], b[4];   // row group i: a[2i] hi, a[2i+1] lo; column group j: b[2j] hi, b[2j+1] lo
};

__device__ __forceinline__ void x3sk_load(X3SkFrag &F, const X3Args &a, const Seg &sg, int ki, int lane, int mtile16,
                                          int nfm, int n0) {
    const int nfk = (sg.k + 31) / 32;
    const int64_t ao = sg.a_off + (int64_t)ki * X3_FRAG + lane * 8;
    const uint16_t *ah = a.ahi[sg.a_buf] + ao;
    const uint16_t *al = a.alo[sg.a_buf] + ao;
#pragma unroll
    for (int rg = 0; rg < 8; ++rg) {
        const int64_t o = (int64_t)min(mtile16 + rg, nfm - 1) * nfk * X3_FRAG;
        F.a[2 * rg] = *reinterpret_cast<const u32x4_t *>(ah + o);
        F.a[2 * rg + 1] = *reinterpret_cast<const u32x4_t *>(al + o);
    }
    const int64_t pb = (sg.b_row + (int64_t)ki * 32) * a.nrhs + 8 * (lane >> 4);
#pragma unroll
    for (int cg = 0; cg < 2; ++cg) {
        const int64_t o = pb + (int64_t)min(n0 + cg * 16 + (lane & 15), a.nrhs - 1) * 32;
        F.b[2 * cg] = *reinterpret_cast<const u32x4_t *>(a.phi[sg.b_buf] + o);
        F.b[2 * cg + 1] = *reinterpret_cast<const u32x4_t *>(a.plo[sg.b_buf] + o);
    }
}

__device__ __forceinline__ void x3sk_compute(x3_f4 (&acc)[8][2], const X3SkFrag &F, int mtile16, int nfm) {
    const u32x4_t z = u32x4_t{0, 0, 0, 0};
    bf16x8_t bh[2], bl[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        bh[j] = __builtin_bit_cast(bf16x8_t, F.b[2 * j]);
        bl[j] = __builtin_bit_cast(bf16x8_t, F.b[2 * j + 1]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const bool ok = mtile16 + i < nfm;
        const bf16x8_t ah = __builtin_bit_cast(bf16x8_t, ok ? F.a[2 * i] : z);
        const bf16x8_t al = __builtin_bit_cast(bf16x8_t, ok ? F.a[2 * i + 1] : z);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = mfma_bf16(al, bh[j], acc[i][j]);
            acc[i][j] = mfma_bf16(ah, bl[j], acc[i][j]);
            acc[i][j] = mfma_bf16(ah, bh[j], acc[i][j]);
        }
    }
}

// One 128 x 32 tile of task (rows m0.., columns n0..) by the 8 waves of the
// workgroup; part: 8 x 64 x 64 floats of LDS, sseg: the segment list.
// on_seg(si) runs in a wave before it loads from segment si, for every
// segment its chunk cursor enters or passes over (the sweep's late wait);
// DEV as in x3_tile.  AHEAD: load the wave's next chunk while computing
// one (two register sets); without it the other resident wave of the SIMD
// covers the loads.  The caller synchronises before reusing the LDS.
template <bool DEV, bool AHEAD, typename OnSeg>
__device__ __forceinline__ void x3sk_tile(const X3Args &a, const Task &task, int m0, int n0, float *part, Seg *sseg,
                                          OnSeg &&on_seg) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nfm = (task.m + 15) / 16, mtile16 = m0 / 16;
    const int nseg = task.seg_end - task.seg_begin;   // <= X3_MAXSEG (checked by the plan)
    for (int i = tid; i < nseg; i += 64 * X3SK_WAVES) sseg[i] = a.segs[task.seg_begin + i];
    __syncthreads();
    int nch = 0;
    for (int si = 0; si < nseg; ++si) nch += (sseg[si].k + 31) / 32;

    x3_f4 acc[8][2];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = x3_f4{0.f, 0.f, 0.f, 0.f};
    // segment cursor: chunk c lies in segment si, whose first chunk is cbase
    int si = 0, cbase = 0;
    bool entered = false;
    auto load = [&](X3SkFrag &F, int c) {
        if (!entered) {
            on_seg(0);
            entered = true;
        }
        while (c >= cbase + (sseg[si].k + 31) / 32) {
            cbase += (sseg[si].k + 31) / 32;
            ++si;
            on_seg(si);
        }
        x3sk_load(F, a, sseg[si], c - cbase, lane, mtile16, nfm, n0);
    };
    if constexpr (AHEAD) {
        X3SkFrag F0, F1;
        if (wv < nch) load(F0, wv);
        for (int c = wv; c < nch; c += 2 * X3SK_WAVES) {
            if (c + X3SK_WAVES < nch) load(F1, c + X3SK_WAVES);
            x3sk_compute(acc, F0, mtile16, nfm);
            if (c + 2 * X3SK_WAVES < nch) load(F0, c + 2 * X3SK_WAVES);
            if (c + X3SK_WAVES < nch) x3sk_compute(acc, F1, mtile16, nfm);
        }
    } else {
        for (int c = wv; c < nch; c += X3SK_WAVES) {
            X3SkFrag F;
            load(F, c);
            x3sk_compute(acc, F, mtile16, nfm);
        }
    }
    float *mine = part + wv * 64 * 64;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) mine[((i * 2 + j) * 4 + r) * 64 + lane] = acc[i][j][r];
    __syncthreads();
    x3_f4 fin[4][2];
    if (wv < 2) {   // wave 0 sums row groups 0-3, wave 1 row groups 4-7, over the waves in order
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = (((wv * 4 + ii) * 2 + j) * 4 + r) * 64 + lane;
                    float v = 0.f;
#pragma unroll
                    for (int q = 0; q < X3SK_WAVES; ++q) v += part[q * 64 * 64 + e];
                    fin[ii][j][r] = v;
                }
    }
    __syncthreads();   // the partials are read; the epilogue stages in the same LDS
    if (wv < 2) x3_epilogue<4, 2, DEV>(fin, reinterpret_cast<uint16_t *>(part), a, task, lane, wv, wv, 0, m0, n0);
}

__global__ __launch_bounds__(64 * X3SK_WAVES) void x3sk_kernel(X3Args a) {
    // per wave its partial tile, element-major (8 x 2 fragments x 4 values,
    // then lane), so that a wave's stores and loads hit consecutive banks
    __shared__ __attribute__((aligned(16))) float part[X3SK_WAVES * 64 * 64];
    __shared__ Seg sseg[X3_MAXSEG];
    const int total = a.ntiles * a.n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / a.n_ntiles, nt = w % a.n_ntiles;
    const int2 tl = a.tiles[tile_id];
    const Task task = a.tasks[tl.x];
    x3sk_tile<false, true>(a, task, tl.y * X3_BM, nt * 32, part, sseg, [](int) {});
}

int x3_variant_bm(int variant) {
    return variant == X3_V256x128_W8S3 || variant == X3_V256x256_W8S2 ? 2 * X3_BM : X3_BM;
}

int x3_variant_bn(int variant) {
    switch (variant) {
    case X3_V128x256_S3: return 256;
    case X3_V256x256_W8S2: return 256;
    case X3_V128x64_S2: return 64;
    case X3_V128x32_S2: return 32;
    case X3_V128x32_S6: return 32;
    case X3_V128x64_S5: return 64;
    case X3_V128x32_K8: return 32;
    default: return 128;
    }
}

void launch_x3(const X3Args &a, int variant, hipStream_t st) {
    int64_t total = (int64_t)a.ntiles * a.n_ntiles;
    if (total <= 0) return;
    const dim3 grid((unsigned)total);
    switch (variant) {
    case X3_V128x128_S4:
        hipLaunchKernelGGL((x3_kernel<2, 2, 4>), grid, dim3(256), 0, st, a);
        break;
    case X3_V128x256_S3:
        hipLaunchKernelGGL((x3_kernel<2, 4, 3>), grid, dim3(512), 0, st, a);
        break;
    case X3_V128x128_REG:
        hipLaunchKernelGGL(x3r_kernel, grid, dim3(256), 0, st, a);
        break;
    case X3_V128x64_S2:
        hipLaunchKernelGGL((x3_kernel<2, 2, 2, 2>), grid, dim3(256), 0, st, a);
        break;
    case X3_V128x32_S2:
        hipLaunchKernelGGL((x3_kernel<2, 2, 2, 1>), grid, dim3(256), 0, st, a);
        break;
    case X3_V128x128_W8S3:
        hipLaunchKernelGGL((x3_kernel<2, 4, 3, 2>), grid, dim3(512), 0, st, a);
        break;
    case X3_V128x128_W8S4:
        hipLaunchKernelGGL((x3_kernel<2, 4, 4, 2>), grid, dim3(512), 0, st, a);
        break;
    case X3_V256x128_W8S3:
        hipLaunchKernelGGL((x3_kernel<4, 2, 3, 4>), grid, dim3(512), 0, st, a);
        break;
    case X3_V128x32_S6:
        hipLaunchKernelGGL((x3_kernel<2, 2, 6, 1>), grid, dim3(256), 0, st, a);
        break;
    case X3_V128x64_S5:
        hipLaunchKernelGGL((x3_kernel<2, 2, 5, 2>), grid, dim3(256), 0, st, a);
        break;
    case X3_V128x32_K8:
        hipLaunchKernelGGL(x3sk_kernel, grid, dim3(64 * X3SK_WAVES), 0, st, a);
        break;
    case X3_V256x256_W8S2:
        hipLaunchKernelGGL((x3_kernel<4, 2, 2, 8>), grid, dim3(512), 0, st, a);
        break;
    default:
        hipLaunchKernelGGL((x3_kernel<2, 2, 2>), grid, dim3(256), 0, st, a);
        break;
    }
}

// ---- dependency-driven sweep (protocol in mx_bf16x3.h) ----------------------

__device__ __forceinline__ int sweep_load(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int sweep_cas(int32_t *p, int expect, int want) {
    __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expect;   // the value found
}

// XCD of the executing workgroup (hardware register; only the choice of
// queue, and so L2 locality, depends on it)
__device__ __forceinline__ int sweep_home() {
    int x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return (x & 15) % SWEEP_NQ;
}

// Wave-wide: wait until every counter of waits [b, e) at column tile ct has
// reached its target.  The loads of the block that follow are issued after
// the counters are seen (no cache invalidation is needed: see the coherence
// note in mx_bf16x3.h).  Gives up on an error or after SWEEP_TIMEOUT_TICKS
// (setting the error word).
// cnt: the counters of the task's plan (ncols column tiles each), w its waits.
__device__ __forceinline__ void sweep_wait(const X3SweepArgs &sa, const int32_t *cnt, const SweepWait *wl, int ncols,
                                           int b, int e, int ct, int lane) {
    if (b >= e) return;
    for (int base = b; base < e; base += 64) {
        const int i = base + lane;
        int idx = 0, target = 0;
        if (i < e) {
            const SweepWait w = wl[i];
            idx = w.block * ncols + ct;
            target = w.target;
        }
        uint64_t t0 = 0;
        for (int it = 0;; ++it) {
            const bool ok = i >= e || sweep_load(cnt + idx) >= target;
            if (__ballot(!ok) == 0) break;
            if (sweep_load(sa.ctl + SWEEP_ERR)) break;
            const uint64_t now = __builtin_amdgcn_s_memrealtime();
            if (it == 0) {
                t0 = now;
            } else if (now - t0 > SWEEP_TIMEOUT_TICKS || it > SWEEP_MAX_POLLS) {
                if (lane == 0) __hip_atomic_store(sa.ctl + SWEEP_ERR, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");   // no memory access moves above the polls
}

// Claims (wave 0; lanes 0..7 each look after one queue).  A work item is
// claimed by fetch-and-add on a queue head: the home queue's, or once that
// queue is exhausted, the head of the queue whose next item ranks lowest.
// Such an item may still have a producer nobody has claimed.  Before it
// runs, its holder therefore checks that every producer is claimed (the
// heads have passed the producers' positions) and, while one is not,
// claims the untaken item of lowest rank over all queues, if that ranks
// below the held item (compare-and-swap on that queue's head; its
// producers rank lower, so all are claimed), and runs that first.  A held
// item thus ranks above everything its holder runs before it, so waits only
// ever point to lower ranks and the lowest unfinished item waits for
// nothing.  `heads` (lanes 0..7)
// caches the last heads read: they only grow, so a producer seen claimed
// stays claimed.  Error word: 1 = a wait timed out, 2 = claiming timed out.

// position of the item after `h` in queue z is in range and its rank
__device__ __forceinline__ int sweep_rank(const X3SweepArgs &sa, int2 qi, int h) {
    return h < qi.y * sa.ncols ? sa.q[qi.x + h / sa.ncols].g * sa.ncols + h % sa.ncols : INT_MAX;
}

// lanes 0..7: the queue (z) of the lowest rank among lanes 0..7, and that rank
__device__ __forceinline__ void sweep_argmin(int &rank, int &bz) {
#pragma unroll
    for (int o = SWEEP_NQ / 2; o >= 1; o >>= 1) {
        const int r2 = __shfl_xor(rank, o), z2 = __shfl_xor(bz, o);
        if (r2 < rank || (r2 == rank && z2 < bz)) {
            rank = r2;
            bz = z2;
        }
    }
    rank = __shfl(rank, 0);
    bz = __shfl(bz, 0);
}

__device__ __forceinline__ bool sweep_timed_out(const X3SweepArgs &sa, uint64_t t0, int it, int lane, int code) {
    if (sweep_load(sa.ctl + SWEEP_ERR)) return true;
    if (__builtin_amdgcn_s_memrealtime() - t0 > SWEEP_TIMEOUT_TICKS || it > SWEEP_MAX_POLLS) {
        if (lane == 0) __hip_atomic_store(sa.ctl + SWEEP_ERR, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    return false;
}

__device__ __forceinline__ int sweep_fetch_head(const X3SweepArgs &sa, int q) {
    return __hip_atomic_fetch_add(sa.ctl + q * SWEEP_HEAD_STRIDE, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Any next item: q = -1 when every queue is exhausted (or on an error).
// pre (lane 0): the result of a fetch-and-add on the home queue's head that
// the caller has already issued, or -1.
__device__ __forceinline__ void sweep_claim(const X3SweepArgs &sa, int home, int lane, int pre, int &heads,
                                            int &q_out, int &pos_out) {
    const int z = lane & (SWEEP_NQ - 1);
    const int2 qi = sa.qinfo[z];
    const int len = qi.y * sa.ncols;
    int cand = home;
    pre = __shfl(pre, 0);
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0;; ++it) {
        if (sweep_timed_out(sa, t0, it, lane, 2)) {
            q_out = -1;
            return;
        }
        int got = 0;
        if (it == 0 && pre >= 0) got = pre;
        else if (lane == 0) got = sweep_fetch_head(sa, cand);
        got = __shfl(got, 0);
        if (got < __shfl(len, cand)) {
            q_out = cand;
            pos_out = got;
            return;
        }
        // that queue is exhausted: the queue whose next item ranks lowest
        heads = lane < SWEEP_NQ ? sweep_load(sa.ctl + z * SWEEP_HEAD_STRIDE) : 0;
        int rank = lane < SWEEP_NQ ? sweep_rank(sa, qi, heads) : INT_MAX, bz = z;
        sweep_argmin(rank, bz);
        if (rank == INT_MAX) {
            q_out = -1;
            return;
        }
        cand = bz;
    }
}

// Are all producers of item (q, pos) claimed?  Reads the heads again only
// when the cached ones do not show it.
__device__ __forceinline__ bool sweep_producers_claimed(const X3SweepArgs &sa, int q, int pos, int lane, int &heads) {
    const int tix = pos / sa.ncols, ct = pos % sa.ncols;
    const int need = lane < SWEEP_NQ ? sa.q[sa.qinfo[q].x + tix].maxtile[lane] : -1;
    if (__ballot(!(need < 0 || heads > need * sa.ncols + ct)) == 0) return true;
    if (lane < SWEEP_NQ) heads = sweep_load(sa.ctl + lane * SWEEP_HEAD_STRIDE);
    return __ballot(!(need < 0 || heads > need * sa.ncols + ct)) == 0;
}

// The untaken item of lowest rank, if it ranks below `limit` (the held
// item's rank): compare-and-swap on its queue's head.  q = -1 when there is
// none (then every item below the held one, its producers included, is
// claimed) or on an error.  The bound matters: the held item's unclaimed
// producer seen a moment ago may have been claimed since, and the lowest
// untaken item may then depend on the held item itself.
__device__ __forceinline__ void sweep_claim_lowest(const X3SweepArgs &sa, int lane, int limit, int &heads, int &q_out,
                                                   int &pos_out) {
    const int z = lane & (SWEEP_NQ - 1);
    const int2 qi = sa.qinfo[z];
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0;; ++it) {
        if (sweep_timed_out(sa, t0, it, lane, 2)) {
            q_out = -1;
            return;
        }
        heads = lane < SWEEP_NQ ? sweep_load(sa.ctl + z * SWEEP_HEAD_STRIDE) : 0;
        int rank = lane < SWEEP_NQ ? sweep_rank(sa, qi, heads) : INT_MAX, bz = z;
        sweep_argmin(rank, bz);
        if (rank >= limit) {
            q_out = -1;
            return;
        }
        const int hb = __shfl(heads, bz);
        int got = 0;
        if (lane == 0) got = sweep_cas(sa.ctl + bz * SWEEP_HEAD_STRIDE, hb, hb + 1);
        if (__shfl(got, 0) == hb) {
            q_out = bz;
            pos_out = hb;
            return;
        }
    }
}

// Persistent workgroups of 4 waves (2 x 2, tiles of 128 x 16 NJ x 2
// columns, 2 stages): claim an item, wait for its early producers, run the
// tile with the late wait before its first u^ segment, signal the written
// block (every wave waits for its own panel stores, then adds 1), repeat.  Two
// workgroups per CU, as the level kernel: the accumulators then live in
// VGPRs.
// NJ = 0: the split-K tile (8 waves; its first SWEEP_WAVES waves signal,
// waves 0 and 1 after the epilogue's stores).
template <int NJ>
struct SweepTileShape {
    typedef X3Shape<2, 2, NJ> S;
    static constexpr int WAVES = 4;
    static constexpr int LDS16 = 2 * S::STAGE;   // uint16 of LDS for stages
};
template <>
struct SweepTileShape<0> {
    static constexpr int WAVES = X3SK_WAVES;
    static constexpr int LDS16 = X3SK_WAVES * 64 * 64 * 2;   // the partial tiles
};

template <int NJ, bool BATCH>
__global__ __launch_bounds__(64 * SweepTileShape<NJ>::WAVES) __attribute__((amdgpu_waves_per_eu(2))) void x3_sweep_kernel(
    X3SweepArgs sa) {
    typedef SweepTileShape<NJ> T;
    static_assert(T::WAVES >= SWEEP_WAVES, "the first SWEEP_WAVES waves signal once per tile");
    __shared__ __attribute__((aligned(16))) uint16_t stg[T::LDS16];
    __shared__ Seg sseg[X3_MAXSEG];
    // the item to run; wave 0's claim state between items (in LDS, so that
    // no register stays live across a tile): the claimed item waiting to
    // run, whether to claim a new one, the cached queue heads
    __shared__ int s_take[2], s_held[3], s_heads[SWEEP_NQ];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int home = sweep_home();
    // wave 0: pick the next item to run into s_take (pre as in sweep_claim)
    auto pick = [&](int pre) {
        int held_q = s_held[0], held_pos = s_held[1], heads = s_heads[lane & (SWEEP_NQ - 1)];
        if (s_held[2]) sweep_claim(sa, home, lane, pre, heads, held_q, held_pos);
        int q = held_q, pos = held_pos, fresh = 1;
        if (held_q >= 0 && !sweep_producers_claimed(sa, held_q, held_pos, lane, heads)) {
            // run the lowest untaken item first (if it ranks below the held one), keep the held one
            const int limit = sweep_rank(sa, sa.qinfo[held_q], held_pos);
            sweep_claim_lowest(sa, lane, limit, heads, q, pos);
            if (q >= 0) fresh = 0;
            else if (!sweep_load(sa.ctl + SWEEP_ERR)) q = held_q, pos = held_pos;   // its producers are claimed
        }
        if (lane < SWEEP_NQ) s_heads[lane] = heads;
        if (lane == 0) {
            s_take[0] = q;
            s_take[1] = pos;
            s_held[0] = held_q;
            s_held[1] = held_pos;
            s_held[2] = fresh;
        }
    };
    if (wv == 0) {
        if (lane == 0) {
            s_held[0] = -1;
            s_held[2] = 1;
        }
        if (lane < SWEEP_NQ) s_heads[lane] = 0;
        pick(-1);
    }
    __syncthreads();
    for (;;) {
        // uniform across the workgroup: scalar registers and loads from here
        const int q = __builtin_amdgcn_readfirstlane(s_take[0]), pos = __builtin_amdgcn_readfirstlane(s_take[1]);
        if (q < 0) break;
        const int tix = pos / sa.ncols;
        const SweepTile tile = sa.q[sa.qinfo[q].x + tix];
        int ct = pos % sa.ncols, rt = tile.rt, gcols = sa.ncols;
        const int32_t *cnt = sa.ctl + SWEEP_CNT;
        const SweepTask *stl = sa.st;
        const SweepWait *wl = sa.w;
        const Task *tasks = sa.x.tasks;
        if constexpr (BATCH) {
            const SweepGroup &G = sa.groups[tile.grp];
            ct = rt >> 16;
            rt &= 0xffff;
            gcols = G.ncols;
            cnt = sa.ctl + G.cbase;
            stl = G.st;
            wl = G.w;
            tasks = G.x.tasks;
        }
        const Task task = tasks[tile.task];
        const SweepTask stk = stl[tile.task];
        sweep_wait(sa, cnt, wl, gcols, stk.ew_b, stk.ew_e, ct, lane);
        auto late = [&](int si) {
            if (si == stk.late_seg) sweep_wait(sa, cnt, wl, gcols, stk.lw_b, stk.lw_e, ct, lane);
        };
        // the operands by reference, in kernel arguments or (batched) in the
        // group table: a private copy would be indexed through scratch
        const X3Args &xr = BATCH ? sa.groups[tile.grp].x : sa.x;
        if constexpr (NJ == 0) {
            x3sk_tile<true, false>(xr, task, rt * X3_BM, ct * 32, reinterpret_cast<float *>(stg), sseg, late);
        } else {
            typedef typename T::S S;
            x3_tile<S, 2, NJ, true>(xr, task, rt * S::BM, ct * S::BN, stg, sseg, late);
        }
        // wave 0 starts the next claim now, so that its round trip overlaps
        // the completion of the panel stores (every wave reads q, pos and
        // the task above, before this point)
        int pre = -1;
        const bool claim_next = wv == 0 && s_held[2];
        if (claim_next && lane == 0) pre = sweep_fetch_head(sa, home);
        if (stk.sig >= 0 && wv < SWEEP_WAVES) {
            // this wave's device-scope panel stores have completed before its count
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (lane == 0)
                __hip_atomic_fetch_add(const_cast<int32_t *>(cnt) + stk.sig * gcols + ct, 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
        if (wv == 0) pick(claim_next ? pre : -1);
        __syncthreads();   // s_take is ready; stages and segment list are free
    }
}

__global__ void sweep_reset_kernel(int32_t *ctl, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctl[i] = 0;
}

bool sweep_variant_ok(int variant) {
    return variant == X3_V128x128_S2 || variant == X3_V128x64_S2 || variant == X3_V128x32_S2 ||
           variant == X3_V128x32_K8;
}

int sweep_bn(int variant) { return x3_variant_bn(variant); }

template <int NJ>
static int sweep_grid_of() {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, x3_sweep_kernel<NJ, false>, 64 * SweepTileShape<NJ>::WAVES,
                                                     0) != hipSuccess)
        return 0;
    return std::max(1, per) * cus;
}

int sweep_max_grid(int variant) {
    switch (variant) {
    case X3_V128x64_S2: return sweep_grid_of<2>();
    case X3_V128x32_S2: return sweep_grid_of<1>();
    case X3_V128x32_K8: return sweep_grid_of<0>();
    default: return sweep_grid_of<4>();
    }
}

template <bool BATCH>
static void launch_sweep(const X3SweepArgs &a, int variant, int grid, int ctl_ints, hipStream_t st) {
    hipLaunchKernelGGL(sweep_reset_kernel, dim3((unsigned)std::min(64, (ctl_ints + 255) / 256)), dim3(256), 0, st,
                       a.ctl, ctl_ints);
    if (grid <= 0) return;
    switch (variant) {
    case X3_V128x64_S2:
        hipLaunchKernelGGL((x3_sweep_kernel<2, BATCH>), dim3((unsigned)grid), dim3(256), 0, st, a);
        break;
    case X3_V128x32_S2:
        hipLaunchKernelGGL((x3_sweep_kernel<1, BATCH>), dim3((unsigned)grid), dim3(256), 0, st, a);
        break;
    case X3_V128x32_K8:
        hipLaunchKernelGGL((x3_sweep_kernel<0, BATCH>), dim3((unsigned)grid), dim3(64 * X3SK_WAVES), 0, st, a);
        break;
    default:
        hipLaunchKernelGGL((x3_sweep_kernel<4, BATCH>), dim3((unsigned)grid), dim3(256), 0, st, a);
        break;
    }
}

void launch_x3_sweep(const X3SweepArgs &a, int variant, int grid, int ctl_ints, hipStream_t st) {
    launch_sweep<false>(a, variant, grid, ctl_ints, st);
}

void launch_x3_sweep_batch(const X3SweepArgs &a, int variant, int grid, int ctl_ints, hipStream_t st) {
    launch_sweep<true>(a, variant, grid, ctl_ints, st);
}

// ---- split-BF16x3 gradient products -------------------------------------
//
// K-fragment packing (layout in mx_grad.h).  Column blocks: one wave per
// fragment, consecutive waves on consecutive row groups of one chunk, so the
// 16-row runs a wave reads per column continue in its neighbours' loads.
template <typename T>
__global__ __launch_bounds__(256) void kpack_cols_kernel(const T *__restrict__ src, const KBlock *__restrict__ blocks,
                                                         int nrhs, uint16_t *__restrict__ hi,
                                                         uint16_t *__restrict__ lo) {
    const KBlock b = blocks[blockIdx.y];
    const int nfr = (b.m + 15) / 16, nfk = (nrhs + 31) / 32;
    const int64_t nfrag = (int64_t)nfr * nfk;
    const int lane = threadIdx.x & 63;
    const int64_t step = (int64_t)gridDim.x * 4;
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < nfrag; w += step) {
        const int g = (int)(w % nfr), c = (int)(w / nfr);
        const int i = g * 16 + (lane & 15);
        const int k0 = c * 32 + 8 * (lane >> 4);
        const T *p = src + b.row * (int64_t)nrhs + i;
        T v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (i < b.m && k0 + e < nrhs) ? p[(int64_t)(k0 + e) * b.m] : T(0);
        uint32_t h[4], l[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) split2(v[2 * e], v[2 * e + 1], h[e], l[e]);
        const int64_t o = ((b.fr + g) * nfk + c) * X3_FRAG + lane * 8;
        *reinterpret_cast<u32x4_t *>(hi + o) = u32x4_t{h[0], h[1], h[2], h[3]};
        *reinterpret_cast<u32x4_t *>(lo + o) = u32x4_t{l[0], l[1], l[2], l[3]};
    }
}

template <typename T>
void launch_kpack_cols(const T *src, const KBlock *blocks, int nblocks, int nrhs, uint16_t *hi, uint16_t *lo,
                       hipStream_t st) {
    const int gx = std::min(1024, ((nrhs + 31) / 32) * 64);   // 64 fragments per column chunk cover 1024 rows
    for (int b0 = 0; b0 < nblocks; b0 += 65535) {
        const dim3 grid((unsigned)gx, (unsigned)std::min(65535, nblocks - b0));
        hipLaunchKernelGGL(kpack_cols_kernel<T>, grid, dim3(256), 0, st, src, blocks + b0, nrhs, hi, lo);
    }
}
template void launch_kpack_cols<float>(const float *, const KBlock *, int, int, uint16_t *, uint16_t *, hipStream_t);
template void launch_kpack_cols<double>(const double *, const KBlock *, int, int, uint16_t *, uint16_t *,
                                        hipStream_t);

// Split panels: a 32-row panel block holds each column's 32 rows as 64
// contiguous bytes, so a 128-column tile is 8 KB per half.  A workgroup
// reads the tile (16-byte loads, coalesced), transposes it through LDS and
// writes its 2 row groups x 4 chunks of fragments per half (one chunk per
// wave).  Columns past nrhs and rows past the block are zero.
__global__ __launch_bounds__(256) void kpack_panels_kernel(const uint16_t *__restrict__ phi,
                                                           const uint16_t *__restrict__ plo,
                                                           const KBlock *__restrict__ blocks, int nrhs,
                                                           uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
    __shared__ __attribute__((aligned(16))) uint16_t t[2][128 * 32];   // [half][column][row]
    const KBlock b = blocks[blockIdx.y];
    const int nfr = (b.m + 15) / 16, nfk = (nrhs + 31) / 32;
    const int npb = (b.m + 31) / 32, nct = (nrhs + 127) / 128;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int u = blockIdx.x; u < npb * nct; u += gridDim.x) {
        const int pb = u % npb, ct = u / npb;
        const int j0 = ct * 128;
        const int64_t base = (b.row + (int64_t)pb * 32) * nrhs + (int64_t)j0 * 32;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int piece = r * 256 + tid;   // 16 bytes: rows rr..rr+7 of column q
            const int q = piece >> 2, rr = (piece & 3) * 8;
            u32x4_t vh = u32x4_t{0, 0, 0, 0}, vl = u32x4_t{0, 0, 0, 0};
            if (j0 + q < nrhs) {
                vh = *reinterpret_cast<const u32x4_t *>(phi + base + q * 32 + rr);
                vl = *reinterpret_cast<const u32x4_t *>(plo + base + q * 32 + rr);
            }
            *reinterpret_cast<u32x4_t *>(&t[0][q * 32 + rr]) = vh;
            *reinterpret_cast<u32x4_t *>(&t[1][q * 32 + rr]) = vl;
        }
        __syncthreads();
        const int c = ct * 4 + wv;
        if (c < nfk) {
#pragma unroll
            for (int gg = 0; gg < 2; ++gg) {
                const int g = 2 * pb + gg;
                if (g >= nfr) break;
                const int r = gg * 16 + (lane & 15);
                const bool rok = pb * 32 + r < b.m;
                const int q0 = wv * 32 + 8 * (lane >> 4);
                const int64_t o = ((b.fr + g) * nfk + c) * X3_FRAG + lane * 8;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    uint32_t w[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const uint32_t a0 = rok ? t[h][(q0 + 2 * e) * 32 + r] : 0u;
                        const uint32_t a1 = rok ? t[h][(q0 + 2 * e + 1) * 32 + r] : 0u;
                        w[e] = a0 | (a1 << 16);
                    }
                    *reinterpret_cast<u32x4_t *>((h ? lo : hi) + o) = u32x4_t{w[0], w[1], w[2], w[3]};
                }
            }
        }
        __syncthreads();
    }
}

void launch_kpack_panels(const uint16_t *phi, const uint16_t *plo, const KBlock *blocks, int nblocks, int nrhs,
                         uint16_t *hi, uint16_t *lo, hipStream_t st) {
    const int gx = std::min(256, 8 * ((nrhs + 127) / 128));   // 8 panel blocks (256 rows) per column tile
    for (int b0 = 0; b0 < nblocks; b0 += 65535) {
        const dim3 grid((unsigned)gx, (unsigned)std::min(65535, nblocks - b0));
        hipLaunchKernelGGL(kpack_panels_kernel, grid, dim3(256), 0, st, phi, plo, blocks + b0, nrhs, hi, lo);
    }
}

// Gradient products dC = sum_s X_s Z_s^T on the evaluation kernel's pipeline
// (128 x 128 tiles, 4 waves of 64 x 64, two LDS stages, LDS-DMA loads): A
// fragments are X's K-fragments (row groups of the tile), B fragments Z's
// (row groups of the tile's columns), each advancing one fragment per
// 32-deep chunk.  Row groups past X are zeroed at the MFMA (as in the
// evaluation); column groups past Z load a valid neighbour and their results
// are not stored.  The tile leaves through LDS as 64 consecutive rows per
// column.
template <typename S>
struct KLoader {
    const uint16_t *p[S::PER];
    int seg, left;
};

template <typename S, typename O>
__device__ __forceinline__ void kg_seg_begin(KLoader<S> &L, const GradX3Args<O> &a, const KSeg &sg, int wv, int lane,
                                             int mtile16, int nfm, int ntile16, int nfn, int nfk) {
#pragma unroll
    for (int t = 0; t < S::PER; ++t) {
        const int f = wv * S::PER + t;   // wave-uniform
        if (f < S::FA) {
            const int64_t g = sg.x_fr + min(mtile16 + (f >> 1), nfm - 1);
            L.p[t] = ((f & 1) ? a.klo[sg.x_buf] : a.khi[sg.x_buf]) + g * nfk * X3_FRAG + lane * 8;
        } else {
            const int q = f - S::FA;
            const int64_t g = sg.z_fr + min(ntile16 + (q >> 1), nfn - 1);
            L.p[t] = ((q & 1) ? a.klo[sg.z_buf] : a.khi[sg.z_buf]) + g * nfk * X3_FRAG + lane * 8;
        }
    }
    L.left = nfk;
}

template <typename S, typename O>
__device__ __forceinline__ void kg_issue(KLoader<S> &L, const GradX3Args<O> &a, const KSeg *sseg, uint16_t *stage,
                                         int wv, int lane, int mtile16, int nfm, int ntile16, int nfn, int nfk) {
    if (L.left == 0) {
        ++L.seg;
        kg_seg_begin<S, O>(L, a, sseg[L.seg], wv, lane, mtile16, nfm, ntile16, nfn, nfk);
    }
#pragma unroll
    for (int t = 0; t < S::PER; ++t) {
        lds_dma16(L.p[t], stage + (wv * S::PER + t) * X3_FRAG);
        L.p[t] += X3_FRAG;
    }
    --L.left;
}

// WM x WN waves of 64 x 16 NJ: 2 x 2 x 4 (128 x 128 tiles, two workgroups
// per CU) or 4 x 2 x 8 (256 x 256 tiles, one workgroup per CU, the
// pipelined 64 x 128 wave compute of the evaluation kernel, rows past X
// cleared after the chunks).
template <typename O, int WM, int WN, int NJ>
__global__ __launch_bounds__(64 * WM * WN) void x3_grad_kernel(GradX3Args<O> a) {
    typedef X3Shape<WM, WN, NJ> S;
    constexpr int NST = 2;
    constexpr int JG = NJ > 4 ? 4 : NJ;   // column groups staged at a time in the epilogue
    static_assert(S::BM == S::BN && (S::BM == GRAD_X3_TILE || S::BM == GRAD_X3_WTILE), "tile size");
    static_assert(S::NW * 64 * 16 * JG * 4 <= NST * S::STAGE * 2, "the epilogue staging must fit in the stages");
    __shared__ __attribute__((aligned(16))) uint16_t stg[NST * S::STAGE];
    __shared__ KSeg sseg[GRAD_KMAXSEG];

    const int tid = threadIdx.x;
    const int4 tl = a.tiles[xcd_remap(blockIdx.x, a.ntiles)];
    const GTask task = a.tasks[tl.x];
    O *out = a.out[task.c_buf];
    if (out == nullptr) return;   // this block type's gradient was not requested (uniform)
    const int M = task.m, N = task.n;
    const int m0 = tl.y * S::BM, n0 = tl.z * S::BN;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wv / WN, wc = wv % WN;
    const int nfm = (M + 15) / 16, nfn = (N + 15) / 16;
    const int mtile16 = m0 / 16, ntile16 = n0 / 16;
    const int nfk = (a.nrhs + 31) / 32;

    const int nseg = task.seg_end - task.seg_begin;   // <= GRAD_KMAXSEG (checked by the plan)
    for (int i = tid; i < nseg; i += S::THREADS) sseg[i] = a.segs[task.seg_begin + i];
    __syncthreads();
    const int nch = nseg * nfk;

    x3_f4 acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = x3_f4{0.f, 0.f, 0.f, 0.f};

    KLoader<S> ld;
    ld.seg = -1;
    ld.left = 0;
    int issued = 0;
    auto issue = [&]() {
        if (issued < nch) {
            kg_issue<S, O>(ld, a, sseg, stg + (issued % NST) * S::STAGE, wv, lane, mtile16, nfm, ntile16, nfn, nfk);
            ++issued;
        }
    };
#pragma unroll
    for (int p = 0; p < NST - 1; ++p) issue();
    for (int c = 0; c < nch; ++c) {
        wait_dma<S::PER, NST - 1>(issued - c - 1);
        wg_barrier();
        issue();
        x3_compute<NJ, NJ != 8>(acc, stg + (c % NST) * S::STAGE, S::FA, lane, wr, wc, mtile16, nfm);
    }
    if constexpr (NJ == 8) x3_clear_rows<NJ>(acc, wr, mtile16, nfm);
    __syncthreads();

    // C(row, col) of block (i, j), register q: row 16 i + 4 (lane / 16) + q,
    // column 16 j + lane % 16 of the wave's 64 x 16 NJ; staged 64 x 16 JG at a time
    float *Cs = reinterpret_cast<float *>(stg) + wv * 64 * 16 * JG;
    O *C = out + task.c_off;
#pragma unroll
    for (int g0 = 0; g0 < NJ; g0 += JG) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < JG; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int col = j * 16 + (lane & 15), row = i * 16 + 4 * (lane >> 4) + q;
                    Cs[col * 64 + (row ^ ((col & 15) << 2))] = acc[i][g0 + j][q];
                }
        __builtin_amdgcn_wave_barrier();
        const int rb = m0 + wr * 64, cb = n0 + wc * 16 * NJ + g0 * 16;
        for (int e = lane; e < 64 * 16 * JG; e += 64) {
            const int col = e >> 6, row = e & 63;
            const int gm = rb + row, gn = cb + col;
            if (gm < M && gn < N) C[(int64_t)gn * task.ldc + gm] = (O)Cs[col * 64 + (row ^ ((col & 15) << 2))];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <typename O>
void launch_grad_x3(const GradX3Args<O> &a, hipStream_t st) {
    if (a.ntiles <= 0) return;
    hipLaunchKernelGGL((x3_grad_kernel<O, 2, 2, 4>), dim3((unsigned)a.ntiles), dim3(256), 0, st, a);
}
template <typename O>
void launch_grad_x3w(const GradX3Args<O> &a, hipStream_t st) {
    if (a.ntiles <= 0) return;
    hipLaunchKernelGGL((x3_grad_kernel<O, 4, 2, 8>), dim3((unsigned)a.ntiles), dim3(512), 0, st, a);
}
template void launch_grad_x3<float>(const GradX3Args<float> &, hipStream_t);
template void launch_grad_x3<double>(const GradX3Args<double> &, hipStream_t);
template void launch_grad_x3w<float>(const GradX3Args<float> &, hipStream_t);
template void launch_grad_x3w<double>(const GradX3Args<double> &, hipStream_t);

}  // namespace mx
