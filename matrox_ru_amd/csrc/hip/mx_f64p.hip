// FP64 block products on packed A operands (layouts in mx_f64p.h).
#include "mx_f64p.h"
#include "mx_internal.h"

namespace mx {

typedef double f64x4_t __attribute__((ext_vector_type(4)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));

__global__ void pack_f64_blocks(const double *__restrict__ src, const X3Block *__restrict__ blocks,
                                double *__restrict__ dst) {
    const X3Block bk = blocks[blockIdx.y];
    const int nkg = (bk.K + 15) / 16, nrg = (bk.M + 15) / 16;
    const int64_t n = (int64_t)nrg * nkg * F64P_FRAG;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frag = t >> 8;
        const int pos = (int)(t & 255);
        const int rg = (int)(frag / nkg), kg = (int)(frag % nkg);
        const int l = (pos & 127) >> 1;   // the MFMA lane that reads this element
        const int row = rg * 16 + (l & 15);
        const int k = kg * 16 + 4 * (l >> 4) + 2 * (pos >> 7) + (pos & 1);
        double x = 0.0;
        if (row < bk.M && k < bk.K)
            x = bk.trans ? src[bk.src_off + (int64_t)row * bk.lda + k] : src[bk.src_off + (int64_t)k * bk.lda + row];
        dst[bk.dst_off + t] = x;
    }
}

constexpr int F64P_WAVES = 8, F64P_THREADS = 64 * F64P_WAVES;
// Tile columns: 4 waves of 16 NJ (NJ = 2: 128, NJ = 4: 256).  A stage holds
// the 8 A fragments (row groups of the tile), then 16 k of every tile
// column: 32 KB (NJ = 2) or 48 KB (NJ = 4) of doubles.
template <int NJ> struct F64PShape {
    static constexpr int BN = 64 * NJ;
    static constexpr int STAGE = 8 * F64P_FRAG + BN * F64P_BK;   // doubles
    static constexpr int PER = 2 + 4 * NJ;   // DMA loads per wave per chunk: one A fragment (2 x 1 KB), 8 B columns per 4 NJ loads of 256 B
};

__device__ __forceinline__ void dma16(const double *src, double *dst_wave) {
    __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)dst_wave, 16, 0, 0);
}
__device__ __forceinline__ void dma4(const uint32_t *src, double *dst_wave) {
    __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)dst_wave, 4, 0, 0);
}

__device__ __forceinline__ f64x4_t mfma_f64(double a, double b, f64x4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Half a chunk's operands: the 4 A row groups of the wave (16 bytes per lane
// at ab + 2048 i) and its NJ B column groups (at bb + 2048 j), read in one
// asm block that waits for its own reads (LDS reads the compiler can see
// would wait for every outstanding LDS-DMA load).
__device__ __forceinline__ void f64p_read(uint32_t ab, uint32_t bb, u32x4_t (&A)[4], u32x4_t (&B)[2]) {
    asm volatile(
        "ds_read_b128 %0, %6 offset:0\n\t"
        "ds_read_b128 %1, %6 offset:2048\n\t"
        "ds_read_b128 %2, %6 offset:4096\n\t"
        "ds_read_b128 %3, %6 offset:6144\n\t"
        "ds_read_b128 %4, %7 offset:0\n\t"
        "ds_read_b128 %5, %7 offset:2048\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(B[0]), "=&v"(B[1])
        : "v"(ab), "v"(bb)
        : "memory");
}
__device__ __forceinline__ void f64p_read(uint32_t ab, uint32_t bb, u32x4_t (&A)[4], u32x4_t (&B)[4]) {
    asm volatile(
        "ds_read_b128 %0, %8 offset:0\n\t"
        "ds_read_b128 %1, %8 offset:2048\n\t"
        "ds_read_b128 %2, %8 offset:4096\n\t"
        "ds_read_b128 %3, %8 offset:6144\n\t"
        "ds_read_b128 %4, %9 offset:0\n\t"
        "ds_read_b128 %5, %9 offset:2048\n\t"
        "ds_read_b128 %6, %9 offset:4096\n\t"
        "ds_read_b128 %7, %9 offset:6144\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(B[0]), "=&v"(B[1]), "=&v"(B[2]), "=&v"(B[3])
        : "v"(ab), "v"(bb)
        : "memory");
}

// The 16 NJ MFMAs of a chunk for a wave of 64 x 16 NJ (row groups 4 wr + i,
// tile columns 16 NJ wc + 16 j + l%16).  B column n keeps logical slot s
// (k = 2s, 2s+1) at physical slot s ^ ((n >> 1) & 7), the same for every j.
template <int NJ>
__device__ __forceinline__ void f64p_compute(f64x4_t (&acc)[4][NJ], uint32_t sbase, int lane, int wr, int wc) {
    const int n = 16 * NJ * wc + (lane & 15), g = lane >> 4, f = (n >> 1) & 7;
    const uint32_t ab = sbase + wr * 4 * (F64P_FRAG * 8) + lane * 16;
    const uint32_t bb = sbase + 8 * F64P_FRAG * 8 + n * (F64P_BK * 8);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        u32x4_t A[4], B[NJ];
        f64p_read(ab + h * 1024, bb + (((2 * g + h) ^ f) << 4), A, B);
        f64x2_t a[4], b[NJ];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = __builtin_bit_cast(f64x2_t, A[i]);
#pragma unroll
        for (int j = 0; j < NJ; ++j) b[j] = __builtin_bit_cast(f64x2_t, B[j]);
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = mfma_f64(a[i][e], b[j][e], acc[i][j]);
    }
}

// Chunk cursor of a wave: segment, 16-deep chunk within it, and the
// segment's operands (wave-uniform).
struct F64PLoader {
    const double *a;    // the wave's A fragment row of the segment (chunk 0)
    const double *bv;   // B block: column buffer + first row * nrhs
    int seg, kc, nk, K, b_ld;
};

// 128 x 64 NJ tile of a task by 8 waves (2 x 4 of 64 x 16 NJ).  NST stages
// of LDS; loads run NST-1 chunks ahead of the MFMAs.  NJ = 2 with 2 stages
// fits two workgroups per CU; the others run one.
template <int NST, int NJ>
__global__ __launch_bounds__(F64P_THREADS)
__attribute__((amdgpu_waves_per_eu(NST == 2 && NJ == 2 ? 4 : 2))) void f64p_kernel(F64PArgs a) {
    typedef F64PShape<NJ> S;
    static_assert(NST * S::STAGE >= F64P_WAVES * 64 * 16, "the epilogue staging must fit in the stages");
    __shared__ __attribute__((aligned(16))) double stg[NST * S::STAGE];
    // the task's segment list, read from LDS inside the pipeline (a global
    // load there would wait for every outstanding LDS-DMA load)
    __shared__ Seg sseg[F64P_MAXSEG];

    const int total = a.ntiles * a.n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / a.n_ntiles, nt = w % a.n_ntiles;
    const int2 tl = a.tiles[tile_id];
    const Task task = a.tasks[tl.x];
    const int M = task.m;
    if (M <= 0) return;   // padding tile (the whole workgroup)
    const int m0 = tl.y * F64P_BM, n0 = nt * S::BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wv >> 2, wc = wv & 3;
    const int nrg = (M + 15) / 16;
    const int arg = min(m0 / 16 + wv, nrg - 1);   // the wave's A row group (rows past M: any valid group)
    // B loads of the wave: tile columns 8 NJ wv + 2 j + lane/32 (j < 4 NJ),
    // the lane's 4 bytes at dword lane%4 of physical 16-byte slot
    // (lane/4)%8, which holds logical slot ((lane/4)%8) ^ (j%8)
    const int bcol = n0 + 8 * NJ * wv + (lane >> 5);
    const int bps = (lane >> 2) & 7, bkh = (lane >> 1) & 1, bhalf = lane & 1;

    const int nseg = task.seg_end - task.seg_begin;   // <= F64P_MAXSEG (checked by the plan)
    for (int i = tid; i < nseg; i += F64P_THREADS) sseg[i] = a.segs[task.seg_begin + i];
    __syncthreads();

    f64x4_t acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = f64x4_t{0.0, 0.0, 0.0, 0.0};

    F64PLoader L;
    L.seg = -1;
    L.kc = L.nk = 0;
    L.K = L.b_ld = 0;
    L.a = L.bv = nullptr;
    bool more = true;
    int issued = 0;
    auto issue = [&]() {
        if (!more) return;
        while (L.kc == L.nk) {
            if (++L.seg >= nseg) {
                more = false;
                return;
            }
            const Seg sg = sseg[L.seg];
            L.K = sg.k;
            L.nk = (sg.k + F64P_BK - 1) / F64P_BK;
            L.kc = 0;
            L.a = a.a[sg.a_buf] + sg.a_off + (int64_t)arg * sg.lda * F64P_FRAG;
            L.bv = a.vec[sg.b_buf] + sg.b_row * (int64_t)a.nrhs;
            L.b_ld = sg.b_ld;
        }
        double *stage = stg + (issued % NST) * S::STAGE;
        const double *ap = L.a + (int64_t)L.kc * F64P_FRAG + lane * 2;
        dma16(ap, stage + wv * F64P_FRAG);
        dma16(ap + 128, stage + wv * F64P_FRAG + 128);
        const int k0 = L.kc * F64P_BK;
        double *bst = stage + 8 * F64P_FRAG + 8 * NJ * wv * F64P_BK;
#pragma unroll
        for (int j = 0; j < 4 * NJ; ++j) {
            const int col = min(bcol + 2 * j, a.nrhs - 1);
            const int k = min(k0 + 2 * (bps ^ (j & 7)) + bkh, L.K - 1);
            const uint32_t *p = reinterpret_cast<const uint32_t *>(L.bv + ((int64_t)col * L.b_ld + k)) + bhalf;
            dma4(p, bst + j * 2 * F64P_BK);
        }
        ++L.kc;
        ++issued;
    };
#pragma unroll
    for (int p = 0; p < NST - 1; ++p) issue();
    for (int c = 0; c < issued; ++c) {
        // chunk c has landed (this wave's part, then everyone's), and every
        // wave is done with the stage chunk c-1 used, which chunk c+NST-1 refills
        wait_dma<S::PER, NST - 1>(issued - c - 1);
        wg_barrier();
        issue();
        f64p_compute<NJ>(acc, lds_addr(stg + (c % NST) * S::STAGE), lane, wr, wc);
    }
    __syncthreads();

    // epilogue: per column group, the wave's 64 x 16 block goes through its
    // own 8 KB of LDS (rows XOR-swizzled by column: the 32 lanes of a write
    // hit distinct banks) and out as 64 consecutive rows per column
    double *Cs = stg + wv * 64 * 16;
    double *C = a.vec[task.c_buf] + task.c_row * (int64_t)a.nrhs;
    const int rb = m0 + 64 * wr;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = lane & 15, row = 16 * i + (lane >> 4) + 4 * r;
                Cs[col * 64 + (row ^ (2 * col))] = acc[i][j][r];
            }
        __builtin_amdgcn_wave_barrier();
        const int cb = n0 + 16 * NJ * wc + 16 * j;
#pragma unroll 4
        for (int e = lane; e < 64 * 16; e += 64) {
            const int col = e >> 6, row = e & 63;
            const int gm = rb + row, gn = cb + col;
            if (gm < M && gn < a.nrhs) {
                const double v = Cs[col * 64 + (row ^ (2 * col))];
                double *y = C + (int64_t)gn * M + gm;
                *y = task.beta ? *y + v : v;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

void launch_f64p(const F64PArgs &a, int nst, int wide, hipStream_t st) {
    const int64_t total = (int64_t)a.ntiles * a.n_ntiles;
    if (total <= 0) return;
    if (total > INT32_MAX) throw MxError(MX_ERR_ARG, "launch too large");
    const dim3 grid((unsigned)total), block(F64P_THREADS);
    if (wide) {
        if (nst == 3) hipLaunchKernelGGL((f64p_kernel<3, 4>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((f64p_kernel<2, 4>), grid, block, 0, st, a);
    } else {
        if (nst == 3) hipLaunchKernelGGL((f64p_kernel<3, 2>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((f64p_kernel<2, 2>), grid, block, 0, st, a);
    }
}

}  // namespace mx

// ---- FP32 on packed A operands -------------------------------------------------

namespace mx {

typedef float f32x4_t __attribute__((ext_vector_type(4)));

__global__ void pack_f32_blocks(const float *__restrict__ src, const X3Block *__restrict__ blocks,
                                float *__restrict__ dst) {
    const X3Block bk = blocks[blockIdx.y];
    const int nkg = (bk.K + 15) / 16, nrg = (bk.M + 15) / 16;
    const int64_t n = (int64_t)nrg * nkg * F64P_FRAG;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frag = t >> 8;
        const int pos = (int)(t & 255);
        const int rg = (int)(frag / nkg), kg = (int)(frag % nkg);
        const int l = pos >> 2;   // the MFMA lane that reads this element
        const int row = rg * 16 + (l & 15);
        const int k = kg * 16 + 4 * (l >> 4) + (pos & 3);
        float x = 0.f;
        if (row < bk.M && k < bk.K)
            x = bk.trans ? src[bk.src_off + (int64_t)row * bk.lda + k] : src[bk.src_off + (int64_t)k * bk.lda + row];
        dst[bk.dst_off + t] = x;
    }
}

constexpr int F32P_WAVES = 8, F32P_THREADS = 64 * F32P_WAVES;
// Tile columns: 4 waves of 16 NJ.  A stage: the 8 A fragments, then 16 k
// of every tile column: 16 KB (NJ = 2) or 24 KB (NJ = 4) of floats.
template <int NJ> struct F32PShape {
    static constexpr int BN = 64 * NJ;
    static constexpr int STAGE = 8 * F64P_FRAG + BN * F64P_BK;   // floats
    static constexpr int PER = 1 + 2 * NJ;   // DMA loads per wave per chunk: one A fragment (1 KB), 16 NJ B columns (2 NJ x 256 B)
};

__device__ __forceinline__ void dma16f(const float *src, float *dst_wave) {
    __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)dst_wave, 16, 0, 0);
}
__device__ __forceinline__ void dma4f(const float *src, float *dst_wave) {
    __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)dst_wave, 4, 0, 0);
}

// A chunk's operands of a wave of 64 x 16 NJ: its 4 A row groups (16 bytes
// per lane at ab + 1024 i) and NJ B column groups (at bb + 1024 j), read in
// one asm block that waits for its own reads.
__device__ __forceinline__ void f32p_read(uint32_t ab, uint32_t bb, u32x4_t (&A)[4], u32x4_t (&B)[2]) {
    asm volatile(
        "ds_read_b128 %0, %6 offset:0\n\t"
        "ds_read_b128 %1, %6 offset:1024\n\t"
        "ds_read_b128 %2, %6 offset:2048\n\t"
        "ds_read_b128 %3, %6 offset:3072\n\t"
        "ds_read_b128 %4, %7 offset:0\n\t"
        "ds_read_b128 %5, %7 offset:1024\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(B[0]), "=&v"(B[1])
        : "v"(ab), "v"(bb)
        : "memory");
}
__device__ __forceinline__ void f32p_read(uint32_t ab, uint32_t bb, u32x4_t (&A)[4], u32x4_t (&B)[4]) {
    asm volatile(
        "ds_read_b128 %0, %8 offset:0\n\t"
        "ds_read_b128 %1, %8 offset:1024\n\t"
        "ds_read_b128 %2, %8 offset:2048\n\t"
        "ds_read_b128 %3, %8 offset:3072\n\t"
        "ds_read_b128 %4, %9 offset:0\n\t"
        "ds_read_b128 %5, %9 offset:1024\n\t"
        "ds_read_b128 %6, %9 offset:2048\n\t"
        "ds_read_b128 %7, %9 offset:3072\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(B[0]), "=&v"(B[1]), "=&v"(B[2]), "=&v"(B[3])
        : "v"(ab), "v"(bb)
        : "memory");
}

// The 16 NJ MFMAs of a chunk for a wave of 64 x 16 NJ (row groups 4 wr + i,
// tile columns 16 NJ wc + 16 j + l%16).  Step s takes k = 4 (l/16) + s of
// the chunk in A and B alike.  B column n keeps logical slot g (k =
// 4g..4g+3) at physical slot g ^ ((n >> 2) & 3), the same for every j.
template <int NJ>
__device__ __forceinline__ void f32p_compute(f32x4_t (&acc)[4][NJ], uint32_t sbase, int lane, int wr, int wc) {
    const int g = lane >> 4, n = 16 * NJ * wc + (lane & 15);
    const uint32_t ab = sbase + wr * 4 * 1024 + lane * 16;
    const uint32_t bb = sbase + 8 * 1024 + n * (F64P_BK * 4) + ((g ^ ((n >> 2) & 3)) << 4);
    u32x4_t A[4], B[NJ];
    f32p_read(ab, bb, A, B);
    f32x4_t a[4], b[NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = __builtin_bit_cast(f32x4_t, A[i]);
#pragma unroll
    for (int j = 0; j < NJ; ++j) b[j] = __builtin_bit_cast(f32x4_t, B[j]);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
}

// 128 x 64 NJ tile of a task by 8 waves (2 x 4 of 64 x 16 NJ).  NST stages
// of LDS; loads run NST-1 chunks ahead of the MFMAs.
template <int NST, int NJ>
__global__ __launch_bounds__(F32P_THREADS)
__attribute__((amdgpu_waves_per_eu(NST == 2 ? 4 : 2))) void f32p_kernel(F32PArgs a) {
    typedef F32PShape<NJ> S;
    static_assert(NST * S::STAGE >= F32P_WAVES * 64 * 16, "the epilogue staging must fit in the stages");
    __shared__ __attribute__((aligned(16))) float stg[NST * S::STAGE];
    __shared__ Seg sseg[F64P_MAXSEG];

    const int total = a.ntiles * a.n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / a.n_ntiles, nt = w % a.n_ntiles;
    const int2 tl = a.tiles[tile_id];
    const Task task = a.tasks[tl.x];
    const int M = task.m;
    if (M <= 0) return;   // padding tile (the whole workgroup)
    const int m0 = tl.y * F64P_BM, n0 = nt * S::BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wv >> 2, wc = wv & 3;
    const int nrg = (M + 15) / 16;
    const int arg = min(m0 / 16 + wv, nrg - 1);   // the wave's A row group (rows past M: any valid group)
    // B loads of the wave: instruction j moves tile column 8 NJ wv + 4 j +
    // l/16 (j < 2 NJ), the lane's 4 bytes at physical dword l%16 of the
    // column's 64 bytes, whose 16-byte slot (l%16)/4 holds logical slot
    // ((l%16)/4) ^ (j%4)
    const int bq = lane & 3, bps = (lane >> 2) & 3;

    const int nseg = task.seg_end - task.seg_begin;   // <= F64P_MAXSEG (checked by the plan)
    for (int i = tid; i < nseg; i += F32P_THREADS) sseg[i] = a.segs[task.seg_begin + i];
    __syncthreads();

    f32x4_t acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

    const float *la = nullptr, *lbv = nullptr;
    int lseg = -1, lkc = 0, lnk = 0, lK = 0, lbld = 0;
    bool more = true;
    int issued = 0;
    auto issue = [&]() {
        if (!more) return;
        while (lkc == lnk) {
            if (++lseg >= nseg) {
                more = false;
                return;
            }
            const Seg sg = sseg[lseg];
            lK = sg.k;
            lnk = (sg.k + F64P_BK - 1) / F64P_BK;
            lkc = 0;
            la = a.a[sg.a_buf] + sg.a_off + (int64_t)arg * sg.lda * F64P_FRAG;
            lbv = a.vec[sg.b_buf] + sg.b_row * (int64_t)a.nrhs;
            lbld = sg.b_ld;
        }
        float *stage = stg + (issued % NST) * S::STAGE;
        dma16f(la + (int64_t)lkc * F64P_FRAG + lane * 4, stage + wv * F64P_FRAG);
        const int k0 = lkc * F64P_BK;
        float *bst = stage + 8 * F64P_FRAG + 8 * NJ * wv * F64P_BK;
#pragma unroll
        for (int j = 0; j < 2 * NJ; ++j) {
            const int col = min(n0 + 8 * NJ * wv + 4 * j + (lane >> 4), a.nrhs - 1);
            const int k = min(k0 + 4 * (bps ^ (j & 3)) + bq, lK - 1);
            dma4f(lbv + ((int64_t)col * lbld + k), bst + j * 4 * F64P_BK);
        }
        ++lkc;
        ++issued;
    };
#pragma unroll
    for (int p = 0; p < NST - 1; ++p) issue();
    for (int c = 0; c < issued; ++c) {
        wait_dma<S::PER, NST - 1>(issued - c - 1);
        wg_barrier();
        issue();
        f32p_compute<NJ>(acc, lds_addr(stg + (c % NST) * S::STAGE), lane, wr, wc);
    }
    __syncthreads();

    // epilogue: per column group, the wave's 64 x 16 block through its own
    // 4 KB of LDS (rows XOR-swizzled by column) and out as 64 consecutive
    // rows per column
    float *Cs = stg + wv * 64 * 16;
    float *C = a.vec[task.c_buf] + task.c_row * (int64_t)a.nrhs;
    const int rb = m0 + 64 * wr;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = lane & 15, row = 16 * i + 4 * (lane >> 4) + r;
                Cs[col * 64 + (row ^ (4 * col))] = acc[i][j][r];
            }
        __builtin_amdgcn_wave_barrier();
        const int cb = n0 + 16 * NJ * wc + 16 * j;
#pragma unroll 4
        for (int e = lane; e < 64 * 16; e += 64) {
            const int col = e >> 6, row = e & 63;
            const int gm = rb + row, gn = cb + col;
            if (gm < M && gn < a.nrhs) {
                const float v = Cs[col * 64 + (row ^ (4 * col))];
                float *y = C + (int64_t)gn * M + gm;
                *y = task.beta ? *y + v : v;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

void launch_f32p(const F32PArgs &a, int nst, int wide, hipStream_t st) {
    const int64_t total = (int64_t)a.ntiles * a.n_ntiles;
    if (total <= 0) return;
    if (total > INT32_MAX) throw MxError(MX_ERR_ARG, "launch too large");
    const dim3 grid((unsigned)total), block(F32P_THREADS);
    if (wide) {
        if (nst == 4) hipLaunchKernelGGL((f32p_kernel<4, 4>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((f32p_kernel<2, 4>), grid, block, 0, st, a);
    } else {
        if (nst == 4) hipLaunchKernelGGL((f32p_kernel<4, 2>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((f32p_kernel<2, 2>), grid, block, 0, st, a);
    }
}

}  // namespace mx
