// Segmented grouped GEMM on MFMA (FP64 and FP32), the workhorse of every
// evaluation phase.
//
// One workgroup (256 threads = 4 wave64) computes a 64 x 64 tile of one output
// block C; each wave owns a 32 x 32 quarter as 2 x 2 MFMA 16x16x4 blocks
// (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32).  The K loop walks the
// task's segments in 16-deep chunks, staging the A chunk (64 x 16, padded,
// m fastest) and the B chunk (16 x 64, k fastest, padded) in LDS with a
// double buffer; the global loads of chunk i+1 (DEPTH 1) or of chunks i+1
// and i+2 (DEPTH 2, two register stages) are in flight while chunk i is
// multiplied.  DEPTH 2 serves launches whose chains of chunks are
// latency-bound (few, deep tasks); DEPTH 1 keeps more workgroups per CU.
// Ragged edges (block sizes that are not multiples of the tile) load zeros.
// The finished tile goes through LDS so that the global stores (and, for
// beta = 1, loads) of C are 64 consecutive rows per column.
//
// seg_gemm_wide_kernel is the FP32 form with 128 x 128 tiles (4 waves of
// 64 x 64, 4 x 4 MFMA blocks each): half the L2 traffic per flop of the
// 64 x 64 tile, for the wide launches.
//
// MFMA 16x16x4 operand layout (lane l of 64): A element (row l%16, k l/16),
// B element (k l/16, col l%16).  Result register r holds C(4*(l/16)+r, l%16)
// for the FP32 instruction but C(l/16+4*r, l%16) for the FP64 one (measured
// with a lane-layout probe on gfx950).
#pragma once
#include "mx_common.h"

namespace mx {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<double> {
    typedef d4_t type;
    // row of result register r within the 16x16 block, for lane group g = l/16
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct MfmaAcc<float> {
    typedef f4_t type;
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

__device__ __forceinline__ d4_t mfma16x16x4(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16x16x4(float a, float b, f4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int SG_BM = 64, SG_BN = 64, SG_BK = 16;
constexpr int SG_AP = SG_BM + 1;   // padded row stride of the A chunk in LDS
constexpr int SG_BP = SG_BK + 1;   // padded column stride of the B chunk in LDS
constexpr int SG_CP = SG_BM + 1;   // padded column stride of the C tile in LDS
constexpr int SG_THREADS = 256;

template <typename T>
struct SegCursor {
    int seg;   // current segment
    int k0;    // first k of the chunk within the segment
};

// Depth-1 staging: 4 A values and 4 B values per thread, loaded only where
// they lie inside the block (zeros elsewhere).
template <typename T>
struct Stage1 {
    T a[4];
    T b[4];
};

template <typename T>
__device__ __forceinline__ void load_chunk1(Stage1<T> &st, const Seg &sg, int k0, int m0, int M, int n0,
                                           int nrhs, const Bufs<T> &bufs, int tid) {
    const T *A = bufs.mat[sg.a_buf] + sg.a_off;
    const int K = sg.k;
    if (!sg.a_trans) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            int kk = e / SG_BM, mm = e % SG_BM;
            int gm = m0 + mm, gk = k0 + kk;
            st.a[r] = (gm < M && gk < K) ? A[(int64_t)gk * sg.lda + gm] : T(0);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            int mm = e / SG_BK, kk = e % SG_BK;
            int gm = m0 + mm, gk = k0 + kk;
            st.a[r] = (gm < M && gk < K) ? A[(int64_t)gm * sg.lda + gk] : T(0);
        }
    }
    const T *Bv = bufs.vec[sg.b_buf] + sg.b_row * (int64_t)nrhs;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int e = r * SG_THREADS + tid;
        int nn = e / SG_BK, kk = e % SG_BK;
        int gn = n0 + nn, gk = k0 + kk;
        st.b[r] = (gn < nrhs && gk < K) ? Bv[(int64_t)gn * sg.b_ld + gk] : T(0);
    }
}

template <typename T>
__device__ __forceinline__ void store_chunk1(const Stage1<T> &st, int a_trans, T *As, T *Bs, int tid) {
    if (!a_trans) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            As[(e / SG_BM) * SG_AP + (e % SG_BM)] = st.a[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            As[(e % SG_BK) * SG_AP + (e / SG_BK)] = st.a[r];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int e = r * SG_THREADS + tid;
        Bs[(e / SG_BK) * SG_BP + (e % SG_BK)] = st.b[r];
    }
}

// Per-thread register staging of one chunk: 4 A values and 4 B values, as
// loaded, plus a bit per value that says whether it lies inside the block
// (out-of-range elements are loaded from a clamped, valid address and zeroed
// when the chunk is stored to LDS, so the loads carry no branches and the
// compiler can wait for one chunk's loads while the next chunk's are still in
// flight).
template <typename T>
struct Stage {
    T a[4];
    T b[4];
    uint32_t in;   // bits 0-3: a[r] in range, bits 4-7: b[r]
};

template <typename T>
__device__ __forceinline__ void load_chunk(Stage<T> &st, const Seg &sg, int k0, int m0, int M, int n0,
                                           int nrhs, const Bufs<T> &bufs, int tid) {
    const T *A = bufs.mat[sg.a_buf] + sg.a_off;
    const int K = sg.k;
    uint32_t in = 0;
    if (!sg.a_trans) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            int kk = e / SG_BM, mm = e % SG_BM;
            int gm = m0 + mm, gk = k0 + kk;
            in |= (gm < M && gk < K ? 1u : 0u) << r;
            st.a[r] = A[(int64_t)min(gk, K - 1) * sg.lda + min(gm, M - 1)];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            int mm = e / SG_BK, kk = e % SG_BK;
            int gm = m0 + mm, gk = k0 + kk;
            in |= (gm < M && gk < K ? 1u : 0u) << r;
            st.a[r] = A[(int64_t)min(gm, M - 1) * sg.lda + min(gk, K - 1)];
        }
    }
    const T *Bv = bufs.vec[sg.b_buf] + sg.b_row * (int64_t)nrhs;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int e = r * SG_THREADS + tid;
        int nn = e / SG_BK, kk = e % SG_BK;
        int gn = n0 + nn, gk = k0 + kk;
        in |= (gn < nrhs && gk < K ? 1u : 0u) << (4 + r);
        st.b[r] = Bv[(int64_t)min(gn, nrhs - 1) * sg.b_ld + min(gk, K - 1)];
    }
    st.in = in;
}

template <typename T>
__device__ __forceinline__ void store_chunk(const Stage<T> &st, int a_trans, T *As, T *Bs, int tid) {
    if (!a_trans) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            As[(e / SG_BM) * SG_AP + (e % SG_BM)] = (st.in >> r) & 1u ? st.a[r] : T(0);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int e = r * SG_THREADS + tid;
            As[(e % SG_BK) * SG_AP + (e / SG_BK)] = (st.in >> r) & 1u ? st.a[r] : T(0);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int e = r * SG_THREADS + tid;
        Bs[(e / SG_BK) * SG_BP + (e % SG_BK)] = (st.in >> (4 + r)) & 1u ? st.b[r] : T(0);
    }
}

// Advance to the next 16-deep chunk of the task's segment sequence; returns
// false when the task is exhausted.  Segments with k == 0 are skipped.
__device__ __forceinline__ bool next_chunk(const Seg *segs, int seg_end, int &seg, int &k0) {
    k0 += SG_BK;
    while (seg < seg_end && k0 >= segs[seg].k) {
        ++seg;
        k0 = 0;
    }
    return seg < seg_end;
}

__device__ __forceinline__ bool first_chunk(const Seg *segs, int seg_begin, int seg_end, int &seg, int &k0) {
    seg = seg_begin;
    k0 = 0;
    while (seg < seg_end && segs[seg].k <= 0) ++seg;
    return seg < seg_end;
}

// DEPTH: chunks whose global loads are in flight while a chunk is multiplied
// (1: one register stage, more workgroups per CU; 2: two register stages,
// for latency-bound launches).  The plan picks it per launch.
template <typename T, int DEPTH>
__global__ __launch_bounds__(SG_THREADS) void seg_gemm_kernel(const Task *__restrict__ tasks,
                                                              const Seg *__restrict__ segs,
                                                              const int2 *__restrict__ tiles, int ntiles,
                                                              int nrhs, int n_ntiles, Bufs<T> bufs) {
    typedef typename MfmaAcc<T>::type acc_t;
    constexpr int A_ELEMS = SG_BK * SG_AP;
    constexpr int B_ELEMS = SG_BN * SG_BP;
    constexpr int STAGE_ELEMS = A_ELEMS + B_ELEMS;
    constexpr int SMEM_ELEMS = (2 * STAGE_ELEMS > SG_BN * SG_CP) ? 2 * STAGE_ELEMS : SG_BN * SG_CP;
    __shared__ T smem[SMEM_ELEMS];

    const int tid = threadIdx.x;
    const int total = ntiles * n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / n_ntiles, nt = w % n_ntiles;
    const int2 tl = tiles[tile_id];
    const Task task = tasks[tl.x];
    const int M = task.m;
    const int m0 = tl.y * SG_BM, n0 = nt * SG_BN;

    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;

    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // Chunk c is multiplied from LDS stage c % 2 while the loads of chunks
    // c+1 and c+2 are in flight in the two register stages st0 and st1; the
    // loop is unrolled by two so that each stage is a fixed set of registers.
    // Every step issues the same loads whether or not its chunk exists (a
    // missing chunk re-reads the task's first chunk and is never stored), so
    // the wait before storing chunk c+1 leaves chunk c+2's loads in flight
    // on every path.
    int seg, k0;
    bool have = first_chunk(segs, task.seg_begin, task.seg_end, seg, k0);
    if (have && DEPTH == 1) {
        Stage1<T> st;
        int cur = 0;
        {
            Seg sg = segs[seg];
            load_chunk1(st, sg, k0, m0, M, n0, nrhs, bufs, tid);
            store_chunk1(st, sg.a_trans, smem, smem + A_ELEMS, tid);
        }
        __syncthreads();
        while (have) {
            int nseg = seg, nk0 = k0;
            bool nhave = next_chunk(segs, task.seg_end, nseg, nk0);
            Seg nsg;
            if (nhave) {
                nsg = segs[nseg];
                load_chunk1(st, nsg, nk0, m0, M, n0, nrhs, bufs, tid);
            }
            const T *As = smem + cur * STAGE_ELEMS;
            const T *Bs = As + A_ELEMS;
#pragma unroll
            for (int kk = 0; kk < SG_BK; kk += 4) {
                T a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) a[i] = As[(kk + lk) * SG_AP + wr * 32 + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < 2; ++j) b[j] = Bs[(wc * 32 + j * 16 + lr) * SG_BP + kk + lk];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = mfma16x16x4(a[i], b[j], acc[i][j]);
            }
            if (nhave) {
                T *Asn = smem + (cur ^ 1) * STAGE_ELEMS;
                store_chunk1(st, nsg.a_trans, Asn, Asn + A_ELEMS, tid);
            }
            __syncthreads();
            cur ^= 1;
            seg = nseg;
            k0 = nk0;
            have = nhave;
        }
    } else if (have) {
        const int sf = seg, kf = k0;   // a chunk that always exists
        Stage<T> st0, st1;
        // cursors of the chunks c+1 and c+2 (valid when h1 / h2), their a_trans
        int s1 = seg, k1 = k0, s2 = 0, k2 = 0;
        bool h1, h2 = false;
        int t1 = 0, t2 = 0;
        auto load_or_dummy = [&](Stage<T> &st, bool h, int sc, int kc) {
            const Seg sg = segs[h ? sc : sf];
            load_chunk(st, sg, h ? kc : kf, m0, M, n0, nrhs, bufs, tid);
            return sg.a_trans;
        };
        {
            const Seg sg = segs[seg];
            load_chunk(st0, sg, k0, m0, M, n0, nrhs, bufs, tid);
            store_chunk(st0, sg.a_trans, smem, smem + A_ELEMS, tid);
        }
        h1 = next_chunk(segs, task.seg_end, s1, k1);
        t1 = load_or_dummy(st0, h1, s1, k1);
        s2 = s1;
        k2 = k1;
        h2 = h1 && next_chunk(segs, task.seg_end, s2, k2);
        t2 = load_or_dummy(st1, h2, s2, k2);
        __syncthreads();
        // one chunk: multiply LDS stage cur, move chunk c+1 (register stage
        // sp) to the other LDS stage, load chunk c+3 into sp
        auto step = [&](Stage<T> &sp, int cur) {
            const T *As = smem + cur * STAGE_ELEMS;
            const T *Bs = As + A_ELEMS;
#pragma unroll
            for (int kk = 0; kk < SG_BK; kk += 4) {
                T a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) a[i] = As[(kk + lk) * SG_AP + wr * 32 + i * 16 + lr];
#pragma unroll
                for (int j = 0; j < 2; ++j) b[j] = Bs[(wc * 32 + j * 16 + lr) * SG_BP + kk + lk];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = mfma16x16x4(a[i], b[j], acc[i][j]);
            }
            if (h1) {
                T *Asn = smem + (cur ^ 1) * STAGE_ELEMS;
                store_chunk(sp, t1, Asn, Asn + A_ELEMS, tid);
            }
            int s3 = s2, k3 = k2;
            const bool h3 = h2 && next_chunk(segs, task.seg_end, s3, k3);
            const int t3 = load_or_dummy(sp, h3, s3, k3);
            __syncthreads();
            have = h1;
            h1 = h2;
            s1 = s2;
            k1 = k2;
            t1 = t2;
            h2 = h3;
            s2 = s3;
            k2 = k3;
            t2 = t3;
        };
        while (have) {
            step(st0, 0);
            if (!have) break;
            step(st1, 1);
        }
    }
    // C tile -> LDS (column-major, padded) -> coalesced global update
    T *Cs = smem;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = wr * 32 + i * 16 + MfmaAcc<T>::row(lk, r);
                int col = wc * 32 + j * 16 + lr;
                Cs[col * SG_CP + row] = acc[i][j][r];
            }
    __syncthreads();
    T *C = bufs.vec[task.c_buf] + task.c_row * (int64_t)nrhs;
#pragma unroll 4
    for (int e = tid; e < SG_BM * SG_BN; e += SG_THREADS) {
        int col = e / SG_BM, row = e % SG_BM;
        int gm = m0 + row, gn = n0 + col;
        if (gm < M && gn < nrhs) {
            T v = Cs[col * SG_CP + row];
            T *p = C + (int64_t)gn * M + gm;
            *p = task.beta ? *p + v : v;
        }
    }
}

// ---- FP32, 128 x 128 tiles --------------------------------------------------

constexpr int SW_BM = 128, SW_BN = 128, SW_BK = 16;
constexpr int SW_AP = SW_BM + 16;   // A chunk [k][m]: 16 rows x 4 k of a fragment read hit distinct banks
constexpr int SW_BP = SW_BK + 4;    // B chunk [n][k]: likewise for 16 columns x 4 k
constexpr int SW_CP = SW_BM + 1;    // C tile [n][m] in the epilogue

struct StageW {
    float a[8];
    float b[8];
};

__device__ __forceinline__ void load_chunk_w(StageW &st, const Seg &sg, int k0, int m0, int M, int n0, int nrhs,
                                             const Bufs<float> &bufs, int tid) {
    const float *A = bufs.mat[sg.a_buf] + sg.a_off;
    const int K = sg.k;
    if (!sg.a_trans) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int e = r * SG_THREADS + tid;
            int kk = e / SW_BM, mm = e % SW_BM;
            int gm = m0 + mm, gk = k0 + kk;
            st.a[r] = (gm < M && gk < K) ? A[(int64_t)gk * sg.lda + gm] : 0.f;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            int e = r * SG_THREADS + tid;
            int mm = e / SW_BK, kk = e % SW_BK;
            int gm = m0 + mm, gk = k0 + kk;
            st.a[r] = (gm < M && gk < K) ? A[(int64_t)gm * sg.lda + gk] : 0.f;
        }
    }
    const float *Bv = bufs.vec[sg.b_buf] + sg.b_row * (int64_t)nrhs;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        int e = r * SG_THREADS + tid;
        int nn = e / SW_BK, kk = e % SW_BK;
        int gn = n0 + nn, gk = k0 + kk;
        st.b[r] = (gn < nrhs && gk < K) ? Bv[(int64_t)gn * sg.b_ld + gk] : 0.f;
    }
}

__device__ __forceinline__ void store_chunk_w(const StageW &st, int a_trans, float *As, float *Bs, int tid) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        int e = r * SG_THREADS + tid;
        if (!a_trans) As[(e / SW_BM) * SW_AP + (e % SW_BM)] = st.a[r];
        else As[(e % SW_BK) * SW_AP + (e / SW_BK)] = st.a[r];
        Bs[(e / SW_BK) * SW_BP + (e % SW_BK)] = st.b[r];
    }
}

// (a template, instantiated for float only, so that every translation unit
// that includes this header shares one definition)
template <typename T>
__global__ __launch_bounds__(SG_THREADS) void seg_gemm_wide_kernel(const Task *__restrict__ tasks,
                                                                   const Seg *__restrict__ segs,
                                                                   const int2 *__restrict__ tiles, int ntiles,
                                                                   int nrhs, int n_ntiles, Bufs<T> bufs) {
    static_assert(sizeof(T) == sizeof(float), "FP32 only");
    constexpr int A_ELEMS = SW_BK * SW_AP;
    constexpr int STAGE = A_ELEMS + SW_BN * SW_BP;
    constexpr int SMEM = (2 * STAGE > SW_BN * SW_CP) ? 2 * STAGE : SW_BN * SW_CP;
    __shared__ float smem[SMEM];

    const int tid = threadIdx.x;
    const int total = ntiles * n_ntiles;
    const int w = xcd_remap(blockIdx.x, total);
    const int tile_id = w / n_ntiles, nt = w % n_ntiles;
    const int2 tl = tiles[tile_id];
    const Task task = tasks[tl.x];
    const int M = task.m;
    const int m0 = tl.y * SW_BM, n0 = nt * SW_BN;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;

    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

    int seg, k0;
    bool have = first_chunk(segs, task.seg_begin, task.seg_end, seg, k0);
    StageW st;
    int cur = 0;
    if (have) {
        Seg sg = segs[seg];
        load_chunk_w(st, sg, k0, m0, M, n0, nrhs, bufs, tid);
        store_chunk_w(st, sg.a_trans, smem, smem + A_ELEMS, tid);
    }
    __syncthreads();
    while (have) {
        int nseg = seg, nk0 = k0;
        bool nhave = next_chunk(segs, task.seg_end, nseg, nk0);
        Seg nsg;
        if (nhave) {
            nsg = segs[nseg];
            load_chunk_w(st, nsg, nk0, m0, M, n0, nrhs, bufs, tid);
        }
        const float *As = smem + cur * STAGE;
        const float *Bs = As + A_ELEMS;
#pragma unroll
        for (int kk = 0; kk < SW_BK; kk += 4) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[(kk + lk) * SW_AP + wr * 64 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[(wc * 64 + j * 16 + lr) * SW_BP + kk + lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma16x16x4(a[i], b[j], acc[i][j]);
        }
        if (nhave) {
            float *Asn = smem + (cur ^ 1) * STAGE;
            store_chunk_w(st, nsg.a_trans, Asn, Asn + A_ELEMS, tid);
        }
        __syncthreads();
        cur ^= 1;
        seg = nseg;
        k0 = nk0;
        have = nhave;
    }

    float *Cs = smem;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = wr * 64 + i * 16 + MfmaAcc<float>::row(lk, r);
                int col = wc * 64 + j * 16 + lr;
                Cs[col * SW_CP + row] = acc[i][j][r];
            }
    __syncthreads();
    float *C = bufs.vec[task.c_buf] + task.c_row * (int64_t)nrhs;
#pragma unroll 4
    for (int e = tid; e < SW_BM * SW_BN; e += SG_THREADS) {
        int col = e / SW_BM, row = e % SW_BM;
        int gm = m0 + row, gn = n0 + col;
        if (gm < M && gn < nrhs) {
            float v = Cs[col * SW_CP + row];
            float *p = C + (int64_t)gn * M + gm;
            *p = task.beta ? *p + v : v;
        }
    }
}

}  // namespace mx
