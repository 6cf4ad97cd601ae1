// Kernel blocks computed on the device from the points: D_k = K(points of
// leaf x, points of leaf y), B_k = K(skeleton points of x, skeleton points of
// y), Gaussian kernel, written in FP64, FP32 or packed split-BF16.
#include "mx_internal.h"

namespace mx {

//
// D_k = K(points of leaf x, points of leaf y) and B_k = K(skeleton points of
// x, skeleton points of y) computed where they are needed instead of being
// uploaded, for matrices whose FP64 blocks do not fit the host (the
// reference computes them on the host, matroxTest-V1/compression.h:256-360,
// with the same norm expansion |a|^2 + |b|^2 - 2 a.b).



__global__ void point_norms(const double *__restrict__ X, int64_t n, int d, double *__restrict__ nrm) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < d; ++q) s += X[p * d + q] * X[p * d + q];
        nrm[p] = s;
    }
}

__device__ __forceinline__ double gauss_entry(const double *__restrict__ X, const double *__restrict__ nrm, int d,
                                              double kscal, int64_t pa, int64_t pb) {
    const double *xa = X + pa * d, *xb = X + pb * d;
    double dot = 0.0;
    for (int q = 0; q < d; ++q) dot += xa[q] * xb[q];
    return exp(kscal * (-2.0 * dot + nrm[pa] + nrm[pb]));
}

// blocks[blockIdx.y] in its stored layout (column-major na x nb), FP64 or FP32
template <typename T>
__global__ void gen_blocks(const double *__restrict__ X, const double *__restrict__ nrm, int d, double kscal,
                           const int32_t *__restrict__ idx, const GenBlock *__restrict__ blocks, T *__restrict__ out) {
    const GenBlock g = blocks[blockIdx.y];
    const int64_t ne = (int64_t)g.na * g.nb;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % g.na), j = (int)(e / g.na);
        out[g.out + e] = (T)gauss_entry(X, nrm, d, kscal, idx[g.a + i], idx[g.b + j]);
    }
}

// Packed split-BF16 operand xb[blockIdx.y] (fragment order as pack_x3_blocks)
// computed from its generating block gb[blockIdx.y]: op(A)(i, k) is entry
// (i, k) of the block, or (k, i) when transposed.
__global__ void gen_pack_x3(const double *__restrict__ X, const double *__restrict__ nrm, int d, double kscal,
                            const int32_t *__restrict__ idx, const X3Block *__restrict__ xb,
                            const GenBlock *__restrict__ gb, uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
    const X3Block bk = xb[blockIdx.y];
    const GenBlock g = gb[blockIdx.y];
    const int nfk = (bk.K + 31) / 32, nfm = (bk.M + 15) / 16;
    const int64_t lanes = (int64_t)nfm * nfk * 64;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < lanes;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frag = t >> 6;
        const int lane = (int)(t & 63);
        const int mi = (int)(frag / nfk), ki = (int)(frag % nfk);
        const int row = mi * 16 + (lane & 15);
        const int kb = ki * 32 + 8 * (lane >> 4);
        const int64_t dst = bk.dst_off + frag * 512 + (int64_t)lane * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = kb + e;
            double x = 0.0;
            if (row < bk.M && k < bk.K) {
                const int r = bk.trans ? k : row, c = bk.trans ? row : k;
                x = gauss_entry(X, nrm, d, kscal, idx[g.a + r], idx[g.b + c]);
            }
            const uint16_t h = bf16_rne((float)x);
            hi[dst + e] = h;
            lo[dst + e] = bf16_rne((float)(x - (double)bf16_f(h)));
        }
    }
}

GenState::~GenState() {
    for (void *p : {(void *)Xd, (void *)nrm, (void *)idx[0], (void *)idx[1]})
        if (p) (void)hipFree(p);
}

void GenState::init(const mx_structure &s, const Topology &t, const GenInfo &gen) {
    arg_check(gen.X && gen.n >= 1 && gen.d >= 1 && gen.h > 0 && gen.lids && gen.skels && gen.skoff,
              "bad block-generation arguments");
    int64_t N = 0;
    for (int l = 0; l < s.nleaf; ++l) N = std::max<int64_t>(N, (int64_t)s.wrow[l] + s.Ddim[l]);
    for (int64_t r = 0; r < N; ++r) arg_check(gen.lids[r] >= 0 && gen.lids[r] < gen.n, "lids out of range");
    for (int64_t i = 0; i < gen.nskels; ++i)
        arg_check(gen.skels[i] >= 0 && gen.skels[i] < gen.n, "skeleton point id out of range");
    for (int v = 0; v < s.numnodes; ++v)
        arg_check(s.slen[v] == 0 || (gen.skoff[v] >= 0 && gen.skoff[v] + s.slen[v] <= gen.nskels),
                  "skeleton offsets out of range");
    d = gen.d;
    kscal = -0.5 / (gen.h * gen.h);
    Xd = upload(gen.X, (size_t)gen.n * gen.d);
    HIPCHK(hipMalloc(&nrm, sizeof(double) * (size_t)gen.n));
    point_norms<<<cvt_grid(gen.n), 256>>>(Xd, gen.n, gen.d, nrm);
    HIPCHK(hipGetLastError());
    idx[0] = upload(gen.lids, (size_t)N);
    idx[1] = upload(gen.skels, (size_t)gen.nskels);
    for (int64_t k = 0; k < s.ncount; ++k) {
        int x = s.is_hss ? (int)k : s.npairx[k], y = s.is_hss ? (int)k : s.npairy[k];
        if (s.Ddim[x] == 0 || s.Ddim[y] == 0) continue;
        at[0][(int64_t)s.Dptr[k]] = (int)blk[0].size();
        blk[0].push_back(GenBlock{(int64_t)s.Dptr[k], s.wrow[x], s.wrow[y], s.Ddim[x], s.Ddim[y]});
    }
    for (int64_t k = 0; k < s.fcount; ++k) {
        int x, y;
        if (s.is_hss) {
            x = (int)k + 1;
            int p = t.parent[x];
            y = s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
        } else {
            x = s.fpairx[k];
            y = s.fpairy[k];
        }
        if (s.slen[x] == 0 || s.slen[y] == 0) continue;
        at[1][(int64_t)s.Bptr[k]] = (int)blk[1].size();
        blk[1].push_back(GenBlock{(int64_t)s.Bptr[k], gen.skoff[x], gen.skoff[y], s.slen[x], s.slen[y]});
    }
}

template <typename T>
static void emit_blocks(const GenState &G, int w, T *out) {
    const std::vector<GenBlock> &b = G.blk[w];
    for (size_t b0 = 0; b0 < b.size(); b0 += 65535) {
        const size_t nb = std::min<size_t>(65535, b.size() - b0);
        GenBlock *db = upload(b.data() + b0, nb);
        gen_blocks<T><<<dim3(64, (unsigned)nb), 256>>>(G.Xd, G.nrm, G.d, G.kscal, G.idx[w], db, out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(db);
        if (e != hipSuccess) throw MxError(MX_ERR_HIP, std::string("block generation: ") + hipGetErrorString(e));
    }
}

void GenState::emit(int w, double *out) const { emit_blocks(*this, w, out); }
void GenState::emit(int w, float *out) const { emit_blocks(*this, w, out); }

void GenState::emit_x3(int w, const std::vector<X3Block> &xb, uint16_t *hi, uint16_t *lo) const {
    std::vector<X3Block> todo;
    std::vector<GenBlock> src;
    for (const X3Block &x : xb) {
        if (x.M == 0 || x.K == 0) continue;
        auto it = at[w].find(x.src_off);
        if (it == at[w].end()) throw MxError(MX_ERR_ARG, "a packed operand does not start at a generated block");
        todo.push_back(x);
        src.push_back(blk[w][it->second]);
    }
    for (size_t b0 = 0; b0 < todo.size(); b0 += 65535) {
        const size_t nb = std::min<size_t>(65535, todo.size() - b0);
        X3Block *dx = upload(todo.data() + b0, nb);
        GenBlock *dg = nullptr;
        hipError_t e = hipMalloc(&dg, nb * sizeof(GenBlock));
        if (e == hipSuccess) e = hipMemcpy(dg, src.data() + b0, nb * sizeof(GenBlock), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            gen_pack_x3<<<dim3(64, (unsigned)nb), 256>>>(Xd, nrm, d, kscal, idx[w], dx, dg, hi, lo);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipDeviceSynchronize();
        }
        (void)hipFree(dx);
        if (dg) (void)hipFree(dg);
        if (e != hipSuccess) throw MxError(MX_ERR_HIP, std::string("block generation: ") + hipGetErrorString(e));
    }
}

}  // namespace mx
