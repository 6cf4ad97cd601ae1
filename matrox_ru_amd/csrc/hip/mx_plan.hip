// Plans: device uploads, the launch sequence of one evaluation (forward and
// adjoint), the backward pass, the autotuner and the C-ABI of
// include/matrox_hgemm.h (batches are in mx_batch.hip, task lists in
// mx_lists.hip, device-generated blocks in mx_generate.hip).
//
// Evaluation order (Y = K~ W, semantics of symGen/HGEMM_gen.h:117-238):
//   UP        one launch per tree level, leaves first:
//             w_v = VT_v W_leaf  |  w_v = VT_v[:, :s_l] w_l + VT_v[:, s_l:] w_r
//   couplings fused schedule: one launch per level, root's children first;
//             each task owns one node c and sums its far couplings and its
//             parent's transfer:
//             u_c = beta*u_c + sum_{far (c,y)} B_cy w_y + VT_p[:, part(c)]^T u_p
//             split schedule: one launch of every far coupling of every level,
//             then one transfer launch per level with beta = 1
//   NEAR+LEAF one launch; each task owns one leaf block of Y:
//             Y_x = beta*Y_x + sum_{near (x,y)} D_xy W_y + VT_x^T u_x
// The reference runs NEAR first, then UP, FAR and DOWN with beta = 1 updates
// (HGEMM_gen.h:125-236); fusing the transfers into the coupling and near
// tasks gives the same sums with each output block written once per launch.
#include "mx_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <tuple>

namespace mx {

thread_local std::string g_err;

static void upload_lists(const TaskLists &L, Task **t, Seg **s, int2 **tl) {
    *t = upload(L.tasks.data(), L.tasks.size());
    *s = upload(L.segs.data(), L.segs.size());
    *tl = upload(L.tiles.data(), L.tiles.size());
}

__global__ void cvt_f64_f32(const double *__restrict__ in, float *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (float)in[i];
}

__global__ void cvt_f32_f64(const float *__restrict__ in, double *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (double)in[i];
}

// Pack the split-BF16 operands of D, B and VT from FP64 sources.
static void pack_all(mx_plan *P, const double *const src[3], hipStream_t st) {
    for (int w = 0; w < 3; ++w) {
        if (!src[w]) continue;
        for (int64_t b0 = 0; b0 < P->n_xblk[w]; b0 += 65535) {
            int64_t nb = std::min<int64_t>(65535, P->n_xblk[w] - b0);
            pack_x3_blocks<<<dim3(64, (unsigned)nb), 256, 0, st>>>(src[w], P->d_xblk[w] + b0, P->xhi[w], P->xlo[w]);
            HIPCHK(hipGetLastError());
        }
    }
}

static void pack_f32_all(mx_plan *P, const float *const src[3], hipStream_t st) {
    for (int w = 0; w < 3; ++w) {
        if (!src[w] || !P->f32p[w]) continue;
        for (int64_t b0 = 0; b0 < P->n_fblk[w]; b0 += 65535) {
            int64_t nb = std::min<int64_t>(65535, P->n_fblk[w] - b0);
            pack_f32_blocks<<<dim3(32, (unsigned)nb), 256, 0, st>>>(src[w], P->d_fblk[w] + b0, P->f32p[w]);
            HIPCHK(hipGetLastError());
        }
    }
}

static void pack_f64_all(mx_plan *P, const double *const src[3], hipStream_t st) {
    for (int w = 0; w < 3; ++w) {
        if (!src[w] || !P->f64p[w]) continue;
        for (int64_t b0 = 0; b0 < P->n_fblk[w]; b0 += 65535) {
            int64_t nb = std::min<int64_t>(65535, P->n_fblk[w] - b0);
            pack_f64_blocks<<<dim3(32, (unsigned)nb), 256, 0, st>>>(src[w], P->d_fblk[w] + b0, P->f64p[w]);
            HIPCHK(hipGetLastError());
        }
    }
}

// The coarse schedule's composite operands op(A1) op(A2) in FP64: one 32 x
// 32 tile of composite blockIdx.y per workgroup (blockIdx.x: tile, column
// of tiles major), written column-major (ld = m) at its c_off of out.  A2
// is a stored block, or a composite of an earlier generation (in out, at
// a2_off, column-major with ld a2_lda).
__global__ void comp_gemm(const double *__restrict__ B, const double *__restrict__ VT,
                          const CompDesc *__restrict__ cd, double *out) {
    const CompDesc d = cd[blockIdx.y];
    const Composite &c = d.c;
    const int tm = (c.m + 31) / 32, tn = (c.n + 31) / 32;
    if ((int)blockIdx.x >= tm * tn) return;
    const int i0 = ((int)blockIdx.x % tm) * 32, j0 = ((int)blockIdx.x / tm) * 32;
    const double *A1 = (c.a1_buf == MAT_B ? B : VT) + c.a1_off;
    // a second factor that is a composite itself was formed by an earlier launch
    const double *A2 = (c.a2_buf == MAT_COMP ? out : c.a2_buf == MAT_B ? B : VT) + c.a2_off;
    __shared__ double sa[32][33], sb[32][33];   // sa[l][i] = op(A1)(i0+i, l0+l), sb[l][j] = op(A2)(l0+l, j0+j)
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int l0 = 0; l0 < c.kmid; l0 += 32) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = ty + 8 * q;
            const int i = i0 + tx, l = l0 + r;
            double a = 0.0, b = 0.0;
            if (i < c.m && l < c.kmid)
                a = c.a1_trans ? A1[(int64_t)i * c.a1_lda + l] : A1[(int64_t)l * c.a1_lda + i];
            const int j = j0 + tx;
            if (l < c.kmid && j < c.n)
                b = c.a2_trans ? A2[(int64_t)l * c.a2_lda + j] : A2[(int64_t)j * c.a2_lda + l];
            sa[r][tx] = a;
            sb[r][tx] = b;
        }
        __syncthreads();
#pragma unroll 8
        for (int l = 0; l < 32; ++l) {
            const double a = sa[l][tx];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = fma(a, sb[l][ty + 8 * q], acc[q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = i0 + tx, j = j0 + ty + 8 * q;
        if (i < c.m && j < c.n) out[d.c_off + (int64_t)j * c.m + i] = acc[q];
    }
}

// Form the coarse schedule's composites from FP64 B and VT (device arrays)
// and pack them into the split VT operands.
static void form_composites(mx_plan *P, const double *B, const double *VT, hipStream_t st) {
    if (P->n_comp == 0) return;
    // (a structure without far blocks has no composite over B)
    if (!VT) throw MxError(MX_ERR_STATE, "composite operands need the FP64 VT blocks");
    double *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, (size_t)std::max<int64_t>(P->comp_elems, 1) * sizeof(double)));
    try {
        // generation by generation (the descriptors are in generation order)
        for (size_t g = 0; g + 1 < P->comp_gen.size(); ++g)
            for (int64_t c0 = P->comp_gen[g]; c0 < P->comp_gen[g + 1]; c0 += 65535) {
                const int64_t nc = std::min<int64_t>(65535, P->comp_gen[g + 1] - c0);
                comp_gemm<<<dim3((unsigned)P->comp_tiles, (unsigned)nc), 256, 0, st>>>(B, VT, P->comp_desc + c0, tmp);
                HIPCHK(hipGetLastError());
            }
        for (int64_t c0 = 0; c0 < P->n_comp; c0 += 65535) {
            const int64_t nc = std::min<int64_t>(65535, P->n_comp - c0);
            pack_x3_blocks<<<dim3(64, (unsigned)nc), 256, 0, st>>>(tmp, P->comp_xblk + c0, P->xhi[MAT_VT],
                                                                    P->xlo[MAT_VT]);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipFree(tmp);
        throw;
    }
    HIPCHK(hipFree(tmp));
}

// First padded panel row of every W leaf block and skeleton block (BF16x3
// split panels), and the rows of each panel buffer.
static void panel_rows(const mx_structure &s, std::vector<int64_t> &wprow, std::vector<int64_t> &sprow,
                       int64_t *wp_rows, int64_t *sp_rows) {
    wprow.assign(s.nleaf, 0);
    sprow.assign(s.numnodes, 0);
    int64_t r = 0;
    for (int l = 0; l < s.nleaf; ++l) {
        wprow[l] = r;
        r += x3_pad_rows(s.Ddim[l]);
    }
    *wp_rows = r;
    r = 0;
    for (int v = 0; v < s.numnodes; ++v) {
        sprow[v] = r;
        r += x3_pad_rows(s.slen[v]);
    }
    *sp_rows = r;
}

mx_plan *create_plan(const mx_structure &s, uint32_t precisions, int beta_u, int beta_y, const GenInfo *gen) {
    arg_check(s.nleaf > 0 && s.numnodes >= 1, "empty structure");
    arg_check(s.numnodes == 2 * s.nleaf - 1, "numnodes must be 2*nleaf-1 for a binary tree");
    arg_check(s.lchildren && s.rchildren && s.lm && s.Ddim && s.slen && s.VTptr && s.wrow && s.skelrow,
              "null structure array");
    arg_check(s.ncount >= 0 && s.fcount >= 0, "negative pair count");
    arg_check(s.ncount == 0 || (s.Dptr && (s.D || gen)), "near blocks missing");
    arg_check(s.fcount == 0 || (s.Bptr && (s.B || gen)), "far blocks missing");
    arg_check(s.is_hss || s.ncount == 0 || (s.npairx && s.npairy), "near pair arrays missing");
    arg_check(s.is_hss || s.fcount == 0 || (s.fpairx && s.fpairy), "far pair arrays missing");
    arg_check((precisions & 7u) != 0 && (precisions & ~(7u | MX_PLAN_ADJOINT)) == 0, "bad precision mask");
    Topology t = topology(s);
    validate(s, t);

    std::unique_ptr<mx_plan> P(new mx_plan());
    HIPCHK(hipGetDevice(&P->device));
    P->nleaf = s.nleaf;
    P->numnodes = s.numnodes;
    P->is_hss = s.is_hss;
    P->wrow.assign(s.wrow, s.wrow + s.nleaf);
    P->Ddim.assign(s.Ddim, s.Ddim + s.nleaf);
    P->skelrow.assign(s.skelrow, s.skelrow + s.numnodes);
    P->slen.assign(s.slen, s.slen + s.numnodes);
    int64_t N = 0, SK = 0;
    for (int l = 0; l < s.nleaf; ++l) N = std::max(N, (int64_t)s.wrow[l] + s.Ddim[l]);
    for (int v = 0; v < s.numnodes; ++v) SK = std::max(SK, (int64_t)s.skelrow[v] + s.slen[v]);
    P->N = N;
    P->skel_rows = SK;
    P->D_len = s.D_len;
    P->B_len = s.B_len;
    P->VT_len = s.VT_len;
    P->precisions = precisions & 7u;
    P->adjoint = (precisions & MX_PLAN_ADJOINT) != 0;

    std::vector<int64_t> wrow64(s.wrow, s.wrow + s.nleaf), skel64(s.skelrow, s.skelrow + s.numnodes);
    const int ndir = P->adjoint ? 2 : 1;
    // FP64 and FP32 plans: every A operand of the FP64/FP32 programs is also
    // packed in 16 x 16 fragments for f64p_kernel / f32p_kernel; each
    // distinct (buffer, offset, orientation, shape) gets its packed offset
    // (in elements, the same for both precisions), shared by all programs
    std::vector<X3Block> fblk[3];
    std::map<std::tuple<int64_t, int, int, int, int>, int64_t> fseen[3];
    int64_t fnext[3] = {0, 0, 0};
    for (int dir = 0; dir < ndir; ++dir)
        for (int sc = 0; sc < NSCHED; ++sc) {
            if (sched_coarse(sc)) continue;   // BF16x3 only
            TaskLists L = build_lists(s, t, beta_u, beta_y, 0, dir, sc, wrow64.data(), skel64.data());
            compute_launch_flops(L);
            // The 128-row tiles (packed-operand kernels, FP32 128 x 128
            // kernel) of the big launches go in groups of tasks that share B
            // operands, with each group's segments aligned (as the BF16x3
            // lists): at the headline the FP64 evaluation fetches 15.6
            // instead of 19.0 GB and runs 1.5% faster.  The 64-row tiles
            // keep the tree order and the segments their natural order: the
            // 64 x 64 kernel ran 2.8% slower on grouped lists (N = 524,288).
            // Both lists hold the same tasks, each with the same segments.
            TaskLists Lg = build_lists(s, t, beta_u, beta_y, 0, dir, sc, wrow64.data(), skel64.data(), true);
            if (Lg.tasks.size() != L.tasks.size() || Lg.segs.size() != L.segs.size() ||
                Lg.launches.size() != L.launches.size())
                throw MxError(MX_ERR_STATE, "grouped task lists differ from the tree-order lists");
            pair_tiles(Lg);
            Program &pg = P->prog[dir][0][sc];
            upload_lists(L, &pg.tasks, &pg.segs, &pg.tiles);
            if (precisions & (MX_PLAN_FP64 | MX_PLAN_FP32)) {
                std::vector<Seg> sp(Lg.segs);
                for (const Task &tk : L.tasks)
                    for (int si = tk.seg_begin; si < tk.seg_end; ++si) {
                        Seg &sg = sp[si];
                        const int w = sg.a_buf;
                        auto key = std::make_tuple(sg.a_off, sg.a_trans, sg.lda, tk.m, sg.k);
                        auto it = fseen[w].find(key);
                        int64_t dst;
                        if (it == fseen[w].end()) {
                            dst = fnext[w];
                            fseen[w][key] = dst;
                            fblk[w].push_back(X3Block{sg.a_off, dst, tk.m, sg.k, sg.lda, sg.a_trans});
                            fnext[w] += f64p_packed_elems(tk.m, sg.k);
                        } else {
                            dst = it->second;
                        }
                        sg.a_off = dst;
                        sg.a_trans = 0;
                        sg.lda = (sg.k + F64P_BK - 1) / F64P_BK;
                    }
                pg.segs_p = upload(sp.data(), sp.size());
                pg.maxseg.assign(L.launches.size(), 0);
                for (size_t li = 0; li < L.launches.size(); ++li)
                    for (int ti = L.launches[li].tile_offset; ti < L.launches[li].tile_offset + L.launches[li].ntiles;
                         ++ti) {
                        const Task &tk = L.tasks[L.tiles[ti].x];
                        pg.maxseg[li] = std::max(pg.maxseg[li], tk.seg_end - tk.seg_begin);
                    }
            }
            pg.launches = L.launches;
            pg.tiles2 = upload(Lg.tiles2.data(), Lg.tiles2.size());
            pg.launches2 = Lg.launches2;
            pg.flops = L.flops;
            std::copy(L.last_write, L.last_write + 4, pg.last_write);
            pg.built = true;
        }
    std::vector<X3Block> xblk[3];   // blocks to pack per stored buffer (D, B, VT)
    std::vector<CompDesc> comps;     // the coarse schedule's composites and their packing list
    std::vector<X3Block> comp_xblk;
    std::vector<int> cgen;           // per composite: generation (1: both factors stored blocks)
    std::vector<int64_t> cdst;       // per composite: packed offset in the VT operands
    int64_t comp_elems = 0;
    if (precisions & MX_PLAN_BF16X3) {
        // split panels: padded row offsets of W leaf blocks and skeleton blocks
        panel_rows(s, P->wprow, P->sprow, &P->wp_rows, &P->sp_rows);
        std::vector<X3Rows> wr(s.nleaf);
        for (int l = 0; l < s.nleaf; ++l) wr[l] = X3Rows{s.wrow[l], P->wprow[l], s.Ddim[l], 0};
        P->d_wrows = upload(wr.data(), wr.size());
        std::vector<X3Rows> sr;
        for (int v = 0; v < s.numnodes; ++v)
            if (s.slen[v] > 0) sr.push_back(X3Rows{s.skelrow[v], P->sprow[v], s.slen[v], 0});
        P->d_srows = upload(sr.data(), sr.size());
        P->n_srows = (int)sr.size();
        // every A operand is read from packed split-BF16 fragments: assign each
        // distinct (buffer, offset, orientation, shape) its packed offset,
        // shared by the forward and adjoint lists
        std::map<std::tuple<int64_t, int, int, int, int>, int64_t> seen[3];
        int64_t next[3] = {0, 0, 0};
        // composites of the coarse schedule: each distinct one formed once
        // (a nested one after the composite it multiplies), packed into the
        // VT operands; keyed with the global id of a composite factor
        std::map<std::tuple<int, int64_t, int, int, int, int64_t, int, int, int, int, int>, int> cid;
        for (int dsc = 0; dsc < ndir * NSCHED; ++dsc) {
            const int dir = dsc / NSCHED, sc = dsc % NSCHED;
            // composites need FP64 B and VT, which a plan of generated blocks
            // may never hold
            if (sched_coarse(sc) && gen) continue;
            TaskLists X = build_lists(s, t, beta_u, beta_y, 1, dir, sc, P->wprow.data(), P->sprow.data());
            split_long_tasks(X, X3_MAXSEG);
            std::vector<int> local(X.comps.size(), -1);   // this program's composites -> global ids
            std::function<int(int)> resolve = [&](int li) -> int {
                if (local[li] >= 0) return local[li];
                Composite c = X.comps[li];
                int g = 1;
                if (c.a2_buf == MAT_COMP) {
                    const int inner = resolve((int)c.a2_off);
                    c.a2_off = inner;
                    g = cgen[inner] + 1;
                }
                auto key = std::make_tuple(c.a1_buf, c.a1_off, c.a1_trans, c.a1_lda, c.a2_buf, c.a2_off,
                                           c.a2_trans, c.a2_lda, c.m, c.kmid, c.n);
                auto it = cid.find(key);
                if (it != cid.end()) return local[li] = it->second;
                const int id = (int)comps.size();
                cid[key] = id;
                comps.push_back(CompDesc{c, comp_elems});
                comp_elems += (int64_t)c.m * c.n;
                cgen.push_back(g);
                cdst.push_back(-1);
                return local[li] = id;
            };
            for (size_t ti = 0; ti < X.tasks.size(); ++ti) {
                if (X.tasks[ti].seg_end - X.tasks[ti].seg_begin > X3_MAXSEG)
                    throw MxError(MX_ERR_ARG, "internal: a BF16x3 task longer than its segment window");
                for (int si = X.tasks[ti].seg_begin; si < X.tasks[ti].seg_end; ++si) {
                    Seg &sg = X.segs[si];
                    if (sg.a_buf == MAT_COMP) {
                        const int id = resolve((int)sg.a_off);
                        if (cdst[id] < 0) {
                            cdst[id] = next[MAT_VT];
                            next[MAT_VT] += x3_packed_elems(comps[id].c.m, comps[id].c.n);
                        }
                        sg.a_buf = MAT_VT;
                        sg.a_off = cdst[id];
                        sg.a_trans = 0;
                        sg.lda = (sg.k + 31) / 32;
                        continue;
                    }
                    const int w = sg.a_buf, m = X.tasks[ti].m;
                    auto key = std::make_tuple(sg.a_off, sg.a_trans, sg.lda, m, sg.k);
                    auto it = seen[w].find(key);
                    int64_t dst;
                    if (it == seen[w].end()) {
                        dst = next[w];
                        seen[w][key] = dst;
                        X3Block xb{};
                        xb.src_off = sg.a_off;
                        xb.dst_off = dst;
                        xb.M = m;
                        xb.K = sg.k;
                        xb.lda = sg.lda;
                        xb.trans = sg.a_trans;
                        xblk[w].push_back(xb);
                        next[w] += x3_packed_elems(m, sg.k);
                    } else {
                        dst = it->second;
                    }
                    sg.a_off = dst;
                    sg.a_trans = 0;
                    sg.lda = (sg.k + 31) / 32;
                }
            }
            if (sc == SCHED_FUSED) {
                SweepLists SW = build_sweep(X, s.numnodes, P->sprow);
                if (SW.ok) {
                    mx_plan::Sweep &sw = P->sweep[dir];
                    sw.tasks = upload(SW.tasks.data(), SW.tasks.size());
                    sw.segs = upload(SW.segs.data(), SW.segs.size());
                    sw.st = upload(SW.stasks.data(), SW.stasks.size());
                    sw.w = upload(SW.waits.data(), SW.waits.size());
                    sw.q = upload(SW.tiles.data(), SW.tiles.size());
                    sw.qinfo = upload(SW.qinfo.data(), SW.qinfo.size());
                    sw.qinfo_h = SW.qinfo;
                    sw.h_tiles = SW.tiles;
                    sw.h_st = SW.stasks;
                    sw.h_w = SW.waits;
                    sw.nblocks = SW.nblocks;
                    sw.ok = true;
                }
            }
            compute_launch_flops(X);
            pair_tiles(X);
            Program &pg = P->prog[dir][1][sc];
            upload_lists(X, &pg.tasks, &pg.segs, &pg.tiles);
            pg.launches = X.launches;
            pg.tiles2 = upload(X.tiles2.data(), X.tiles2.size());
            pg.launches2 = X.launches2;
            pg.flops = X.flops;
            std::copy(X.last_write, X.last_write + 4, pg.last_write);
            pg.built = true;
        }
        for (int w = 0; w < 3; ++w) P->x3_elems[w] = next[w];
    }
    if (P->adjoint) {
        GradLists G = build_grad_lists(s, t);
        P->g_tasks = upload(G.tasks.data(), G.tasks.size());
        P->g_segs = upload(G.segs.data(), G.segs.size());
        P->g_tiles = upload(G.tiles.data(), G.tiles.size());
        if (G.tiles.size() > (size_t)INT32_MAX) throw MxError(MX_ERR_ARG, "gradient launch too large");
        P->g_ntiles = (int)G.tiles.size();
        P->g_tiles_x3 = upload(G.tiles_x3.data(), G.tiles_x3.size());
        P->g_ntiles_x3 = (int)G.tiles_x3.size();
        P->g_tiles_x3w = upload(G.tiles_x3w.data(), G.tiles_x3w.size());
        P->g_ntiles_x3w = (int)G.tiles_x3w.size();
        // the 256 x 256 tiles where their padded work is within 15% of the
        // 128 x 128 tiles' (blocks of about 256 rows and columns)
        P->g_x3w_auto = G.area_x3w <= 1.15 * G.area_x3;
        P->g_flops = G.flops;
        if (precisions & MX_PLAN_BF16X3) {
            // K-fragment blocks: leaves (W, dY) and skeleton nodes, each
            // starting at its own 16-row group, in index order
            std::vector<KBlock> kl, ks, kp;
            std::map<int64_t, int64_t> leaf_fr, skel_fr;   // first row -> first group (non-empty blocks)
            int64_t fr = 0;
            for (int l = 0; l < s.nleaf; ++l) {
                if (s.Ddim[l] == 0) continue;
                kl.push_back(KBlock{s.wrow[l], fr, s.Ddim[l], 0});
                leaf_fr[s.wrow[l]] = fr;
                fr += (s.Ddim[l] + 15) / 16;
            }
            P->kfr_leaf = fr;
            fr = 0;
            for (int v = 0; v < s.numnodes; ++v) {
                if (s.slen[v] == 0) continue;
                ks.push_back(KBlock{s.skelrow[v], fr, s.slen[v], 0});
                kp.push_back(KBlock{P->sprow[v], fr, s.slen[v], 0});
                skel_fr[s.skelrow[v]] = fr;
                fr += (s.slen[v] + 15) / 16;
            }
            P->kfr_skel = fr;
            std::vector<KSeg> kseg(G.segs.size());
            for (size_t i = 0; i < G.segs.size(); ++i) {
                const GSeg &g = G.segs[i];
                auto group = [&](int buf, int64_t row) {
                    const auto &m = (buf == GB_W || buf == GB_DY) ? leaf_fr : skel_fr;
                    auto it = m.find(row);
                    if (it == m.end()) throw MxError(MX_ERR_ARG, "internal: gradient operand without a block");
                    return it->second;
                };
                kseg[i] = KSeg{group(g.x_buf, g.x_row), group(g.z_buf, g.z_row), g.x_buf, g.z_buf};
            }
            for (const GTask &t : G.tasks)
                if (t.seg_end - t.seg_begin > GRAD_KMAXSEG)
                    throw MxError(MX_ERR_ARG, "internal: a gradient task with too many segments");
            P->g_ksegs = upload(kseg.data(), kseg.size());
            P->kb_leaf = upload(kl.data(), kl.size());
            P->kb_skel = upload(ks.data(), ks.size());
            P->kb_panel = upload(kp.data(), kp.size());
            P->n_kb_leaf = (int)kl.size();
            P->n_kb_skel = (int)ks.size();
        }
    }
    // stored blocks: uploaded, or generated from the points straight into
    // each precision's representation (the FP64 blocks of a large matrix
    // need not exist anywhere for an FP32 or BF16x3 plan)
    std::unique_ptr<GenState> G;
    if (gen) {
        G.reset(new GenState());
        G->init(s, t, *gen);
    }
    const bool need64 = !gen || (precisions & MX_PLAN_FP64);
    if (need64) {
        P->D64 = upload(gen ? nullptr : s.D, (size_t)s.D_len);
        P->B64 = upload(gen ? nullptr : s.B, (size_t)s.B_len);
        if (gen) {
            G->emit(0, P->D64);
            G->emit(1, P->B64);
        }
    }
    P->VT64 = upload(s.VT, (size_t)s.VT_len);
    if (precisions & MX_PLAN_FP32) {
        P->VT32 = upload<float>(nullptr, (size_t)s.VT_len);
        cvt_f64_f32<<<cvt_grid(s.VT_len), 256>>>(P->VT64, P->VT32, s.VT_len);
        P->D32 = upload<float>(nullptr, (size_t)s.D_len);
        P->B32 = upload<float>(nullptr, (size_t)s.B_len);
        if (gen) {
            G->emit(0, P->D32);
            G->emit(1, P->B32);
        } else {
            cvt_f64_f32<<<cvt_grid(s.D_len), 256>>>(P->D64, P->D32, s.D_len);
            cvt_f64_f32<<<cvt_grid(s.B_len), 256>>>(P->B64, P->B32, s.B_len);
        }
    }
    if (precisions & MX_PLAN_BF16X3) {
        for (int w = 0; w < 3; ++w) {
            P->xhi[w] = upload<uint16_t>(nullptr, (size_t)P->x3_elems[w]);
            P->xlo[w] = upload<uint16_t>(nullptr, (size_t)P->x3_elems[w]);
            P->d_xblk[w] = upload(xblk[w].data(), xblk[w].size());
            P->n_xblk[w] = (int64_t)xblk[w].size();
        }
        if (gen) {
            const double *src[3] = {nullptr, nullptr, P->VT64};
            pack_all(P.get(), src, nullptr);
            G->emit_x3(0, xblk[0], P->xhi[0], P->xlo[0]);
            G->emit_x3(1, xblk[1], P->xhi[1], P->xlo[1]);
        } else {
            const double *src[3] = {P->D64, P->B64, P->VT64};
            pack_all(P.get(), src, nullptr);
            if (!comps.empty()) {
                // descriptors in generation order, a composite factor named by
                // its offset in the FP64 buffer; every composite is read by a
                // segment, so each has its packed place
                std::vector<int> order(comps.size());
                for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
                std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cgen[a] < cgen[b]; });
                std::vector<CompDesc> cd;
                for (int i : order) {
                    if (cdst[i] < 0) throw MxError(MX_ERR_STATE, "internal: a composite no segment reads");
                    CompDesc d = comps[i];
                    if (d.c.a2_buf == MAT_COMP) d.c.a2_off = comps[d.c.a2_off].c_off;
                    cd.push_back(d);
                    X3Block xb{};
                    xb.src_off = d.c_off;
                    xb.dst_off = cdst[i];
                    xb.M = d.c.m;
                    xb.K = d.c.n;
                    xb.lda = d.c.m;
                    xb.trans = 0;
                    comp_xblk.push_back(xb);
                }
                P->comp_gen.assign(1, 0);
                for (size_t k = 1; k <= order.size(); ++k)
                    if (k == order.size() || cgen[order[k]] != cgen[order[k - 1]]) P->comp_gen.push_back((int64_t)k);
                P->comp_desc = upload(cd.data(), cd.size());
                P->comp_xblk = upload(comp_xblk.data(), comp_xblk.size());
                P->n_comp = (int64_t)cd.size();
                P->comp_elems = comp_elems;
                for (const CompDesc &d : cd)   // 32 x 32 tiles of the largest composite
                    P->comp_tiles = std::max<int64_t>(P->comp_tiles, (int64_t)((d.c.m + 31) / 32) * ((d.c.n + 31) / 32));
                HIPCHK(hipDeviceSynchronize());
                form_composites(P.get(), P->B64, P->VT64, nullptr);
            }
        }
    }
    // The packed copies double an FP64 plan's block memory (more with the
    // adjoint's transposed views).  They are made only when they fit beside
    // what the plan already holds with room left for the work buffers and
    // the caller's W and Y (16 GiB + 10% of the device); otherwise every
    // launch runs the 64 x 64 kernel on the stored blocks.
    bool pack64 = (precisions & MX_PLAN_FP64) != 0;
    if (pack64) {
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        const double need = (double)(fnext[0] + fnext[1] + fnext[2]) * sizeof(double);
        pack64 = need + 16.0 * (1ull << 30) + 0.1 * (double)total_b <= (double)free_b;
    }
    if (pack64) {
        for (int w = 0; w < 3; ++w) {
            P->f64p[w] = upload<double>(nullptr, (size_t)std::max<int64_t>(fnext[w], 1));
            P->f64p_elems[w] = fnext[w];
        }
    }
    // the FP32 copies (half the bytes), by the same rule
    bool pack32 = (precisions & MX_PLAN_FP32) != 0;
    if (pack32) {
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        const double need = (double)(fnext[0] + fnext[1] + fnext[2]) * sizeof(float);
        pack32 = need + 16.0 * (1ull << 30) + 0.1 * (double)total_b <= (double)free_b;
    }
    if (pack32)
        for (int w = 0; w < 3; ++w) {
            P->f32p[w] = upload<float>(nullptr, (size_t)std::max<int64_t>(fnext[w], 1));
            P->f64p_elems[w] = fnext[w];
        }
    if (pack64 || pack32)
        for (int w = 0; w < 3; ++w) {
            P->d_fblk[w] = upload(fblk[w].data(), fblk[w].size());
            P->n_fblk[w] = (int64_t)fblk[w].size();
        }
    if (pack64) {
        const double *src[3] = {P->D64, P->B64, P->VT64};
        pack_f64_all(P.get(), src, nullptr);
    }
    if (pack32) {
        const float *src[3] = {P->D32, P->B32, P->VT32};
        pack_f32_all(P.get(), src, nullptr);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    if (!(precisions & MX_PLAN_FP64)) {   // FP64 blocks were only a source
        // (B and VT stay while composites are formed from them)
        for (double **p : {&P->D64, &P->B64, &P->VT64}) {
            if (*p && (P->n_comp == 0 || p == &P->D64)) {
                HIPCHK(hipFree(*p));
                *p = nullptr;
            }
        }
    }
    return P.release();
}

// Work buffers grow to the largest request and are never shrunk.  Every
// reallocation of a buffer that an evaluation reads or writes bumps
// buf_gen, so that graphs captured with the old address are dropped before
// the next replay (eval_graph), and forgets the skeletons the old buffer held.
void ensure_adj_work(mx_plan *P, int nrhs, size_t elem) {
    size_t sb = (size_t)std::max<int64_t>(P->skel_rows, 1) * nrhs * elem;
    if (sb > P->skel_adj_bytes) {
        if (P->wskel_adj) HIPCHK(hipFree(P->wskel_adj));
        if (P->uskel_adj) HIPCHK(hipFree(P->uskel_adj));
        P->wskel_adj = P->uskel_adj = nullptr;
        P->skel_adj_bytes = 0;
        ++P->buf_gen;
        HIPCHK(hipMalloc(&P->wskel_adj, sb));
        HIPCHK(hipMalloc(&P->uskel_adj, sb));
        P->skel_adj_bytes = sb;
    }
}

static void ensure_io_g(mx_plan *P, int nrhs) {
    size_t ib = (size_t)std::max<int64_t>(P->N, 1) * nrhs * sizeof(float);
    if (ib > P->io_g_bytes) {
        if (P->io_g) HIPCHK(hipFree(P->io_g));
        P->io_g = nullptr;
        HIPCHK(hipMalloc(&P->io_g, ib));
        P->io_g_bytes = ib;
    }
}

void ensure_work(mx_plan *P, int nrhs, size_t elem, bool need_io) {
    size_t sb = (size_t)std::max<int64_t>(P->skel_rows, 1) * nrhs * elem;
    if (sb > P->skel_bytes) {
        if (P->wskel) HIPCHK(hipFree(P->wskel));
        if (P->uskel) HIPCHK(hipFree(P->uskel));
        P->wskel = P->uskel = nullptr;
        P->skel_bytes = 0;
        ++P->buf_gen;
        if (P->last_prec != MX_BF16X3) {   // the FP64/FP32 forward skeletons are gone
            P->last_prec = -1;
            P->last_nrhs = 0;
        }
        HIPCHK(hipMalloc(&P->wskel, sb));
        HIPCHK(hipMalloc(&P->uskel, sb));
        P->skel_bytes = sb;
    }
    if (need_io) ensure_io(P, nrhs);
}

void ensure_io(mx_plan *P, int nrhs) {
    size_t ib = (size_t)std::max<int64_t>(P->N, 1) * nrhs * sizeof(float);
    if (ib > P->io_bytes) {
        if (P->io_w) HIPCHK(hipFree(P->io_w));
        if (P->io_y) HIPCHK(hipFree(P->io_y));
        P->io_w = P->io_y = nullptr;
        P->io_bytes = 0;
        ++P->buf_gen;
        HIPCHK(hipMalloc(&P->io_w, ib));
        HIPCHK(hipMalloc(&P->io_y, ib));
        P->io_bytes = ib;
    }
}

size_t panel_bytes_of(const mx_plan *P, int nrhs) {
    return ((size_t)2 * P->wp_rows + (size_t)4 * P->sp_rows) * nrhs * sizeof(uint16_t);
}

size_t panel_ctl_bytes_of(const mx_plan *P, int nrhs) {
    return ((panel_bytes_of(P, nrhs) + 255) & ~(size_t)255) + (size_t)sweep_ctl_ints(P->numnodes, nrhs) * sizeof(int32_t);
}

int sweep_grid_for(mx_plan *P, int dir, int nrhs) {
    const int v = P->sweep_var[dir];
    const int k = v == X3_V128x64_S2 ? 1 : v == X3_V128x32_S2 ? 2 : v == X3_V128x32_K8 ? 3 : 0;
    if (P->sweep_grid[k] == 0) P->sweep_grid[k] = sweep_max_grid(v);
    if (P->sweep_grid[k] <= 0) throw MxError(MX_ERR_HIP, "sweep: no occupancy for the kernel");
    const int bn = sweep_bn(v), ncols = (nrhs + bn - 1) / bn;
    int64_t items = 0;
    for (const int2 &qi : P->sweep[dir].qinfo_h) items += (int64_t)qi.y * ncols;
    return (int)std::min<int64_t>(P->sweep_grid[k], std::max<int64_t>(items, 1));
}

// The panels buffer holds the split panels, then (from a 256-byte boundary)
// the sweep's control block.
void ensure_panels(mx_plan *P, int nrhs) {
    size_t bytes = panel_ctl_bytes_of(P, nrhs);
    if (bytes > P->panel_bytes) {
        if (P->panels) HIPCHK(hipFree(P->panels));
        P->panels = nullptr;
        P->panel_bytes = 0;
        ++P->buf_gen;
        P->panels_dir = -1;   // the split skeleton panels are gone
        P->panels_nrhs = 0;
        HIPCHK(hipMalloc(&P->panels, bytes));
        P->panel_bytes = bytes;
    }
}

void note_eval(mx_plan *P, int dir, mx_precision prec, int nrhs) {
    if (prec == MX_BF16X3) {   // forward and adjoint share the panels
        P->panels_dir = dir;
        P->panels_nrhs = nrhs;
    }
    if (dir == DIR_FWD) {
        P->last_prec = prec;
        P->last_nrhs = nrhs;
    }
}

// FP64/FP32 kernel variants of a launch: 64 x 64 tiles with one or two
// chunks of loads in flight, (FP32) 128 x 128 tiles on the plan's 128-row
// tile list, or f64p_kernel / f32p_kernel on packed A operands: 128 x 128
// tiles with two or three (FP32: four) LDS stages, and 128 x 256 tiles
// likewise.  A launch with a task longer than F64P_MAXSEG segments, or a
// plan without packed copies, runs the packed variants on the 64 x 64
// kernel.
enum { SEG_V64_D1 = 1, SEG_V64_D2 = 2, SEG_V128 = 3, SEG_P2 = 4, SEG_P3 = 5, SEG_PW2 = 6, SEG_PW3 = 7 };

static bool seg_variant_ok(mx_precision prec, int v) {
    return prec == MX_FP64 ? (v == SEG_V64_D1 || v == SEG_V64_D2 || (v >= SEG_P2 && v <= SEG_PW3))
                           : (v >= SEG_V64_D1 && v <= SEG_PW3);
}

template <typename T>
static void run_seg_launch(const Program &pg, size_t li, int nrhs, const Bufs<T> &b, int variant, hipStream_t st,
                           const T *const *packed = nullptr) {
    if (variant >= SEG_P2 && variant <= SEG_PW3) {
        if (packed && pg.segs_p && li < pg.maxseg.size() && pg.maxseg[li] <= F64P_MAXSEG) {
            const Launch &ln = pg.launches2[li];
            const int wide = variant >= SEG_PW2, deep = variant == SEG_P3 || variant == SEG_PW3;
            const int bn = wide ? F64P_BNW : F64P_BN;
            if constexpr (sizeof(T) == sizeof(double)) {
                F64PArgs fa{};
                fa.tasks = pg.tasks;
                fa.segs = pg.segs_p;
                fa.tiles = pg.tiles2 + ln.tile_offset;
                fa.ntiles = ln.ntiles;
                fa.nrhs = nrhs;
                fa.n_ntiles = (nrhs + bn - 1) / bn;
                for (int w = 0; w < 3; ++w) fa.a[w] = packed[w];
                for (int v = 0; v < 4; ++v) fa.vec[v] = b.vec[v];
                launch_f64p(fa, deep ? 3 : 2, wide, st);
            } else {
                F32PArgs fa{};
                fa.tasks = pg.tasks;
                fa.segs = pg.segs_p;
                fa.tiles = pg.tiles2 + ln.tile_offset;
                fa.ntiles = ln.ntiles;
                fa.nrhs = nrhs;
                fa.n_ntiles = (nrhs + bn - 1) / bn;
                for (int w = 0; w < 3; ++w) fa.a[w] = packed[w];
                for (int v = 0; v < 4; ++v) fa.vec[v] = b.vec[v];
                launch_f32p(fa, deep ? 4 : 2, wide, st);
            }
            return;
        }
        variant = SEG_V64_D1;
    }
    const bool wide = variant == SEG_V128 && sizeof(T) == sizeof(float);
    const Launch &ln = wide ? pg.launches2[li] : pg.launches[li];
    const int bn = wide ? SW_BN : SG_BN;
    int nnt = (nrhs + bn - 1) / bn;
    int64_t total = (int64_t)ln.ntiles * nnt;
    if (total <= 0) return;
    if (total > INT32_MAX) throw MxError(MX_ERR_ARG, "launch too large");
    const dim3 grid((unsigned)total);
    if (wide)
        hipLaunchKernelGGL(seg_gemm_wide_kernel<float>, grid, dim3(SG_THREADS), 0, st, pg.tasks, pg.segs,
                           pg.tiles2 + ln.tile_offset, ln.ntiles, nrhs, nnt, reinterpret_cast<const Bufs<float> &>(b));
    else if (variant == SEG_V64_D2)
        hipLaunchKernelGGL((seg_gemm_kernel<T, 2>), grid, dim3(SG_THREADS), 0, st, pg.tasks, pg.segs,
                           pg.tiles + ln.tile_offset, ln.ntiles, nrhs, nnt, b);
    else
        hipLaunchKernelGGL((seg_gemm_kernel<T, 1>), grid, dim3(SG_THREADS), 0, st, pg.tasks, pg.segs,
                           pg.tiles + ln.tile_offset, ln.ntiles, nrhs, nnt, b);
}

// kernel variant of launch i of the FP64/FP32 program in use (SEG_V64_D1
// unless the autotuner chose otherwise)
static int seg_variant(const mx_plan *P, int dir, int kd, size_t i) {
    const std::vector<int> &d = P->lsg[dir][kd][P->sched[dir][0]];
    return i < d.size() ? d[i] : SEG_V64_D1;
}

static void prof_mark(mx_plan *P, hipStream_t st, int kind, bool begin) {
    if (!P->profiling) return;
    if (begin) {
        size_t need = 2 * (size_t)P->ev_used + 2;
        while (P->ev.size() < need) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            P->ev.push_back(e);
        }
        if ((int)P->ev_kind.size() <= P->ev_used) P->ev_kind.resize(P->ev_used + 1);
        P->ev_kind[P->ev_used] = kind;
        HIPCHK(hipEventRecord(P->ev[2 * P->ev_used], st));
    } else {
        HIPCHK(hipEventRecord(P->ev[2 * P->ev_used + 1], st));
        ++P->ev_used;
    }
}

// Checks shared by every evaluation entry point.
void eval_checks(const mx_plan *P, int dir, mx_precision prec, const void *W, const void *Y, int nrhs,
                 mx_io_type io) {
    arg_check(W && Y, "null W or Y");
    arg_check(nrhs >= 1, "nrhs must be >= 1");
    arg_check(io == MX_IO_F64 || io == MX_IO_F32, "bad io type");
    int dev;
    HIPCHK(hipGetDevice(&dev));
    if (dev != P->device) throw MxError(MX_ERR_STATE, "plan was created on another device");
    if (!(P->precisions & (1u << prec))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the plan");
    if (dir == DIR_ADJ && !P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
    if (prec == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 evaluation needs FP64 W and Y");
}

void size_work(mx_plan *P, int dir, mx_precision prec, mx_io_type io, int nrhs) {
    if (prec == MX_BF16X3) {
        ensure_panels(P, nrhs);
        if (P->sweep_on[dir] && P->sweep[dir].ok) (void)sweep_grid_for(P, dir, nrhs);   // occupancy, before any capture
        return;
    }
    const size_t elem = prec == MX_FP64 ? sizeof(double) : sizeof(float);
    const bool need_io = prec == MX_FP32 && io == MX_IO_F64;
    if (dir == DIR_ADJ) {
        ensure_adj_work(P, nrhs, elem);
        if (need_io) ensure_io(P, nrhs);
    } else {
        ensure_work(P, nrhs, elem, need_io);
    }
}

void x3_args_of(const mx_plan *P, int nrhs, uint16_t *panels, X3Args &xa) {
    const size_t wp = (size_t)P->wp_rows * nrhs, sp = (size_t)P->sp_rows * nrhs;
    xa.nrhs = nrhs;
    for (int w = 0; w < 3; ++w) {
        xa.ahi[w] = P->xhi[w];
        xa.alo[w] = P->xlo[w];
    }
    xa.phi[VEC_W] = panels;
    xa.plo[VEC_W] = panels + wp;
    xa.phi[VEC_WSKEL] = panels + 2 * wp;
    xa.plo[VEC_WSKEL] = panels + 2 * wp + sp;
    xa.phi[VEC_USKEL] = panels + 2 * wp + 2 * sp;
    xa.plo[VEC_USKEL] = panels + 2 * wp + 3 * sp;
}

void eval_core(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
               hipStream_t st, const Work &wk) {
    const Program &pg0 = P->prog[dir][0][P->sched[dir][0]];
    const Program &pg1 = P->prog[dir][1][P->sched[dir][1]];
    const int64_t nel = P->N * (int64_t)nrhs;
    auto mark = [&](int kind, bool begin) {
        if (wk.prof) prof_mark(P, st, kind, begin);
    };
    auto phase = [&](size_t i) {   // events after the last writer of each skeleton buffer
        if (wk.after_wskel && (int)i == pg0.last_write[VEC_WSKEL]) HIPCHK(hipEventRecord(wk.after_wskel, st));
        if (wk.after_uskel && (int)i == pg0.last_write[VEC_USKEL]) HIPCHK(hipEventRecord(wk.after_uskel, st));
    };
    if (prec == MX_FP64) {
        Bufs<double> b;
        b.mat[MAT_D] = P->D64;
        b.mat[MAT_B] = P->B64;
        b.mat[MAT_VT] = P->VT64;
        b.vec[VEC_W] = const_cast<double *>(static_cast<const double *>(W));
        b.vec[VEC_WSKEL] = (double *)wk.wsk;
        b.vec[VEC_USKEL] = (double *)wk.usk;
        b.vec[VEC_Y] = (double *)Y;
        for (size_t i = 0; i < pg0.launches.size(); ++i) {
            mark(MX_KIND_SEG_F64, true);
            run_seg_launch<double>(pg0, i, nrhs, b, seg_variant(P, dir, 0, i), st,
                                   P->f64p[0] ? P->f64p : nullptr);
            mark(MX_KIND_SEG_F64, false);
            phase(i);
        }
    } else if (prec == MX_FP32) {
        const float *Wf;
        float *Yf;
        if (io == MX_IO_F64) {
            mark(MX_KIND_CONVERT, true);
            cvt_f64_f32<<<cvt_grid(nel), 256, 0, st>>>((const double *)W, (float *)wk.io_w, nel);
            mark(MX_KIND_CONVERT, false);
            Wf = (const float *)wk.io_w;
            Yf = (float *)wk.io_y;
        } else {
            Wf = (const float *)W;
            Yf = (float *)Y;
        }
        Bufs<float> b;
        b.mat[MAT_D] = P->D32;
        b.mat[MAT_B] = P->B32;
        b.mat[MAT_VT] = P->VT32;
        b.vec[VEC_W] = const_cast<float *>(Wf);
        b.vec[VEC_WSKEL] = (float *)wk.wsk;
        b.vec[VEC_USKEL] = (float *)wk.usk;
        b.vec[VEC_Y] = Yf;
        for (size_t i = 0; i < pg0.launches.size(); ++i) {
            mark(MX_KIND_SEG_F32, true);
            run_seg_launch<float>(pg0, i, nrhs, b, seg_variant(P, dir, 1, i), st, P->f32p[0] ? P->f32p : nullptr);
            mark(MX_KIND_SEG_F32, false);
            phase(i);
        }
        if (io == MX_IO_F64) {
            mark(MX_KIND_CONVERT, true);
            cvt_f32_f64<<<cvt_grid(nel), 256, 0, st>>>(Yf, (double *)Y, nel);
            mark(MX_KIND_CONVERT, false);
        }
    } else {
        X3Args xa{};
        x3_args_of(P, nrhs, wk.panels, xa);
        const std::vector<int> &lv = P->lvar[dir][P->sched[dir][1]];
        xa.y_f64 = io == MX_IO_F64;
        xa.Y = Y;
        xa.tasks = pg1.tasks;
        xa.segs = pg1.segs;
        mark(MX_KIND_CONVERT, true);
        launch_x3_split_w(W, io == MX_IO_F64, P->d_wrows, P->nleaf, nrhs, xa.phi[VEC_W], xa.plo[VEC_W], st);
        mark(MX_KIND_CONVERT, false);
        if (P->sweep_on[dir] && P->sweep[dir].ok) {
            const mx_plan::Sweep &sw = P->sweep[dir];
            if (!wk.sweep_ctl) throw MxError(MX_ERR_STATE, "internal: no sweep control block");
            X3SweepArgs sa{};
            sa.x = xa;
            sa.x.tasks = sw.tasks;
            sa.x.segs = sw.segs;
            sa.q = sw.q;
            sa.qinfo = sw.qinfo;
            sa.st = sw.st;
            sa.w = sw.w;
            sa.ctl = wk.sweep_ctl;
            const int v = P->sweep_var[dir], bn = sweep_bn(v);
            sa.ncols = (nrhs + bn - 1) / bn;
            const int ctl_ints = SWEEP_CNT + 2 * P->numnodes * sa.ncols;
            const int grid = sweep_grid_for(P, dir, nrhs);
            mark(MX_KIND_BF16X3, true);
            launch_x3_sweep(sa, v, grid, ctl_ints, st);
            mark(MX_KIND_BF16X3, false);
            P->sweep_last_ctl = wk.sweep_ctl;
            HIPCHK(hipGetLastError());
            return;
        }
        for (size_t i = 0; i < pg1.launches.size(); ++i) {
            const int v = lv.size() == pg1.launches.size() ? lv[i] : P->x3_variant;
            const int bn = x3_variant_bn(v);
            const bool pairs = x3_variant_bm(v) == 2 * X3_BM;
            const Launch &ln = pairs ? pg1.launches2[i] : pg1.launches[i];
            xa.n_ntiles = (nrhs + bn - 1) / bn;
            xa.tiles = (pairs ? pg1.tiles2 : pg1.tiles) + ln.tile_offset;
            xa.ntiles = ln.ntiles;
            mark(MX_KIND_BF16X3, true);
            launch_x3(xa, v, st);
            mark(MX_KIND_BF16X3, false);
        }
    }
    HIPCHK(hipGetLastError());
}

void eval(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
          hipStream_t st) {
    eval_checks(P, dir, prec, W, Y, nrhs, io);
    size_work(P, dir, prec, io, nrhs);
    Work wk;
    wk.wsk = dir == DIR_ADJ ? P->wskel_adj : P->wskel;
    wk.usk = dir == DIR_ADJ ? P->uskel_adj : P->uskel;
    wk.io_w = P->io_w;
    wk.io_y = P->io_y;
    wk.panels = P->panels;
    if (prec == MX_BF16X3)
        wk.sweep_ctl = reinterpret_cast<int32_t *>(reinterpret_cast<unsigned char *>(P->panels) +
                                                   ((panel_bytes_of(P, nrhs) + 255) & ~(size_t)255));
    wk.prof = P->profiling != 0;
    if (!P->keep_prof) P->ev_used = 0;
    eval_core(P, dir, prec, W, Y, nrhs, io, st, wk);
    note_eval(P, dir, prec, nrhs);
}

// FP32 copies of the skeleton panels of the last BF16x3 evaluation (of
// direction dir, nrhs columns) in the reference skeleton layout.
static void export_panels(mx_plan *P, int dir, int nrhs, float *wdst, float *udst, hipStream_t st) {
    if (P->panels_dir != dir || P->panels_nrhs != nrhs)
        throw MxError(MX_ERR_STATE, "the skeleton panels do not hold that evaluation");
    const size_t wp = (size_t)P->wp_rows * nrhs, sp = (size_t)P->sp_rows * nrhs;
    uint16_t *pw = P->panels;
    launch_x3_panels_to_f32(pw + 2 * wp, pw + 2 * wp + sp, P->d_srows, P->n_srows, nrhs, wdst, st);
    launch_x3_panels_to_f32(pw + 2 * wp + 2 * sp, pw + 2 * wp + 3 * sp, P->d_srows, P->n_srows, nrhs, udst, st);
    HIPCHK(hipGetLastError());
}

// K-fragment buffers of the six gradient operands (hi, lo) for nrhs columns:
// leaf-blocked operands (W, dY) hold kfr_leaf 16-row groups, skeleton ones
// kfr_skel.
static void ensure_kfrag(mx_plan *P, int nrhs) {
    const size_t nfk = (size_t)(nrhs + 31) / 32;
    const size_t need = (2 * (size_t)P->kfr_leaf + 4 * (size_t)P->kfr_skel) * nfk * X3_FRAG * 2 * sizeof(uint16_t);
    if (need > P->kfrag_bytes) {
        if (P->kfrag) HIPCHK(hipFree(P->kfrag));
        P->kfrag = nullptr;
        HIPCHK(hipMalloc(&P->kfrag, std::max<size_t>(need, 1)));
        P->kfrag_bytes = need;
    }
}

// hi and lo K-fragments of gradient operand b (GradBuf) for nrhs columns
static void kfrag_of(mx_plan *P, int b, int nrhs, uint16_t **hi, uint16_t **lo) {
    const int64_t nfk = (nrhs + 31) / 32;
    const int64_t leaf = P->kfr_leaf * nfk * X3_FRAG, skel = P->kfr_skel * nfk * X3_FRAG;
    const int64_t off = b < GB_WHAT ? 2 * leaf * b : 4 * leaf + 2 * skel * (b - GB_WHAT);
    *hi = P->kfrag + off;
    *lo = *hi + (b < GB_WHAT ? leaf : skel);
}

// Skeleton panels of the last BF16x3 evaluation (direction dir) into the
// K-fragments of operands bw (its wskel panel) and bu (its uskel panel); a
// negative operand is skipped.
static void kpack_skel_panels(mx_plan *P, int dir, int nrhs, int bw, int bu, hipStream_t st) {
    if (P->panels_dir != dir || P->panels_nrhs != nrhs)
        throw MxError(MX_ERR_STATE, "the skeleton panels do not hold that evaluation");
    const size_t wp = (size_t)P->wp_rows * nrhs, sp = (size_t)P->sp_rows * nrhs;
    const uint16_t *pw = P->panels;
    uint16_t *hi, *lo;
    if (bw >= 0) {
        kfrag_of(P, bw, nrhs, &hi, &lo);
        launch_kpack_panels(pw + 2 * wp, pw + 2 * wp + sp, P->kb_panel, P->n_kb_skel, nrhs, hi, lo, st);
    }
    if (bu >= 0) {
        kfrag_of(P, bu, nrhs, &hi, &lo);
        launch_kpack_panels(pw + 2 * wp + 2 * sp, pw + 2 * wp + 3 * sp, P->kb_panel, P->n_kb_skel, nrhs, hi, lo,
                            st);
    }
}

template <typename T>
static void kpack_cols(mx_plan *P, const void *src, bool leaf, int b, int nrhs, hipStream_t st) {
    uint16_t *hi, *lo;
    kfrag_of(P, b, nrhs, &hi, &lo);
    launch_kpack_cols((const T *)src, leaf ? P->kb_leaf : P->kb_skel, leaf ? P->n_kb_leaf : P->n_kb_skel, nrhs, hi,
                      lo, st);
}

// Backward pass of Y = K~ W for the upstream gradient dY: dW = K~^T dY by the
// adjoint evaluation, then (when requested) the block gradients from the
// forward skeletons (w^, u^: caller copies, or the plan's own buffers) and the
// adjoint ones (y^, g^).  FP64, FP32 or split-BF16x3 compute; outputs in the
// I/O type.  Split-BF16x3 products read every operand as K-fragments, packed
// from the I/O arrays, from caller FP32 skeletons, or straight from the
// evaluations' skeleton panels.
void backward(mx_plan *P, mx_precision prec, const void *W, const void *dY, void *dW, int nrhs, mx_io_type io,
              const void *wskel, const void *uskel, void *dD, void *dB, void *dVT, hipStream_t st) {
    arg_check(W && dY && dW, "null W, dY or dW");
    arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
    arg_check(io == MX_IO_F64 || io == MX_IO_F32, "bad io type");
    if (prec == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 backward needs FP64 I/O");
    if (!P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
    if (!(P->precisions & (1u << prec))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the plan");
    const bool grads = dD || dB || dVT;
    const bool x3 = prec == MX_BF16X3;
    // operands of the requested products: dD = dY W^T, dB = y^ w^^T,
    // dVT = g^ [W | w^]^T + u^ [dY | y^]^T
    const bool need_w_dy = dD || dVT, need_skel_w = dB || dVT, need_skel_u = dVT != nullptr;
    // the products read W and dY after the adjoint has written dW
    arg_check(!grads || (dW != W && dW != dY), "dW must not alias W or dY when block gradients are requested");
    if (grads && x3) {
        ensure_kfrag(P, nrhs);
        if (wskel && uskel) {   // caller FP32 skeletons
            if (need_skel_w) kpack_cols<float>(P, wskel, false, GB_WHAT, nrhs, st);
            if (need_skel_u) kpack_cols<float>(P, uskel, false, GB_UHAT, nrhs, st);
        } else {                // the forward's panels, before the adjoint reuses them
            if (P->last_prec != prec)
                throw MxError(MX_ERR_STATE, "no forward evaluation in this precision to take skeletons from");
            prof_mark(P, st, MX_KIND_CONVERT, true);
            kpack_skel_panels(P, DIR_FWD, nrhs, need_skel_w ? GB_WHAT : -1, need_skel_u ? GB_UHAT : -1, st);
            prof_mark(P, st, MX_KIND_CONVERT, false);
        }
    } else if (grads && (!wskel || !uskel)) {
        if (P->last_prec != prec) throw MxError(MX_ERR_STATE, "no forward evaluation in this precision to take skeletons from");
        if (P->last_nrhs != nrhs)
            throw MxError(MX_ERR_STATE, "the last forward evaluation had another number of columns");
        wskel = P->wskel;
        uskel = P->uskel;
    }
    if (!P->keep_prof) P->ev_used = 0;
    P->keep_prof = 1;
    try {
        eval(P, DIR_ADJ, prec, dY, dW, nrhs, io, st);
    } catch (...) {
        P->keep_prof = 0;
        throw;
    }
    P->keep_prof = 0;
    if (!grads) return;
    const int64_t nel = P->N * (int64_t)nrhs;
    const int kind = prec == MX_FP64 ? MX_KIND_GRAD_F64 : (prec == MX_FP32 ? MX_KIND_GRAD_F32 : MX_KIND_GRAD_X3);
    if (x3) {
        prof_mark(P, st, MX_KIND_CONVERT, true);
        kpack_skel_panels(P, DIR_ADJ, nrhs, need_skel_w ? GB_YHAT : -1, need_skel_u ? GB_GHAT : -1, st);
        if (need_w_dy) {
            if (io == MX_IO_F64) {
                kpack_cols<double>(P, W, true, GB_W, nrhs, st);
                kpack_cols<double>(P, dY, true, GB_DY, nrhs, st);
            } else {
                kpack_cols<float>(P, W, true, GB_W, nrhs, st);
                kpack_cols<float>(P, dY, true, GB_DY, nrhs, st);
            }
        }
        prof_mark(P, st, MX_KIND_CONVERT, false);
        const bool wide = P->g_x3_tile == GRAD_X3_WTILE || (P->g_x3_tile == 0 && P->g_x3w_auto);
        auto fill = [&](auto &g) {
            g.tasks = P->g_tasks;
            g.segs = P->g_ksegs;
            g.tiles = wide ? P->g_tiles_x3w : P->g_tiles_x3;
            g.ntiles = wide ? P->g_ntiles_x3w : P->g_ntiles_x3;
            g.nrhs = nrhs;
            for (int b = 0; b < GB_COUNT; ++b) {
                uint16_t *hi, *lo;
                kfrag_of(P, b, nrhs, &hi, &lo);
                g.khi[b] = hi;
                g.klo[b] = lo;
            }
        };
        prof_mark(P, st, kind, true);
        if (io == MX_IO_F64) {
            GradX3Args<double> g{};
            fill(g);
            g.out[MAT_D] = (double *)dD;
            g.out[MAT_B] = (double *)dB;
            g.out[MAT_VT] = (double *)dVT;
            if (wide) launch_grad_x3w(g, st);
            else launch_grad_x3(g, st);
        } else {
            GradX3Args<float> g{};
            fill(g);
            g.out[MAT_D] = (float *)dD;
            g.out[MAT_B] = (float *)dB;
            g.out[MAT_VT] = (float *)dVT;
            if (wide) launch_grad_x3w(g, st);
            else launch_grad_x3(g, st);
        }
        prof_mark(P, st, kind, false);
    } else if (prec == MX_FP64) {
        GradArgs<double, double> g{};
        g.tasks = P->g_tasks;
        g.segs = P->g_segs;
        g.tiles = P->g_tiles;
        g.ntiles = P->g_ntiles;
        g.nrhs = nrhs;
        g.src[GB_W] = (const double *)W;
        g.src[GB_DY] = (const double *)dY;
        g.src[GB_WHAT] = (const double *)wskel;
        g.src[GB_UHAT] = (const double *)uskel;
        g.src[GB_YHAT] = (const double *)P->wskel_adj;
        g.src[GB_GHAT] = (const double *)P->uskel_adj;
        g.out[MAT_D] = (double *)dD;
        g.out[MAT_B] = (double *)dB;
        g.out[MAT_VT] = (double *)dVT;
        prof_mark(P, st, kind, true);
        launch_grad(g, st);
        prof_mark(P, st, kind, false);
    } else {
        const float *Wf = (const float *)W, *dYf = (const float *)dY;
        if (io == MX_IO_F64) {   // dY (FP32) is in io_w (left by the FP32 adjoint); W goes to io_g
            ensure_io_g(P, nrhs);
            prof_mark(P, st, MX_KIND_CONVERT, true);
            cvt_f64_f32<<<cvt_grid(nel), 256, 0, st>>>((const double *)W, (float *)P->io_g, nel);
            prof_mark(P, st, MX_KIND_CONVERT, false);
            Wf = (const float *)P->io_g;
            dYf = (const float *)P->io_w;
        }
        auto fill = [&](auto &g) {
            g.tasks = P->g_tasks;
            g.segs = P->g_segs;
            g.tiles = P->g_tiles;
            g.ntiles = P->g_ntiles;
            g.nrhs = nrhs;
            g.src[GB_W] = Wf;
            g.src[GB_DY] = dYf;
            g.src[GB_WHAT] = (const float *)wskel;
            g.src[GB_UHAT] = (const float *)uskel;
            g.src[GB_YHAT] = (const float *)P->wskel_adj;
            g.src[GB_GHAT] = (const float *)P->uskel_adj;
        };
        prof_mark(P, st, kind, true);
        if (io == MX_IO_F64) {
            GradArgs<float, double> g{};
            fill(g);
            g.out[MAT_D] = (double *)dD;
            g.out[MAT_B] = (double *)dB;
            g.out[MAT_VT] = (double *)dVT;
            launch_grad(g, st);
        } else {
            GradArgs<float, float> g{};
            fill(g);
            g.out[MAT_D] = (float *)dD;
            g.out[MAT_B] = (float *)dB;
            g.out[MAT_VT] = (float *)dVT;
            launch_grad(g, st);
        }
        prof_mark(P, st, kind, false);
    }
    HIPCHK(hipGetLastError());
}

// Replace the stored blocks (same structure) and re-derive every precision
// built into the plan.  Captured graphs stay valid: they read the same
// buffers.
static void set_blocks(mx_plan *P, const double *D, const double *B, const double *VT, hipStream_t st) {
    const double *src[3] = {D, B, VT};
    const int64_t len[3] = {P->D_len, P->B_len, P->VT_len};
    double *d64[3] = {P->D64, P->B64, P->VT64};
    float *d32[3] = {P->D32, P->B32, P->VT32};
    for (int w = 0; w < 3; ++w) {
        if (!src[w] || len[w] == 0) continue;
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, src[w]) != hipSuccess || at.type != hipMemoryTypeDevice) {
            (void)hipGetLastError();
            throw MxError(MX_ERR_ARG, "stored blocks must be device arrays");
        }
        if (d64[w]) HIPCHK(hipMemcpyAsync(d64[w], src[w], (size_t)len[w] * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (d32[w]) cvt_f64_f32<<<cvt_grid(len[w]), 256, 0, st>>>(src[w], d32[w], len[w]);
    }
    if (P->precisions & MX_PLAN_BF16X3) {
        pack_all(P, src, st);
        // composites: from the kept FP64 copies, updated above
        if (P->n_comp && (src[MAT_B] || src[MAT_VT])) form_composites(P, P->B64, P->VT64, st);
    }
    if (P->precisions & MX_PLAN_FP64) pack_f64_all(P, src, st);
    if (P->precisions & MX_PLAN_FP32) {   // from the FP32 copies just converted (src[w] null: unchanged)
        const float *s32[3] = {src[0] ? d32[0] : nullptr, src[1] ? d32[1] : nullptr, src[2] ? d32[2] : nullptr};
        pack_f32_all(P, s32, st);
    }
    HIPCHK(hipGetLastError());
}

// Unit-variance pseudo-random fill (splitmix64 of the index, uniform on
// [-sqrt(3), sqrt(3))): timing inputs with realistic bit patterns (MFMA
// power, and so clocks, depend on the data).
__global__ void fill_uniform(void *p, int64_t n, int f64) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = ((double)(z >> 11) * 0x1.0p-53 - 0.5) * 3.4641016151377544;
        if (f64) static_cast<double *>(p)[i] = u;
        else static_cast<float *>(p)[i] = (float)u;
    }
}

void launch_fill_uniform(void *p, int64_t n, bool f64, hipStream_t st) {
    fill_uniform<<<cvt_grid(n), 256, 0, st>>>(p, n, f64 ? 1 : 0);
    HIPCHK(hipGetLastError());
}

static void drop_graphs(mx_plan *P) {
    for (auto &g : P->graphs) HIPCHK(hipGraphExecDestroy(g.exec));
    P->graphs.clear();
}

static void set_schedule(mx_plan *P, int dir, mx_precision prec, int sched, int variant) {
    arg_check(dir == DIR_FWD || dir == DIR_ADJ, "bad direction");
    arg_check((sched >= 0 && sched < SCHED_COARSE) || sched == MX_SCHED_SWEEP || sched == MX_SCHED_COARSE ||
                  sched == MX_SCHED_COARSE_SPLIT,
              "bad schedule");
    arg_check(variant >= 0 && variant <= X3_NVARIANT, "bad BF16x3 kernel variant (0 keeps)");
    if (dir == DIR_ADJ && !P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
    const int kd = prec == MX_BF16X3 ? 1 : 0;
    if (sched == MX_SCHED_SWEEP) {
        if (prec != MX_BF16X3) throw MxError(MX_ERR_UNSUPPORTED, "the sweep schedule is for MX_BF16X3");
        if (!(P->precisions & MX_PLAN_BF16X3)) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the plan");
        if (!P->sweep[dir].ok)
            throw MxError(MX_ERR_UNSUPPORTED, "no sweep for this plan (a block with several writers)");
        arg_check(variant == 0 || sweep_variant_ok(variant),
                  "sweep tile: MX_X3_128x128_S2, MX_X3_128x64_S2, MX_X3_128x32_S2 or MX_X3_128x32_K8 (0 keeps)");
        P->sweep_on[dir] = 1;
        if (variant) P->sweep_var[dir] = variant;
        drop_graphs(P);
        return;
    }
    if (sched == MX_SCHED_COARSE || sched == MX_SCHED_COARSE_SPLIT) {
        sched = sched == MX_SCHED_COARSE ? SCHED_COARSE : SCHED_COARSE_SPLIT;
        if (prec != MX_BF16X3) throw MxError(MX_ERR_UNSUPPORTED, "the coarse schedules are for MX_BF16X3");
        if (!P->prog[dir][1][sched].built)
            throw MxError(MX_ERR_UNSUPPORTED, "no coarse schedule in this plan (BF16x3 lists of stored blocks)");
    }
    if (kd == 1) P->sweep_on[dir] = 0;
    P->sched[dir][kd] = sched;
    if (variant) {
        P->x3_variant = variant;
        for (auto &d : P->lvar)
            for (auto &v : d) v.clear();
    }
    drop_graphs(P);
}

// Time every candidate (schedule x BF16x3 kernel variant) of each direction
// built into the plan on scratch inputs and keep the fastest.  A candidate
// replaces the current choice only when it is more than 2% faster, so that
// timing noise does not flip between near-equal choices.
void autotune(mx_plan *P, mx_precision prec, int nrhs, mx_io_type io, int reps, hipStream_t st,
              float *best_ms) {
    arg_check(nrhs >= 1 && reps >= 1, "bad nrhs or reps");
    if (!(P->precisions & (1u << prec))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the plan");
    const size_t elem = io == MX_IO_F64 ? sizeof(double) : sizeof(float);
    const size_t bytes = (size_t)std::max<int64_t>(P->N, 1) * nrhs * elem;
    void *W = nullptr, *Y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int graph = P->use_graph, prof = P->profiling;
    P->use_graph = 0;
    P->profiling = 0;
    auto cleanup = [&]() {
        if (W) (void)hipFree(W);
        if (Y) (void)hipFree(Y);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        P->use_graph = graph;
        P->profiling = prof;
    };
    try {
        HIPCHK(hipMalloc(&W, bytes));
        HIPCHK(hipMalloc(&Y, bytes));
        const int64_t nel = std::max<int64_t>(P->N, 1) * nrhs;
        fill_uniform<<<cvt_grid(nel), 256, 0, st>>>(W, nel, io == MX_IO_F64);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const int kd = prec == MX_BF16X3 ? 1 : 0;
        const int ndir = P->adjoint ? 2 : 1;
        const int nvar = prec == MX_BF16X3 ? X3_NVARIANT : 1;
        if (prec == MX_BF16X3) P->sweep_on[0] = P->sweep_on[1] = 0;   // level launches first
        {
            // short evaluations are timed over more repetitions (about 2 ms
            // per measurement, at most 20): their launch-to-launch jitter is
            // a larger share of their time
            eval(P, DIR_FWD, prec, W, Y, nrhs, io, st);   // warm-up
            HIPCHK(hipEventRecord(e0, st));
            eval(P, DIR_FWD, prec, W, Y, nrhs, io, st);
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms > 0.f) reps = std::max(reps, std::min(20, (int)std::ceil(2.0f / ms)));
        }
        const int cur_variant = P->x3_variant;
        const int cur_sched[2] = {P->sched[0][kd], P->sched[1][kd]};
        int best_variant = cur_variant;
        int best_pick[2] = {cur_sched[0], cur_sched[1]};
        float best_total = -1.f, best_dir[2] = {0.f, 0.f};
        // the current variant first, so that it is the reference
        for (int si = 0; si < nvar; ++si) {
            if (prec == MX_BF16X3) P->x3_variant = si == 0 ? cur_variant : (si <= cur_variant - 1 ? si : si + 1);
            float dir_ms[2] = {0.f, 0.f};
            float total = 0.f;
            int pick[2] = {cur_sched[0], cur_sched[1]};
            for (int dir = 0; dir < ndir; ++dir) {
                float best = -1.f;
                for (int si2 = 0; si2 < NSCHED; ++si2) {
                    // the current schedule first
                    const int sc = si2 == 0 ? cur_sched[dir] : (si2 <= cur_sched[dir] ? si2 - 1 : si2);
                    if (!P->prog[dir][kd][sc].built) continue;
                    P->sched[dir][kd] = sc;
                    eval(P, dir, prec, W, Y, nrhs, io, st);   // warm-up (sizes work buffers)
                    HIPCHK(hipEventRecord(e0, st));
                    for (int r = 0; r < reps; ++r) eval(P, dir, prec, W, Y, nrhs, io, st);
                    HIPCHK(hipEventRecord(e1, st));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    ms /= reps;
                    if (best < 0.f || ms < 0.98f * best) {
                        best = ms;
                        pick[dir] = sc;
                    }
                }
                P->sched[dir][kd] = pick[dir];
                total += best;
                dir_ms[dir] = best;
            }
            if (best_total < 0.f || total < 0.98f * best_total) {
                best_total = total;
                best_variant = P->x3_variant;
                best_pick[0] = pick[0];
                best_pick[1] = pick[1];
                best_dir[0] = dir_ms[0];
                best_dir[1] = dir_ms[1];
            }
        }
        if (prec == MX_BF16X3) P->x3_variant = best_variant;
        P->sched[0][kd] = best_pick[0];
        P->sched[1][kd] = best_pick[1];
        if (prec != MX_BF16X3) {
            // per launch: 64 x 64 tiles with prefetch depth 2 for
            // latency-bound launches, depth 1 where more workgroups per CU
            // pay, 128 x 128 tiles (FP32; FP64 on packed operands) for the
            // wide launches
            const int seg_kind = prec == MX_FP64 ? MX_KIND_SEG_F64 : MX_KIND_SEG_F32;
            const int pk = prec == MX_FP64 ? 0 : 1;
            std::vector<int> cand = prec == MX_FP64 ? std::vector<int>{SEG_V64_D1, SEG_V64_D2}
                                                    : std::vector<int>{SEG_V64_D1, SEG_V64_D2, SEG_V128};
            if ((prec == MX_FP64 && P->f64p[0]) || (prec == MX_FP32 && P->f32p[0]))
                cand.insert(cand.end(), {SEG_P2, SEG_P3, SEG_PW2, SEG_PW3});
            P->profiling = 1;
            for (int dir = 0; dir < ndir; ++dir) {
                const int sc = P->sched[dir][0];
                const size_t nl = P->prog[dir][0][sc].launches.size();
                std::vector<int> &ld = P->lsg[dir][pk][sc];
                std::vector<int> pick(nl, SEG_V64_D1);
                std::vector<float> best(nl, 0.f);
                for (int depth : cand) {
                    ld.assign(nl, depth);
                    eval(P, dir, prec, W, Y, nrhs, io, st);   // warm-up
                    std::vector<float> t(nl, 0.f);
                    for (int r = 0; r < reps; ++r) {
                        eval(P, dir, prec, W, Y, nrhs, io, st);
                        HIPCHK(hipEventSynchronize(P->ev[2 * P->ev_used - 1]));
                        size_t li = 0;
                        for (int e = 0; e < P->ev_used && li < nl; ++e) {
                            if (P->ev_kind[e] != seg_kind) continue;
                            float ms = 0.f;
                            HIPCHK(hipEventElapsedTime(&ms, P->ev[2 * e], P->ev[2 * e + 1]));
                            t[li++] += ms;
                        }
                    }
                    for (size_t i = 0; i < nl; ++i)
                        if (depth == SEG_V64_D1 || t[i] < 0.97f * best[i]) {
                            best[i] = t[i];
                            pick[i] = depth;
                        }
                }
                ld = pick;
            }
            P->profiling = 0;
        }
        if (prec == MX_BF16X3) {
            // per launch: the deep top-of-tree launches and the wide leaf
            // launches favour different variants.  Tuned for the schedule
            // picked above and for the coarse lists, whose merged launches
            // favour other variants than the plan-wide one it was timed
            // with; the faster of the two tuned evaluations is kept.
            auto tune_launches = [&](int dir, int sc) {
                P->sched[dir][1] = sc;
                const size_t nl = P->prog[dir][1][sc].launches.size();
                std::vector<int> &lv = P->lvar[dir][sc];
                std::vector<int> pick(nl, P->x3_variant);
                std::vector<float> best(nl, 0.f);
                P->profiling = 1;
                for (int vi = 0; vi < X3_NVARIANT; ++vi) {
                    const int v = vi == 0 ? P->x3_variant : (vi <= P->x3_variant - 1 ? vi : vi + 1);
                    lv.assign(nl, v);
                    eval(P, dir, prec, W, Y, nrhs, io, st);   // warm-up
                    std::vector<float> t(nl, 0.f);
                    for (int r = 0; r < reps; ++r) {
                        eval(P, dir, prec, W, Y, nrhs, io, st);
                        HIPCHK(hipEventSynchronize(P->ev[2 * P->ev_used - 1]));
                        // event pairs: the W split, then one per BF16x3 launch
                        for (size_t i = 0; i < nl && (int)i + 1 < P->ev_used; ++i) {
                            float ms = 0.f;
                            HIPCHK(hipEventElapsedTime(&ms, P->ev[2 * (i + 1)], P->ev[2 * (i + 1) + 1]));
                            t[i] += ms;
                        }
                    }
                    for (size_t i = 0; i < nl; ++i)
                        if (vi == 0 || t[i] < 0.97f * best[i]) {
                            best[i] = t[i];
                            pick[i] = v;
                        }
                }
                lv = pick;
                P->profiling = 0;
            };
            auto time_eval = [&](int dir) {
                eval(P, dir, prec, W, Y, nrhs, io, st);   // warm-up
                HIPCHK(hipEventRecord(e0, st));
                for (int r = 0; r < reps; ++r) eval(P, dir, prec, W, Y, nrhs, io, st);
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                return ms / reps;
            };
            for (int dir = 0; dir < ndir; ++dir) {
                const int sc = P->sched[dir][1];
                tune_launches(dir, sc);
                float best = time_eval(dir);
                int pick = sc;
                for (int co : {SCHED_COARSE, SCHED_COARSE_SPLIT}) {
                    if (co == sc || !P->prog[dir][1][co].built) continue;
                    tune_launches(dir, co);
                    const float t = time_eval(dir);
                    if (t < 0.98f * best) {
                        best = t;
                        pick = co;
                    }
                }
                P->sched[dir][1] = pick;
                best_dir[dir] = best;
            }
            // the dependency-driven sweep (three tile widths) against the
            // tuned level launches
            for (int dir = 0; dir < ndir; ++dir) {
                if (!P->sweep[dir].ok) continue;
                P->sweep_on[dir] = 0;
                float best = time_eval(dir);
                int pick = 0;
                for (int v : {X3_V128x128_S2, X3_V128x64_S2, X3_V128x32_S2, X3_V128x32_K8}) {
                    P->sweep_on[dir] = 1;
                    P->sweep_var[dir] = v;
                    const float ms = time_eval(dir);
                    if (ms < 0.98f * best) {
                        best = ms;
                        pick = v;
                    }
                }
                P->sweep_on[dir] = pick != 0;
                if (pick) P->sweep_var[dir] = pick;
                best_dir[dir] = best;
            }
        }
        if (best_ms) {
            best_ms[0] = best_dir[0];
            best_ms[1] = best_dir[1];
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    drop_graphs(P);
}

// Graphs captured before the last reallocation of a work buffer write
// through freed addresses: destroy them.
static void drop_stale_graphs(mx_plan *P) {
    for (size_t i = 0; i < P->graphs.size();) {
        if (P->graphs[i].gen != P->buf_gen) {
            HIPCHK(hipGraphExecDestroy(P->graphs[i].exec));
            P->graphs.erase(P->graphs.begin() + (std::ptrdiff_t)i);
        } else {
            ++i;
        }
    }
}

// Evaluation through a captured hipGraph: the launch sequence of one
// (precision, I/O type, nrhs, W, Y) is captured once and replayed; work
// buffers are sized before the capture and the sequence holds no memsets.
// A graph is replayed only while the work buffers it was captured with are
// still allocated (buf_gen unchanged).  Up to four graphs are kept per plan
// (least recently used is dropped).
static void eval_graph(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
                       hipStream_t st) {
    drop_stale_graphs(P);
    for (auto &g : P->graphs) {
        if (g.dir == dir && g.prec == prec && g.io == io && g.nrhs == nrhs && g.W == W && g.Y == Y) {
            g.used = ++P->graph_clock;
            HIPCHK(hipGraphLaunch(g.exec, st));
            note_eval(P, dir, prec, nrhs);
            return;
        }
    }
    eval_checks(P, dir, prec, W, Y, nrhs, io);
    size_work(P, dir, prec, io, nrhs);
    drop_stale_graphs(P);
    const uint64_t gen = P->buf_gen;
    const int prof = P->profiling;
    P->profiling = 0;
    hipStream_t cs = st;
    bool own = false;
    if (cs == nullptr) {   // the legacy default stream cannot be captured
        HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        own = true;
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    try {
        HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            eval(P, dir, prec, W, Y, nrhs, io, cs);
        } catch (...) {
            hipGraph_t g2 = nullptr;
            (void)hipStreamEndCapture(cs, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            throw;
        }
        HIPCHK(hipStreamEndCapture(cs, &graph));
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
        graph = nullptr;
    } catch (...) {
        P->profiling = prof;
        if (graph) (void)hipGraphDestroy(graph);
        if (own) (void)hipStreamDestroy(cs);
        throw;
    }
    P->profiling = prof;
    if (own) HIPCHK(hipStreamDestroy(cs));
    if (P->buf_gen != gen) {   // the buffers were sized above; a capture must not move them
        (void)hipGraphExecDestroy(exec);
        throw MxError(MX_ERR_STATE, "work buffers moved during graph capture");
    }
    if (P->graphs.size() >= 4) {
        auto lru = std::min_element(P->graphs.begin(), P->graphs.end(),
                                    [](const mx_plan::GraphEntry &a, const mx_plan::GraphEntry &b) {
                                        return a.used < b.used;
                                    });
        HIPCHK(hipGraphExecDestroy(lru->exec));
        P->graphs.erase(lru);
    }
    P->graphs.push_back(mx_plan::GraphEntry{dir, prec, io, nrhs, W, Y, exec, ++P->graph_clock, gen});
    HIPCHK(hipGraphLaunch(exec, st));
    note_eval(P, dir, prec, nrhs);
}

// Kr (nrows x N, column-major) = exp(-|x_rows[r] - x_j|^2 / (2 h^2)), points in
// leaf order; distances summed directly as in matroxTest-V1/Util.h:189-201.
__global__ void dense_kernel_rows(const double *__restrict__ X, int64_t N, int d, double h,
                                  const int32_t *__restrict__ rows, int nrows, double *__restrict__ Kr) {
    const double s = 1.0 / (2.0 * h * h);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N * nrows;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = e / nrows;
        int r = (int)(e % nrows);
        const double *a = X + (int64_t)rows[r] * d;
        const double *b = X + j * d;
        double k2 = 0.0;
        for (int p = 0; p < d; ++p) {
            double t = b[p] - a[p];
            k2 += t * t;
        }
        Kr[e] = exp(-k2 * s);
    }
}

// K(rows, :) W for points X_leaf (device, leaf order) of a leaf-blocked W:
// the kernel rows, then one segmented GEMM over the leaf blocks.
void dense_rows_core(const double *X_leaf, int64_t N, int d, double h, int nleaf, const int32_t *wrow,
                     const int32_t *Ddim, const int32_t *rows, int nrows, const double *W, int nrhs,
                     double *out, hipStream_t st) {
    arg_check(rows && W && out && nrows >= 1 && nrhs >= 1, "bad argument");
    for (int r = 0; r < nrows; ++r) arg_check(rows[r] >= 0 && rows[r] < N, "row index out of range");
    TaskLists L;
    Task tk{};
    tk.c_row = 0;
    tk.m = nrows;
    tk.c_buf = VEC_Y;
    tk.seg_begin = 0;
    tk.beta = 0;
    for (int l = 0; l < nleaf; ++l) {
        if (Ddim[l] == 0) continue;
        Seg sg{};
        sg.a_off = (int64_t)wrow[l] * nrows;
        sg.a_buf = MAT_D;
        sg.a_trans = 0;
        sg.lda = nrows;
        sg.k = Ddim[l];
        sg.b_buf = VEC_W;
        sg.b_row = wrow[l];
        sg.b_ld = Ddim[l];
        L.segs.push_back(sg);
    }
    tk.seg_end = (int)L.segs.size();
    L.tasks.push_back(tk);
    for (int i = 0; i < (nrows + SG_BM - 1) / SG_BM; ++i) L.tiles.push_back(make_int2(0, i));
    Task *dt = nullptr;
    Seg *ds = nullptr;
    int2 *dtl = nullptr;
    int32_t *drows = nullptr;
    double *Kr = nullptr;
    try {
        upload_lists(L, &dt, &ds, &dtl);
        drows = upload(rows, (size_t)nrows);
        HIPCHK(hipMalloc(&Kr, sizeof(double) * (size_t)nrows * N));
        int64_t tot = (int64_t)nrows * N;
        dense_kernel_rows<<<cvt_grid(tot), 256, 0, st>>>(X_leaf, N, d, h, drows, nrows, Kr);
        Bufs<double> b{};
        b.mat[MAT_D] = Kr;
        b.vec[VEC_W] = const_cast<double *>(W);
        b.vec[VEC_Y] = out;
        Program pg;   // views the lists uploaded above; they are freed below
        pg.tasks = dt;
        pg.segs = ds;
        pg.tiles = dtl;
        pg.launches.push_back(Launch{0, (int)L.tiles.size()});
        run_seg_launch<double>(pg, 0, nrhs, b, SEG_V64_D1, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
        for (void *p : {(void *)dt, (void *)ds, (void *)dtl, (void *)drows, (void *)Kr})
            if (p) (void)hipFree(p);
        throw;
    }
    for (void *p : {(void *)dt, (void *)ds, (void *)dtl, (void *)drows, (void *)Kr}) HIPCHK(hipFree(p));
}

static void dense_rows(mx_plan *P, const int32_t *rows, int nrows, const double *W, int nrhs, double *out,
                       hipStream_t st) {
    if (!P->X_leaf) throw MxError(MX_ERR_STATE, "mx_plan_set_points has not been called");
    dense_rows_core(P->X_leaf, P->N, P->pts_d, P->pts_h, P->nleaf, P->wrow.data(), P->Ddim.data(), rows, nrows, W,
                    nrhs, out, st);
}

// ---- evaluation without allocation (mx_eval_noalloc) ----------------------
//
// Every work buffer of the forward evaluation is carved out of one
// file-scope device array of MX_NOALLOC_SCRATCH_BYTES.  Each GPU the library
// runs on holds its own copy until the process ends; its address on the
// current device comes from hipGetSymbolAddress, taken once per device on the
// first call there (host code never uses the array's name as an address).
// Calls on one GPU share the array: calls that overlap in time race.  No
// part of the array needs clearing: every skeleton block and panel row
// (padding included) is written before it is read in every evaluation.

static __device__ __align__(256) unsigned char scratch[MX_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;   // device -> address of its scratch

static void *scratch_base(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        throw MxError(MX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
    return ptr;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Work buffers of a forward evaluation laid out in the scratch area; returns
// the bytes needed (nothing is written when base is null).
static size_t noalloc_layout(const mx_plan *P, mx_precision prec, mx_io_type io, int nrhs, unsigned char *base,
                             Work *wk) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> void * {
        void *p = base ? base + off : nullptr;
        off += round256(std::max<size_t>(bytes, 1));
        return p;
    };
    if (prec == MX_BF16X3) {
        wk->panels = (uint16_t *)take(panel_bytes_of(P, nrhs));
        wk->sweep_ctl = (int32_t *)take((size_t)sweep_ctl_ints(P->numnodes, nrhs) * sizeof(int32_t));
    } else {
        const size_t elem = prec == MX_FP64 ? sizeof(double) : sizeof(float);
        const size_t sb = (size_t)std::max<int64_t>(P->skel_rows, 1) * nrhs * elem;
        wk->wsk = take(sb);
        wk->usk = take(sb);
        if (prec == MX_FP32 && io == MX_IO_F64) {
            const size_t ib = (size_t)std::max<int64_t>(P->N, 1) * nrhs * sizeof(float);
            wk->io_w = take(ib);
            wk->io_y = take(ib);
        }
    }
    return off;
}

static void eval_noalloc(mx_plan *P, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
                         hipStream_t st) {
    eval_checks(P, DIR_FWD, prec, W, Y, nrhs, io);
    Work probe;
    const size_t need = noalloc_layout(P, prec, io, nrhs, nullptr, &probe);
    if (need > MX_NOALLOC_SCRATCH_BYTES)
        throw MxError(MX_ERR_ARG, "mx_eval_noalloc: this evaluation needs " + std::to_string(need) +
                                      " bytes of work space; the static scratch holds " +
                                      std::to_string((size_t)MX_NOALLOC_SCRATCH_BYTES));
    int dev;
    HIPCHK(hipGetDevice(&dev));
    Work wk;
    noalloc_layout(P, prec, io, nrhs, static_cast<unsigned char *>(scratch_base(dev)), &wk);
    wk.prof = false;
    eval_core(P, DIR_FWD, prec, W, Y, nrhs, io, st, wk);
}

}  // namespace mx

using namespace mx;

extern "C" {

mx_status mx_plan_create(mx_plan **plan, const mx_structure *s, uint32_t precisions) {
    return (mx_status)guard([&]() {
        arg_check(plan && s, "null argument");
        *plan = nullptr;
        *plan = create_plan(*s, precisions, 0, 0);
        return 0;
    });
}

mx_status mx_plan_create_generated(mx_plan **plan, const mx_structure *s, uint32_t precisions, const double *X,
                                   int64_t n, int32_t d, double h, const int32_t *lids, const int32_t *skels,
                                   const int64_t *skelsoffset, int64_t nskels) {
    return (mx_status)guard([&]() {
        arg_check(plan && s, "null argument");
        *plan = nullptr;
        GenInfo g;
        g.X = X;
        g.n = n;
        g.d = d;
        g.h = h;
        g.lids = lids;
        g.skels = skels;
        g.skoff = skelsoffset;
        g.nskels = nskels;
        *plan = create_plan(*s, precisions, 0, 0, &g);
        return 0;
    });
}

mx_status mx_plan_destroy(mx_plan *plan) {
    delete plan;
    return MX_OK;
}

mx_status mx_eval(mx_plan *plan, mx_precision prec, const void *W, void *Y, int32_t nrhs, mx_io_type io,
                  void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
        arg_check(W && Y && nrhs >= 1, "bad W, Y or nrhs");
        if (plan->use_graph && !plan->profiling) eval_graph(plan, DIR_FWD, prec, W, Y, nrhs, io, (hipStream_t)stream);
        else eval(plan, DIR_FWD, prec, W, Y, nrhs, io, (hipStream_t)stream);
        return 0;
    });
}

mx_status mx_eval_noalloc(mx_plan *plan, mx_precision prec, const void *W, void *Y, int32_t nrhs, mx_io_type io,
                          void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
        eval_noalloc(plan, prec, W, Y, nrhs, io, (hipStream_t)stream);
        return 0;
    });
}

mx_status mx_eval_adjoint(mx_plan *plan, mx_precision prec, const void *dY, void *dW, int32_t nrhs, mx_io_type io,
                          void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
        arg_check(dY && dW && nrhs >= 1, "bad dY, dW or nrhs");
        if (plan->use_graph && !plan->profiling) eval_graph(plan, DIR_ADJ, prec, dY, dW, nrhs, io, (hipStream_t)stream);
        else eval(plan, DIR_ADJ, prec, dY, dW, nrhs, io, (hipStream_t)stream);
        return 0;
    });
}

mx_status mx_backward(mx_plan *plan, mx_precision prec, const void *W, const void *dY, void *dW, int32_t nrhs,
                      mx_io_type io, const void *wskel, const void *uskel, void *dD, void *dB, void *dVT,
                      void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr && nrhs >= 1, "null plan or bad nrhs");
        backward(plan, prec, W, dY, dW, nrhs, io, wskel, uskel, dD, dB, dVT, (hipStream_t)stream);
        return 0;
    });
}

mx_status mx_plan_set_blocks(mx_plan *plan, const double *D, const double *B, const double *VT, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        set_blocks(plan, D, B, VT, (hipStream_t)stream);
        return 0;
    });
}

mx_status mx_plan_grad_flops(const mx_plan *plan, int32_t nrhs, double *flops) {
    return (mx_status)guard([&]() {
        arg_check(plan && flops && nrhs >= 0, "bad argument");
        if (!plan->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
        *flops = plan->g_flops * nrhs;
        return 0;
    });
}

mx_status mx_plan_set_grad_tile(mx_plan *plan, int32_t tile) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        arg_check(tile == 0 || tile == GRAD_X3_TILE || tile == GRAD_X3_WTILE, "gradient tile: 0, 128 or 256");
        plan->g_x3_tile = tile;
        return 0;
    });
}

mx_status mx_plan_get_grad_tile(const mx_plan *plan, int32_t *tile) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr && tile != nullptr, "null argument");
        *tile = plan->g_x3_tile == 0 ? (plan->g_x3w_auto ? GRAD_X3_WTILE : GRAD_X3_TILE) : plan->g_x3_tile;
        return 0;
    });
}

mx_status mx_plan_skeletons(mx_plan *plan, const void **wskel, const void **uskel) {
    return (mx_status)guard([&]() {
        arg_check(plan && wskel && uskel, "null argument");
        if (plan->last_prec < 0) throw MxError(MX_ERR_STATE, "no evaluation has run on this plan");
        if (plan->last_prec == MX_BF16X3)
            throw MxError(MX_ERR_UNSUPPORTED, "split-BF16x3 evaluations keep skeletons as split panels");
        *wskel = plan->wskel;
        *uskel = plan->uskel;
        return 0;
    });
}

mx_status mx_plan_copy_skeletons(mx_plan *plan, void *wskel_dst, void *uskel_dst, int32_t nrhs, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan && wskel_dst && uskel_dst && nrhs >= 1, "bad argument");
        if (plan->last_prec < 0) throw MxError(MX_ERR_STATE, "no evaluation has run on this plan");
        if (plan->last_prec == MX_BF16X3) {   // FP32 copies of the split panels
            export_panels(plan, DIR_FWD, nrhs, (float *)wskel_dst, (float *)uskel_dst, (hipStream_t)stream);
            return 0;
        }
        if (plan->last_nrhs != nrhs)
            throw MxError(MX_ERR_STATE, "the last forward evaluation had another number of columns");
        size_t elem = plan->last_prec == MX_FP64 ? sizeof(double) : sizeof(float);
        size_t bytes = (size_t)plan->skel_rows * nrhs * elem;
        HIPCHK(hipMemcpyAsync(wskel_dst, plan->wskel, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
        HIPCHK(hipMemcpyAsync(uskel_dst, plan->uskel, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
        return 0;
    });
}

mx_status mx_plan_set_points(mx_plan *plan, const double *X_leaf, int32_t d, double h) {
    return (mx_status)guard([&]() {
        arg_check(plan && X_leaf && d >= 1 && h > 0, "bad argument");
        if (plan->X_leaf) HIPCHK(hipFree(plan->X_leaf));
        plan->X_leaf = nullptr;
        plan->X_leaf = upload(X_leaf, (size_t)plan->N * d);
        plan->pts_d = d;
        plan->pts_h = h;
        return 0;
    });
}

mx_status mx_dense_rows(mx_plan *plan, const int32_t *rows, int32_t nrows, const double *W, int32_t nrhs,
                        double *out, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        dense_rows(plan, rows, nrows, W, nrhs, out, (hipStream_t)stream);
        return 0;
    });
}

mx_status mx_plan_set_graph(mx_plan *plan, int32_t enable) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        plan->use_graph = enable ? 1 : 0;
        if (!enable) drop_graphs(plan);
        return 0;
    });
}

mx_status mx_plan_set_schedule(mx_plan *plan, int32_t direction, mx_precision prec, int32_t schedule,
                               int32_t x3_variant) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        set_schedule(plan, direction, prec, schedule, x3_variant);
        return 0;
    });
}

mx_status mx_lists_describe(const mx_structure *s, int32_t direction, int32_t schedule, int32_t *nlaunch,
                            int32_t *nltask, int32_t *ntask, int32_t *nseg, int32_t *ncomp, int64_t *launches,
                            int64_t *ltasks, int64_t *tasks, int64_t *segs, int64_t *comps) {
    return (mx_status)guard([&]() {
        arg_check(s && nlaunch && nltask && ntask && nseg && ncomp && (direction == 0 || direction == 1),
                  "bad argument");
        arg_check((schedule >= 0 && schedule < SCHED_COARSE) || schedule == MX_SCHED_COARSE ||
                      schedule == MX_SCHED_COARSE_SPLIT,
                  "bad schedule");
        const int sc = schedule == MX_SCHED_COARSE ? SCHED_COARSE
                       : schedule == MX_SCHED_COARSE_SPLIT ? SCHED_COARSE_SPLIT : schedule;
        Topology t = topology(*s);
        validate(*s, t);
        std::vector<int64_t> wrow(s->wrow, s->wrow + s->nleaf), skel(s->skelrow, s->skelrow + s->numnodes);
        TaskLists X = build_lists(*s, t, 0, 0, 1, direction, sc, wrow.data(), skel.data());
        // the tasks of each launch, once each, in tile order
        std::vector<int64_t> off{0}, lt;
        std::vector<int> seen(X.tasks.size(), -1);
        for (size_t li = 0; li < X.launches.size(); ++li) {
            for (int ti = X.launches[li].tile_offset; ti < X.launches[li].tile_offset + X.launches[li].ntiles; ++ti) {
                const int task = X.tiles[ti].x;
                if (seen[task] == (int)li || X.tasks[task].m == 0) continue;
                seen[task] = (int)li;
                lt.push_back(task);
            }
            off.push_back((int64_t)lt.size());
        }
        const bool fill = launches != nullptr;
        arg_check(!fill || (*nlaunch >= (int32_t)X.launches.size() && *nltask >= (int32_t)lt.size() &&
                            *ntask >= (int32_t)X.tasks.size() && *nseg >= (int32_t)X.segs.size() &&
                            *ncomp >= (int32_t)X.comps.size() && ltasks && tasks && segs && (comps || X.comps.empty())),
                  "arrays too small");
        *nlaunch = (int32_t)X.launches.size();
        *nltask = (int32_t)lt.size();
        *ntask = (int32_t)X.tasks.size();
        *nseg = (int32_t)X.segs.size();
        *ncomp = (int32_t)X.comps.size();
        if (!fill) return 0;
        std::copy(off.begin(), off.end(), launches);
        std::copy(lt.begin(), lt.end(), ltasks);
        for (size_t i = 0; i < X.tasks.size(); ++i) {
            const Task &k = X.tasks[i];
            const int64_t v[6] = {k.c_row, k.m, k.c_buf, k.seg_begin, k.seg_end, k.beta};
            std::copy(v, v + 6, tasks + 6 * i);
        }
        for (size_t i = 0; i < X.segs.size(); ++i) {
            const Seg &g = X.segs[i];
            const int64_t v[8] = {g.a_off, g.b_row, g.a_buf, g.a_trans, g.lda, g.k, g.b_buf, g.b_ld};
            std::copy(v, v + 8, segs + 8 * i);
        }
        for (size_t i = 0; i < X.comps.size(); ++i) {
            const Composite &c = X.comps[i];
            const int64_t v[11] = {c.a1_off, c.a2_off, c.a1_buf, c.a1_trans, c.a1_lda, c.a2_buf,
                                   c.a2_trans, c.a2_lda, c.m, c.kmid, c.n};
            std::copy(v, v + 11, comps + 11 * i);
        }
        return 0;
    });
}

mx_status mx_sweep_describe(const mx_structure *s, int32_t direction, int32_t *ntiles, int32_t *ntasks,
                            int32_t *nwaits, int32_t *tiles, int32_t *qinfo, int32_t *tasks, int32_t *waits) {
    return (mx_status)guard([&]() {
        arg_check(s && ntiles && ntasks && nwaits && (direction == 0 || direction == 1), "bad argument");
        Topology t = topology(*s);
        validate(*s, t);
        std::vector<int64_t> wprow, sprow;
        int64_t wp = 0, sp = 0;
        panel_rows(*s, wprow, sprow, &wp, &sp);
        TaskLists X = build_lists(*s, t, 0, 0, 1, direction, SCHED_FUSED, wprow.data(), sprow.data());
        split_long_tasks(X, X3_MAXSEG);
        SweepLists SW = build_sweep(X, s->numnodes, sprow);
        if (!SW.ok) throw MxError(MX_ERR_UNSUPPORTED, "no sweep for this structure (a block with several writers)");
        const bool fill = tiles != nullptr;
        arg_check(!fill || (*ntiles >= (int32_t)SW.tiles.size() && *ntasks >= (int32_t)SW.stasks.size() &&
                            *nwaits >= (int32_t)SW.waits.size() && qinfo && tasks && waits),
                  "arrays too small");
        *ntiles = (int32_t)SW.tiles.size();
        *ntasks = (int32_t)SW.stasks.size();
        *nwaits = (int32_t)SW.waits.size();
        if (!fill) return 0;
        static_assert(sizeof(SweepTile) == 12 * sizeof(int32_t) && sizeof(SweepTask) == 8 * sizeof(int32_t),
                      "descriptor layout");
        std::memcpy(tiles, SW.tiles.data(), SW.tiles.size() * sizeof(SweepTile));
        std::memcpy(tasks, SW.stasks.data(), SW.stasks.size() * sizeof(SweepTask));
        for (size_t i = 0; i < SW.waits.size(); ++i) {
            waits[2 * i] = SW.waits[i].block;
            waits[2 * i + 1] = SW.waits[i].target;
        }
        for (int z = 0; z < SWEEP_NQ; ++z) {
            qinfo[2 * z] = SW.qinfo[z].x;
            qinfo[2 * z + 1] = SW.qinfo[z].y;
        }
        return 0;
    });
}

mx_status mx_plan_sweep_status(const mx_plan *plan, void *stream, int32_t *error) {
    return (mx_status)guard([&]() {
        arg_check(plan && error, "null argument");
        *error = 0;
        HIPCHK(hipStreamSynchronize((hipStream_t)stream));
        if (plan->sweep_last_ctl)
            HIPCHK(hipMemcpy(error, plan->sweep_last_ctl + SWEEP_ERR, sizeof(int32_t), hipMemcpyDeviceToHost));
        return 0;
    });
}

mx_status mx_plan_get_schedule(const mx_plan *plan, int32_t direction, mx_precision prec, int32_t *schedule,
                               int32_t *x3_variant) {
    return (mx_status)guard([&]() {
        arg_check(plan && schedule && x3_variant && (direction == 0 || direction == 1), "bad argument");
        const bool sw = prec == MX_BF16X3 && plan->sweep_on[direction];
        const int sc = plan->sched[direction][prec == MX_BF16X3 ? 1 : 0];
        *schedule = sw ? MX_SCHED_SWEEP
                    : sc == SCHED_COARSE ? MX_SCHED_COARSE
                    : sc == SCHED_COARSE_SPLIT ? MX_SCHED_COARSE_SPLIT : sc;
        *x3_variant = sw ? plan->sweep_var[direction] : plan->x3_variant;
        return 0;
    });
}

mx_status mx_plan_launch_variants(const mx_plan *plan, int32_t direction, int32_t *variants, int32_t cap,
                                  int32_t *count) {
    return (mx_status)guard([&]() {
        arg_check(plan && count && (direction == 0 || direction == 1) && cap >= 0 && (cap == 0 || variants),
                  "bad argument");
        const int sc = plan->sched[direction][1];
        const std::vector<Launch> &ln = plan->prog[direction][1][sc].launches;
        const std::vector<int> &lv = plan->lvar[direction][sc];
        *count = (int32_t)ln.size();
        for (int32_t i = 0; i < (int32_t)ln.size() && i < cap; ++i)
            variants[i] = lv.size() == ln.size() ? lv[i] : plan->x3_variant;
        return 0;
    });
}

mx_status mx_plan_set_launch_variants(mx_plan *plan, int32_t direction, const int32_t *variants, int32_t count) {
    return (mx_status)guard([&]() {
        arg_check(plan && (direction == 0 || direction == 1), "bad argument");
        if (direction == DIR_ADJ && !plan->adjoint)
            throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
        const int sc = plan->sched[direction][1];
        std::vector<int> &lv = plan->lvar[direction][sc];
        if (count == 0) {
            lv.clear();
        } else {
            arg_check(variants != nullptr, "null variants");
            arg_check(count == (int32_t)plan->prog[direction][1][sc].launches.size(),
                      "count must equal the number of launches of the current schedule");
            for (int32_t i = 0; i < count; ++i)
                arg_check(variants[i] >= 1 && variants[i] <= X3_NVARIANT, "bad BF16x3 kernel variant");
            lv.assign(variants, variants + count);
        }
        drop_graphs(plan);
        return 0;
    });
}

mx_status mx_plan_seg_variants(const mx_plan *plan, int32_t direction, mx_precision prec, int32_t *variants,
                               int32_t cap, int32_t *count) {
    return (mx_status)guard([&]() {
        arg_check(plan && count && (direction == 0 || direction == 1) && cap >= 0 && (cap == 0 || variants),
                  "bad argument");
        arg_check(prec == MX_FP64 || prec == MX_FP32, "FP64 or FP32 only");
        const int pk = prec == MX_FP64 ? 0 : 1;
        const int sc = plan->sched[direction][0];
        const size_t nl = plan->prog[direction][0][sc].launches.size();
        *count = (int32_t)nl;
        for (int32_t i = 0; i < (int32_t)nl && i < cap; ++i) variants[i] = seg_variant(plan, direction, pk, i);
        return 0;
    });
}

mx_status mx_plan_set_seg_variants(mx_plan *plan, int32_t direction, mx_precision prec, const int32_t *variants,
                                   int32_t count) {
    return (mx_status)guard([&]() {
        arg_check(plan && (direction == 0 || direction == 1), "bad argument");
        arg_check(prec == MX_FP64 || prec == MX_FP32, "FP64 or FP32 only");
        if (direction == DIR_ADJ && !plan->adjoint)
            throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
        const int pk = prec == MX_FP64 ? 0 : 1;
        const int sc = plan->sched[direction][0];
        std::vector<int> &lv = plan->lsg[direction][pk][sc];
        if (count == 0) {
            lv.clear();
        } else {
            arg_check(variants != nullptr, "null variants");
            arg_check(count == (int32_t)plan->prog[direction][0][sc].launches.size(),
                      "count must equal the number of launches of the current schedule");
            for (int32_t i = 0; i < count; ++i)
                arg_check(seg_variant_ok(prec, variants[i]), "bad FP64/FP32 kernel variant");
            lv.assign(variants, variants + count);
        }
        drop_graphs(plan);
        return 0;
    });
}

mx_status mx_plan_autotune(mx_plan *plan, mx_precision prec, int32_t nrhs, mx_io_type io, int32_t reps,
                           float *best_ms, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        arg_check(io == MX_IO_F64 || io == MX_IO_F32, "bad io type");
        if (prec == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 evaluation needs FP64 I/O");
        autotune(plan, prec, nrhs, io, reps, (hipStream_t)stream, best_ms);
        return 0;
    });
}

mx_status mx_plan_set_profiling(mx_plan *plan, int32_t enable) {
    return (mx_status)guard([&]() {
        arg_check(plan != nullptr, "null plan");
        plan->profiling = enable ? 1 : 0;
        plan->ev_used = 0;
        return 0;
    });
}

mx_status mx_plan_launch_times(mx_plan *plan, float *ms, int32_t *kind, int32_t cap, int32_t *count) {
    return (mx_status)guard([&]() {
        arg_check(plan && count && cap >= 0 && (cap == 0 || (ms && kind)), "bad argument");
        *count = plan->ev_used;
        if (plan->ev_used == 0) return 0;
        HIPCHK(hipEventSynchronize(plan->ev[2 * plan->ev_used - 1]));
        for (int i = 0; i < plan->ev_used && i < cap; ++i) {
            HIPCHK(hipEventElapsedTime(&ms[i], plan->ev[2 * i], plan->ev[2 * i + 1]));
            kind[i] = plan->ev_kind[i];
        }
        return 0;
    });
}

mx_status mx_plan_flops(const mx_plan *plan, mx_precision prec, int32_t nrhs, double *flops, double *total) {
    return (mx_status)guard([&]() {
        arg_check(plan && flops && total && nrhs >= 0, "bad argument");
        for (int i = 0; i < 4; ++i) flops[i] = 0.0;
        if (prec == MX_BF16X3) {
            // algorithmic flops: the coarse schedule's composites add work the
            // reference's evaluation does not have, so its total is the fused
            // lists' (the same for every other schedule)
            const int sc = sched_coarse(plan->sched[DIR_FWD][1]) ? SCHED_FUSED : plan->sched[DIR_FWD][1];
            const Program &pg = plan->prog[DIR_FWD][1][sc];
            if (!pg.built) throw MxError(MX_ERR_UNSUPPORTED, "no BF16x3 lists in the plan");
            for (double f : pg.flops) flops[MX_KIND_BF16X3] += f * nrhs;
        } else {
            int k = prec == MX_FP64 ? MX_KIND_SEG_F64 : MX_KIND_SEG_F32;
            for (double f : plan->prog[DIR_FWD][0][plan->sched[DIR_FWD][0]].flops) flops[k] += f * nrhs;
        }
        *total = flops[0] + flops[1] + flops[2] + flops[3];
        return 0;
    });
}

mx_status mx_plan_launch_flops(const mx_plan *plan, mx_precision prec, int32_t nrhs, double *flops, int32_t cap,
                               int32_t *count) {
    return (mx_status)guard([&]() {
        arg_check(plan && count && cap >= 0 && (cap == 0 || flops), "bad argument");
        const int kd = prec == MX_BF16X3 ? 1 : 0;
        const std::vector<double> &f = plan->prog[DIR_FWD][kd][plan->sched[DIR_FWD][kd]].flops;
        *count = (int32_t)f.size();
        for (int32_t i = 0; i < (int32_t)f.size() && i < cap; ++i) flops[i] = f[i] * nrhs;
        return 0;
    });
}

mx_status mx_plan_dims(const mx_plan *plan, int64_t *N, int64_t *skel_rows) {
    return (mx_status)guard([&]() {
        arg_check(plan && N && skel_rows, "null argument");
        *N = plan->N;
        *skel_rows = plan->skel_rows;
        return 0;
    });
}

const char *mx_last_error(void) { return g_err.c_str(); }

}  // extern "C"
