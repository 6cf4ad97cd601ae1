// Many independent instances evaluated together (mx_batch, C-ABI in
// include/matrox_hgemm.h).
#include "mx_internal.h"

#include <cmath>

namespace mx {

// ---- many independent instances (mx_batch) -------------------------------
//
// Instances are grouped by nrhs.  The instances of a group become one forest
// joined under rank-0 nodes into a single binary tree: a dummy node has no
// skeleton, so it takes part in no product, and every launch of the group's
// plan covers the same tree level (or phase) of all its instances at once.
// HSS instances are rewritten with explicit pair lists (near (l, l) at
// Dptr[l], coupling (v, sibling) at Bptr[v-1]).  The groups run concurrently
// on their own streams, forked from and joined back to the caller's stream,
// and the whole batch can be captured as one hipGraph.

struct BatchInst {
    int group;
    int64_t row_off;                 // first row in the group's W / Y
    int64_t d_off = 0, b_off = 0, vt_off = 0;   // first element of its D, B, VT in the group's blocks
    int64_t N;
    int nleaf;
    std::vector<int32_t> wrow, Ddim; // instance-local leaf layout
    double *X_leaf = nullptr;        // points for dense check rows
    int pts_d = 0;
    double pts_h = 0.0;
};

// Combined structure of a group; owns its arrays.
struct GroupStructure {
    std::vector<int32_t> lc, rc, lm, Ddim, slen, npx, npy, fpx, fpy, VTptr, wrow, skelrow;
    std::vector<uint64_t> Dptr, Bptr;
    std::vector<double> D, B, VT;
    std::vector<int64_t> dbase, bbase, vbase;   // per member: first element of its D, B, VT
    int64_t N = 0;
    mx_structure view() {
        mx_structure s{};
        s.nleaf = (int32_t)Ddim.size();
        s.numnodes = (int32_t)lc.size();
        s.is_hss = 0;
        s.ncount = (int64_t)npx.size();
        s.fcount = (int64_t)fpx.size();
        s.lchildren = lc.data();
        s.rchildren = rc.data();
        s.lm = lm.data();
        s.Ddim = Ddim.data();
        s.slen = slen.data();
        s.npairx = npx.data();
        s.npairy = npy.data();
        s.fpairx = fpx.data();
        s.fpairy = fpy.data();
        s.Dptr = Dptr.data();
        s.Bptr = Bptr.data();
        s.VTptr = VTptr.data();
        s.wrow = wrow.data();
        s.skelrow = skelrow.data();
        s.D = D.data();
        s.B = B.data();
        s.VT = VT.data();
        s.D_len = (int64_t)D.size();
        s.B_len = (int64_t)B.size();
        s.VT_len = (int64_t)VT.size();
        return s;
    }
};

static int64_t struct_rows(const mx_structure &s) {
    int64_t n = 0;
    for (int l = 0; l < s.nleaf; ++l) n = std::max<int64_t>(n, (int64_t)s.wrow[l] + s.Ddim[l]);
    return n;
}

static void combine(const std::vector<const mx_structure *> &in, GroupStructure &G) {
    const int k = (int)in.size();
    int64_t nn = k - 1;
    for (auto *s : in) nn += s->numnodes;
    G.lc.assign(nn, -1);
    G.rc.assign(nn, -1);
    G.lm.assign(nn, -1);
    G.slen.assign(nn, 0);
    G.skelrow.assign(nn, 0);
    G.VTptr.assign(nn, 0);
    std::vector<int64_t> nbase(k), lbase(k), rbase(k), sbase(k), vbase(k), dbase(k), bbase(k);
    int64_t nb = k - 1, lb = 0, rb = 0, sb = 0, vb = 0, db = 0, bb = 0;
    for (int i = 0; i < k; ++i) {
        const mx_structure &s = *in[i];
        nbase[i] = nb;
        lbase[i] = lb;
        rbase[i] = rb;
        sbase[i] = sb;
        vbase[i] = vb;
        dbase[i] = db;
        bbase[i] = bb;
        nb += s.numnodes;
        lb += s.nleaf;
        rb += struct_rows(s);
        int64_t sk = 0;
        for (int v = 0; v < s.numnodes; ++v) sk = std::max<int64_t>(sk, (int64_t)s.skelrow[v] + s.slen[v]);
        sb += sk;
        vb += s.VT_len;
        db += s.D_len;
        bb += s.B_len;
    }
    arg_check(rb < INT32_MAX && sb < INT32_MAX && vb < INT32_MAX, "batch group too large for 32-bit offsets");
    G.N = rb;
    G.dbase = dbase;
    G.bbase = bbase;
    G.vbase = vbase;
    // dummy nodes 0..k-2 in heap order; child ids >= k-1 are instance roots
    for (int j = 0; j < k - 1; ++j) {
        int c[2] = {2 * j + 1, 2 * j + 2};
        int64_t id[2];
        for (int t = 0; t < 2; ++t) id[t] = c[t] < k - 1 ? c[t] : nbase[c[t] - (k - 1)];
        G.lc[j] = (int32_t)id[0];
        G.rc[j] = (int32_t)id[1];
    }
    G.Ddim.reserve(lb);
    G.wrow.reserve(lb);
    G.D.reserve(db);
    G.B.reserve(bb);
    G.VT.reserve(vb);
    for (int i = 0; i < k; ++i) {
        const mx_structure &s = *in[i];
        const int64_t o = nbase[i];
        std::vector<int> parent(s.numnodes, -1);
        for (int v = 0; v < s.numnodes; ++v) {
            if (s.lchildren[v] >= 0) {
                parent[s.lchildren[v]] = v;
                parent[s.rchildren[v]] = v;
                G.lc[o + v] = (int32_t)(s.lchildren[v] + o);
                G.rc[o + v] = (int32_t)(s.rchildren[v] + o);
            } else {
                G.lm[o + v] = (int32_t)(s.lm[v] + lbase[i]);
            }
            G.slen[o + v] = s.slen[v];
            G.skelrow[o + v] = (int32_t)(s.skelrow[v] + sbase[i]);
            G.VTptr[o + v] = (int32_t)(s.VTptr[v] + vbase[i]);
        }
        for (int l = 0; l < s.nleaf; ++l) {
            G.Ddim.push_back(s.Ddim[l]);
            G.wrow.push_back((int32_t)(s.wrow[l] + rbase[i]));
        }
        for (int64_t q = 0; q < s.ncount; ++q) {
            G.npx.push_back((int32_t)((s.is_hss ? q : s.npairx[q]) + lbase[i]));
            G.npy.push_back((int32_t)((s.is_hss ? q : s.npairy[q]) + lbase[i]));
            G.Dptr.push_back(s.Dptr[q] + (uint64_t)dbase[i]);
        }
        for (int64_t q = 0; q < s.fcount; ++q) {
            int x, y;
            if (s.is_hss) {
                x = (int)q + 1;
                arg_check(x < s.numnodes && parent[x] >= 0, "HSS coupling list longer than the tree");
                int p = parent[x];
                y = s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
            } else {
                x = s.fpairx[q];
                y = s.fpairy[q];
            }
            G.fpx.push_back((int32_t)(x + o));
            G.fpy.push_back((int32_t)(y + o));
            G.Bptr.push_back(s.Bptr[q] + (uint64_t)bbase[i]);
        }
        if (s.D_len) G.D.insert(G.D.end(), s.D, s.D + s.D_len);
        if (s.B_len) G.B.insert(G.B.end(), s.B, s.B + s.B_len);
        if (s.VT_len) G.VT.insert(G.VT.end(), s.VT, s.VT + s.VT_len);
    }
}

}  // namespace mx

struct mx_batch {
    std::vector<mx_plan *> plans;          // one per group
    std::vector<int> group_nrhs;
    // groups whose instances the split-BF16x3 accuracy guard moved: they
    // evaluate in FP32 when BF16x3 is requested (their plans hold no split
    // blocks)
    std::vector<int> group_fp32;
    std::vector<double> x3_err;            // per instance: estimated BF16x3 error (-1: not estimated)
    std::vector<int64_t> group_rows;
    std::vector<hipStream_t> streams;      // one per group (when there are several)
    std::vector<hipEvent_t> events;        // fork + one join event per group
    std::vector<mx::BatchInst> inst;
    int use_graph = 0;
    struct GraphEntry {
        int dir, prec, io;
        std::vector<const void *> W;
        std::vector<void *> Y;
        hipGraphExec_t exec;
        uint64_t used;
        uint64_t gen;   // batch_gen at capture
    };
    std::vector<GraphEntry> graphs;
    uint64_t clock = 0;
    // Per direction: the BF16x3 groups evaluated together as one
    // dependency-driven launch (sweep_var = its tile width, 0 = off: every
    // group runs its own plan's launches).  The merged queues are built for
    // one tile width; a group table (operands, panels, Y of every group) is
    // kept per (Y pointers, I/O type, buffer generation).
    int sweep_var[2] = {0, 0};
    struct BSweep {
        int var = 0;                     // tile width the lists were built for (0: none)
        std::vector<int> groups;         // the batch group of each sweep group
        std::vector<int> cbase, ncols;   // per sweep group: first counter, column tiles
        mx::SweepTile *q = nullptr;
        int2 *qinfo = nullptr;
        int64_t items = 0;
        int ctl_ints = 0;
        int32_t *ctl = nullptr;
        struct Table {
            std::vector<void *> Y;
            int io;
            uint64_t gen;
            mx::SweepGroup *d;
        };
        std::vector<Table> tables;
        void release() {
            for (void *p : {(void *)q, (void *)qinfo, (void *)ctl})
                if (p) (void)hipFree(p);
            for (auto &t : tables) (void)hipFree(t.d);
            q = nullptr;
            qinfo = nullptr;
            ctl = nullptr;
            tables.clear();
            groups.clear();
            cbase.clear();
            ncols.clear();
            items = 0;
            ctl_ints = 0;
            var = 0;
        }
    } bsweep[2];
    ~mx_batch() {
        for (auto &b : bsweep) b.release();
        for (auto *p : plans) delete p;
        for (auto st : streams) (void)hipStreamDestroy(st);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        for (auto &i : inst)
            if (i.X_leaf) (void)hipFree(i.X_leaf);
    }
};

namespace mx {

// Instances grouped by (nrhs, fp32[i]); a group of fp32 instances is built
// without the split-BF16x3 blocks.
static mx_batch *create_batch(const mx_structure *const *in, const int32_t *nrhs, int count, uint32_t precisions,
                              const std::vector<char> &fp32) {
    arg_check(in && nrhs && count >= 1, "bad batch arguments");
    std::unique_ptr<mx_batch> B(new mx_batch());
    std::map<std::pair<int, int>, std::vector<int>> by_key;
    for (int i = 0; i < count; ++i) {
        arg_check(in[i] != nullptr && nrhs[i] >= 1, "instance " + std::to_string(i) + ": null structure or bad nrhs");
        by_key[{nrhs[i], fp32[i] ? 1 : 0}].push_back(i);
    }
    B->inst.resize(count);
    B->x3_err.assign(count, -1.0);
    for (auto &kv : by_key) {
        const int g = (int)B->plans.size();
        std::vector<const mx_structure *> members;
        int64_t row = 0;
        for (int i : kv.second) {
            const mx_structure &s = *in[i];
            arg_check(s.nleaf > 0 && s.numnodes == 2 * s.nleaf - 1 && s.wrow && s.Ddim,
                      "instance " + std::to_string(i) + ": bad tree");
            BatchInst &bi = B->inst[i];
            bi.group = g;
            bi.row_off = row;
            bi.N = struct_rows(s);
            bi.nleaf = s.nleaf;
            bi.wrow.assign(s.wrow, s.wrow + s.nleaf);
            bi.Ddim.assign(s.Ddim, s.Ddim + s.nleaf);
            row += bi.N;
            members.push_back(&s);
        }
        GroupStructure G;
        combine(members, G);
        for (size_t m = 0; m < kv.second.size(); ++m) {
            BatchInst &bi = B->inst[kv.second[m]];
            bi.d_off = G.dbase[m];
            bi.b_off = G.bbase[m];
            bi.vt_off = G.vbase[m];
        }
        mx_structure v = G.view();
        const bool routed = kv.first.second != 0;
        B->plans.push_back(create_plan(v, routed ? (precisions & ~(uint32_t)MX_PLAN_BF16X3) : precisions, 0, 0));
        B->group_nrhs.push_back(kv.first.first);
        B->group_rows.push_back(G.N);
        B->group_fp32.push_back(routed ? 1 : 0);
    }
    const int ng = (int)B->plans.size();
    if (ng > 1) {
        B->streams.resize(ng, nullptr);
        for (auto &st : B->streams) HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        B->events.resize(ng + 1, nullptr);
        for (auto &e : B->events) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return B.release();
}

// The precision group g evaluates in when prec is requested.
static mx_precision group_prec(const mx_batch *B, int g, mx_precision prec) {
    return prec == MX_BF16X3 && B->group_fp32[g] ? MX_FP32 : prec;
}

// Per instance i of `ranges` (element ranges [b, e) of two group buffers):
// out[2i] += sum (x - y)^2, out[2i+1] += sum y^2.
__global__ void instance_err_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                    const int64_t *__restrict__ ranges, double *out) {
    const int64_t b = ranges[2 * blockIdx.y], e = ranges[2 * blockIdx.y + 1];
    double d2 = 0.0, y2 = 0.0;
    for (int64_t k = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < e; k += (int64_t)gridDim.x * blockDim.x) {
        const double t = x[k] - y[k];
        d2 += t * t;
        y2 += y[k] * y[k];
    }
    for (int o = 32; o > 0; o >>= 1) {
        d2 += __shfl_down(d2, o);
        y2 += __shfl_down(y2, o);
    }
    __shared__ double red[2][8];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        red[0][w] = d2;
        red[1][w] = y2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int i = 0; i < (int)(blockDim.x / 64); ++i) {
            a += red[0][i];
            c += red[1][i];
        }
        atomicAdd(&out[2 * blockIdx.y], a);
        atomicAdd(&out[2 * blockIdx.y + 1], c);
    }
}

// Estimated relative error of every instance's split-BF16x3 evaluation
// (forward, and adjoint when built): each group evaluated once on a
// unit-variance pseudo-random W in BF16x3 and in the reference precision
// (FP64 when built, else FP32, whose own error is ~1e-6 of that).
static std::vector<double> estimate_x3_error(mx_batch *B, hipStream_t st) {
    std::vector<double> err(B->inst.size(), 0.0);
    for (size_t g = 0; g < B->plans.size(); ++g) {
        mx_plan *P = B->plans[g];
        const int n = B->group_nrhs[g];
        const mx_precision ref = (P->precisions & MX_PLAN_FP64) ? MX_FP64 : MX_FP32;
        std::vector<int> members;
        std::vector<int64_t> ranges;
        for (size_t i = 0; i < B->inst.size(); ++i)
            if (B->inst[i].group == (int)g) {
                members.push_back((int)i);
                ranges.push_back(B->inst[i].row_off * n);
                ranges.push_back((B->inst[i].row_off + B->inst[i].N) * n);
            }
        const int64_t nel = std::max<int64_t>(B->group_rows[g], 1) * n;
        double *buf = nullptr, *acc = nullptr;
        int64_t *dr = nullptr;
        try {
            HIPCHK(hipMalloc(&buf, 3 * (size_t)nel * sizeof(double)));
            HIPCHK(hipMalloc(&acc, 2 * members.size() * sizeof(double)));
            dr = upload(ranges.data(), ranges.size());
            double *W = buf, *Y3 = buf + nel, *Yr = buf + 2 * nel;
            launch_fill_uniform(W, nel, true, st);
            for (int dir = 0; dir < (P->adjoint ? 2 : 1); ++dir) {
                eval(P, dir, MX_BF16X3, W, Y3, n, MX_IO_F64, st);
                eval(P, dir, ref, W, Yr, n, MX_IO_F64, st);
                HIPCHK(hipMemsetAsync(acc, 0, 2 * members.size() * sizeof(double), st));
                instance_err_kernel<<<dim3(16, (unsigned)members.size()), 256, 0, st>>>(Y3, Yr, dr, acc);
                HIPCHK(hipGetLastError());
                std::vector<double> h(2 * members.size());
                HIPCHK(hipMemcpyAsync(h.data(), acc, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                for (size_t m = 0; m < members.size(); ++m)
                    err[members[m]] = std::max(err[members[m]], h[2 * m + 1] > 0 ? std::sqrt(h[2 * m] / h[2 * m + 1])
                                                                                  : 0.0);
            }
        } catch (...) {
            for (void *p : {(void *)buf, (void *)acc, (void *)dr})
                if (p) (void)hipFree(p);
            throw;
        }
        for (void *p : {(void *)buf, (void *)acc, (void *)dr}) HIPCHK(hipFree(p));
    }
    return err;
}

// Build the batch; with BF16x3 and FP32 both requested (and the guard not
// switched off), estimate every instance's BF16x3 error and rebuild with the
// instances above MX_BATCH_X3_GUARD in FP32 groups.
static mx_batch *create_guarded_batch(const mx_structure *const *in, const int32_t *nrhs, int count,
                                      uint32_t precisions) {
    const bool guard = (precisions & MX_PLAN_BF16X3) && (precisions & MX_PLAN_FP32) &&
                       !(precisions & MX_BATCH_NO_X3_GUARD);
    precisions &= ~(uint32_t)MX_BATCH_NO_X3_GUARD;
    std::vector<char> fp32(std::max(count, 0), 0);
    std::unique_ptr<mx_batch> B(create_batch(in, nrhs, count, precisions, fp32));
    if (!guard) return B.release();
    const std::vector<double> err = estimate_x3_error(B.get(), nullptr);
    bool any = false;
    for (int i = 0; i < count; ++i) {
        fp32[i] = err[i] > MX_BATCH_X3_GUARD;
        any = any || fp32[i];
    }
    if (any) {
        B.reset();
        B.reset(create_batch(in, nrhs, count, precisions, fp32));
    }
    B->x3_err = err;
    return B.release();
}

static uint64_t batch_gen(const mx_batch *B);

// ---- batched sweep: every BF16x3 group in one launch ------------------------
//
// The groups' sweep lists (each its plan's, as for mx_eval) are merged into
// one set of eight queues.  Column tiles become explicit entries (row tile
// | column tile << 16, ncols = 1), since the groups' nrhs differ.  The
// groups' DAGs are independent, so any interleaving of their rank orders
// is a rank order of the union: items are ordered by their rank within
// their group divided by the group's item count, which spreads every group
// over the whole launch.  Counters: one block per group, after the others.

static bool batch_sweep_on(const mx_batch *B, int dir, mx_precision prec) {
    return prec == MX_BF16X3 && B->sweep_var[dir] != 0;
}

static void build_batch_sweep(mx_batch *B, int dir, int var) {
    mx_batch::BSweep &bs = B->bsweep[dir];
    bs.release();
    const int bn = sweep_bn(var);
    struct Item {
        double key;
        int grp, r, z, j, ct;
    };
    std::vector<std::vector<Item>> perq(SWEEP_NQ);
    std::vector<std::vector<std::vector<int>>> at;   // [sweep group][queue][tile * ncols + ct] -> merged index
    int cb = SWEEP_CNT;
    for (size_t g = 0; g < B->plans.size(); ++g) {
        const mx_plan *P = B->plans[g];
        if (B->group_fp32[g] || !(P->precisions & MX_PLAN_BF16X3)) continue;
        const mx_plan::Sweep &sw = P->sweep[dir];
        if (!sw.ok) throw MxError(MX_ERR_UNSUPPORTED, "batched sweep: a group without sweep lists");
        const int si = (int)bs.groups.size();
        const int nc = (B->group_nrhs[g] + bn - 1) / bn;
        bs.groups.push_back((int)g);
        bs.ncols.push_back(nc);
        bs.cbase.push_back(cb);
        cb += sw.nblocks * nc;
        int maxg = 0;
        for (const SweepTile &t : sw.h_tiles) maxg = std::max(maxg, t.g);
        const double R = (double)(maxg + 1) * nc;
        at.emplace_back(SWEEP_NQ);
        for (int z = 0; z < SWEEP_NQ; ++z) {
            const int2 qi = sw.qinfo_h[z];
            at[si][z].assign((size_t)qi.y * nc, -1);
            for (int j = 0; j < qi.y; ++j)
                for (int ct = 0; ct < nc; ++ct) {
                    const int r = sw.h_tiles[qi.x + j].g * nc + ct;
                    perq[z].push_back(Item{(r + 0.5) / R, si, r, z, j, ct});
                }
        }
    }
    if (bs.groups.empty()) throw MxError(MX_ERR_UNSUPPORTED, "batched sweep: no BF16x3 group");
    auto less = [](const Item &a, const Item &b) {
        if (a.key != b.key) return a.key < b.key;
        if (a.grp != b.grp) return a.grp < b.grp;
        return a.r < b.r;
    };
    std::vector<Item> all;
    for (int z = 0; z < SWEEP_NQ; ++z) {
        std::sort(perq[z].begin(), perq[z].end(), less);
        for (size_t p = 0; p < perq[z].size(); ++p) {
            const Item &it = perq[z][p];
            at[it.grp][z][(size_t)it.j * bs.ncols[it.grp] + it.ct] = (int)p;
        }
        all.insert(all.end(), perq[z].begin(), perq[z].end());
    }
    std::sort(all.begin(), all.end(), less);
    std::map<std::tuple<int, int, int, int>, int> rank;   // (group, queue, tile, ct) -> global rank
    for (size_t k = 0; k < all.size(); ++k) rank[std::make_tuple(all[k].grp, all[k].z, all[k].j, all[k].ct)] = (int)k;
    std::vector<SweepTile> tiles;
    std::vector<int2> qinfo;
    for (int z = 0; z < SWEEP_NQ; ++z) {
        qinfo.push_back(make_int2((int)tiles.size(), (int)perq[z].size()));
        for (const Item &it : perq[z]) {
            const mx_plan::Sweep &sw = B->plans[bs.groups[it.grp]]->sweep[dir];
            const SweepTile &src = sw.h_tiles[sw.qinfo_h[it.z].x + it.j];
            if (src.rt >= 65536 || it.ct >= 65536) throw MxError(MX_ERR_ARG, "batched sweep: too many tiles");
            SweepTile t{};
            t.task = src.task;
            t.rt = src.rt | (it.ct << 16);
            t.g = rank[std::make_tuple(it.grp, it.z, it.j, it.ct)];
            t.grp = it.grp;
            for (int y = 0; y < SWEEP_NQ; ++y) {
                t.maxtile[y] = src.maxtile[y] < 0 ? -1 : at[it.grp][y][(size_t)src.maxtile[y] * bs.ncols[it.grp] + it.ct];
                if (src.maxtile[y] >= 0 && (t.maxtile[y] < 0 || perq[y][t.maxtile[y]].key >= it.key))
                    throw MxError(MX_ERR_ARG, "internal: batched sweep producer ranked after its consumer");
            }
            tiles.push_back(t);
        }
    }
    bs.items = (int64_t)tiles.size();
    bs.ctl_ints = cb;
    bs.q = upload(tiles.data(), tiles.size());
    bs.qinfo = upload(qinfo.data(), qinfo.size());
    HIPCHK(hipMalloc(&bs.ctl, (size_t)cb * sizeof(int32_t)));
    // zero now: mx_batch_sweep_status reads the error word of both
    // directions, also of one whose sweep has not run yet
    HIPCHK(hipMemset(bs.ctl, 0, (size_t)cb * sizeof(int32_t)));
    bs.var = var;
}

// The group table for outputs Y (batch groups) and io, built (and uploaded)
// on first use; null when it does not exist yet and create is false.
static SweepGroup *batch_sweep_table(mx_batch *B, int dir, mx_io_type io, void *const *Y, bool create,
                                     uint64_t gen) {
    mx_batch::BSweep &bs = B->bsweep[dir];
    for (size_t i = 0; i < bs.tables.size();) {   // tables of moved buffers are stale
        if (bs.tables[i].gen != gen) {
            HIPCHK(hipFree(bs.tables[i].d));
            bs.tables.erase(bs.tables.begin() + (std::ptrdiff_t)i);
        } else {
            ++i;
        }
    }
    std::vector<void *> ys;
    for (int g : bs.groups) ys.push_back(Y[g]);
    for (auto &t : bs.tables)
        if (t.Y == ys && t.io == (int)io) return t.d;
    if (!create) return nullptr;
    std::vector<SweepGroup> h(bs.groups.size());
    for (size_t k = 0; k < bs.groups.size(); ++k) {
        const int g = bs.groups[k];
        mx_plan *P = B->plans[g];
        SweepGroup &G = h[k];
        G = SweepGroup{};
        x3_args_of(P, B->group_nrhs[g], P->panels, G.x);
        G.x.tasks = P->sweep[dir].tasks;
        G.x.segs = P->sweep[dir].segs;
        G.x.Y = Y[g];
        G.x.y_f64 = io == MX_IO_F64;
        G.st = P->sweep[dir].st;
        G.w = P->sweep[dir].w;
        G.ncols = bs.ncols[k];
        G.cbase = bs.cbase[k];
    }
    if (bs.tables.size() >= 8) {
        HIPCHK(hipFree(bs.tables.front().d));
        bs.tables.erase(bs.tables.begin());
    }
    mx_batch::BSweep::Table t{ys, (int)io, gen, upload(h.data(), h.size())};
    bs.tables.push_back(t);
    return t.d;
}

// all groups, forked from and joined back to st; dir DIR_FWD (Y = K~ W) or
// DIR_ADJ (Y = K~^T W).  With the batched sweep on, the BF16x3 groups run
// as one launch on st and the other groups beside it.
static void batch_eval(mx_batch *B, int dir, mx_precision prec, const void *const *W, void *const *Y, mx_io_type io,
                       hipStream_t st) {
    const int ng = (int)B->plans.size();
    for (int g = 0; g < ng; ++g) arg_check(W[g] && Y[g], "null W or Y of group " + std::to_string(g));
    const bool sw = batch_sweep_on(B, dir, prec);
    std::vector<char> in_sweep(ng, 0);
    if (sw) {
        const mx_batch::BSweep &bs = B->bsweep[dir];
        if (bs.var != B->sweep_var[dir]) throw MxError(MX_ERR_STATE, "internal: batched sweep lists not built");
        for (int g : bs.groups) in_sweep[g] = 1;
    }
    if (ng == 1 && !sw) {
        eval(B->plans[0], dir, group_prec(B, 0, prec), W[0], Y[0], B->group_nrhs[0], io, st);
        return;
    }
    if (ng > 1) HIPCHK(hipEventRecord(B->events[0], st));
    for (int g = 0; g < ng; ++g) {
        if (in_sweep[g]) continue;
        HIPCHK(hipStreamWaitEvent(B->streams[g], B->events[0], 0));
        eval(B->plans[g], dir, group_prec(B, g, prec), W[g], Y[g], B->group_nrhs[g], io, B->streams[g]);
        HIPCHK(hipEventRecord(B->events[g + 1], B->streams[g]));
    }
    if (sw) {
        mx_batch::BSweep &bs = B->bsweep[dir];
        SweepGroup *tab = batch_sweep_table(B, dir, io, Y, false, batch_gen(B));
        if (!tab) throw MxError(MX_ERR_STATE, "internal: batched sweep group table missing");
        for (int g : bs.groups) {
            mx_plan *P = B->plans[g];
            const int n = B->group_nrhs[g];
            const size_t wp = (size_t)P->wp_rows * n;
            launch_x3_split_w(W[g], io == MX_IO_F64, P->d_wrows, P->nleaf, n, P->panels, P->panels + wp, st);
        }
        X3SweepArgs sa{};
        sa.q = bs.q;
        sa.qinfo = bs.qinfo;
        sa.ctl = bs.ctl;
        sa.ncols = 1;
        sa.groups = tab;
        const int grid = (int)std::min<int64_t>(std::max(sweep_max_grid(bs.var), 1), std::max<int64_t>(bs.items, 1));
        launch_x3_sweep_batch(sa, bs.var, grid, bs.ctl_ints, st);
        HIPCHK(hipGetLastError());
    }
    if (ng > 1)
        for (int g = 0; g < ng; ++g)
            if (!in_sweep[g]) HIPCHK(hipStreamWaitEvent(st, B->events[g + 1], 0));
}

// Allocate every work buffer an evaluation needs before it is captured
// (and, for the batched sweep, its lists and the group table of Y).
static void batch_size_work(mx_batch *B, int dir, mx_precision prec, mx_io_type io, void *const *Y) {
    for (size_t g = 0; g < B->plans.size(); ++g) {
        mx_plan *P = B->plans[g];
        const mx_precision gp = group_prec(B, (int)g, prec);
        if (!(P->precisions & (1u << gp))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the batch");
        if (dir == DIR_ADJ && !P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "batch was created without MX_PLAN_ADJOINT");
        if (gp == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 evaluation needs FP64 W and Y");
        size_work(P, dir, gp, io, B->group_nrhs[g]);
    }
    if (batch_sweep_on(B, dir, prec)) {
        if (B->bsweep[dir].var != B->sweep_var[dir]) build_batch_sweep(B, dir, B->sweep_var[dir]);
        for (int g : B->bsweep[dir].groups) arg_check(Y[g] != nullptr, "null Y of group " + std::to_string(g));
        (void)batch_sweep_table(B, dir, io, Y, true, batch_gen(B));
    }
}

// Sum of the group plans' buffer generations: it changes whenever any work
// buffer of any group is reallocated (also through mx_batch_plan).
static uint64_t batch_gen(const mx_batch *B) {
    uint64_t g = 0;
    for (const mx_plan *P : B->plans) g += P->buf_gen;
    return g;
}

static void batch_drop_stale_graphs(mx_batch *B) {
    const uint64_t gen = batch_gen(B);
    for (size_t i = 0; i < B->graphs.size();) {
        if (B->graphs[i].gen != gen) {
            HIPCHK(hipGraphExecDestroy(B->graphs[i].exec));
            B->graphs.erase(B->graphs.begin() + (std::ptrdiff_t)i);
        } else {
            ++i;
        }
    }
}

static void batch_note_eval(mx_batch *B, int dir, mx_precision prec) {
    for (size_t g = 0; g < B->plans.size(); ++g) note_eval(B->plans[g], dir, group_prec(B, (int)g, prec), B->group_nrhs[g]);
}

// The whole batch as one captured hipGraph (all groups, forked and joined
// inside it), replayed while no group plan's work buffers have moved.
static void batch_eval_graph(mx_batch *B, int dir, mx_precision prec, const void *const *W, void *const *Y,
                             mx_io_type io, hipStream_t st) {
    const int ng = (int)B->plans.size();
    batch_drop_stale_graphs(B);
    for (auto &g : B->graphs) {
        bool same = g.dir == dir && g.prec == prec && g.io == io;
        for (int i = 0; same && i < ng; ++i) same = g.W[i] == W[i] && g.Y[i] == Y[i];
        if (same) {
            g.used = ++B->clock;
            HIPCHK(hipGraphLaunch(g.exec, st));
            batch_note_eval(B, dir, prec);
            return;
        }
    }
    batch_size_work(B, dir, prec, io, Y);
    batch_drop_stale_graphs(B);
    const uint64_t gen = batch_gen(B);
    for (auto *P : B->plans) P->profiling = 0;
    hipStream_t cs = st;
    bool own = false;
    if (cs == nullptr) {
        HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        own = true;
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    try {
        HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            batch_eval(B, dir, prec, W, Y, io, cs);
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
        if (graph) (void)hipGraphDestroy(graph);
        if (own) (void)hipStreamDestroy(cs);
        throw;
    }
    if (own) HIPCHK(hipStreamDestroy(cs));
    if (batch_gen(B) != gen) {   // sized above; a capture must not move them
        (void)hipGraphExecDestroy(exec);
        throw MxError(MX_ERR_STATE, "work buffers moved during graph capture");
    }
    if (B->graphs.size() >= 4) {
        auto lru = std::min_element(B->graphs.begin(), B->graphs.end(),
                                    [](const mx_batch::GraphEntry &a, const mx_batch::GraphEntry &b) {
                                        return a.used < b.used;
                                    });
        HIPCHK(hipGraphExecDestroy(lru->exec));
        B->graphs.erase(lru);
    }
    mx_batch::GraphEntry e;
    e.dir = dir;
    e.prec = prec;
    e.io = io;
    e.W.assign(W, W + ng);
    e.Y.assign(Y, Y + ng);
    e.exec = exec;
    e.used = ++B->clock;
    e.gen = gen;
    B->graphs.push_back(e);
    HIPCHK(hipGraphLaunch(exec, st));
    batch_note_eval(B, dir, prec);
}

static void batch_drop_graphs(mx_batch *B) {
    for (auto &g : B->graphs) HIPCHK(hipGraphExecDestroy(g.exec));
    B->graphs.clear();
}

}  // namespace mx

using namespace mx;

extern "C" {

mx_status mx_batch_create(mx_batch **batch, const mx_structure *const *instances, const int32_t *nrhs, int32_t count,
                          uint32_t precisions) {
    return (mx_status)guard([&]() {
        arg_check(batch != nullptr, "null argument");
        *batch = nullptr;
        *batch = create_guarded_batch(instances, nrhs, count, precisions);
        return 0;
    });
}

mx_status mx_batch_destroy(mx_batch *batch) {
    delete batch;
    return MX_OK;
}

mx_status mx_batch_groups(const mx_batch *batch, int32_t *ngroups) {
    return (mx_status)guard([&]() {
        arg_check(batch && ngroups, "null argument");
        *ngroups = (int32_t)batch->plans.size();
        return 0;
    });
}

mx_status mx_batch_group_info(const mx_batch *batch, int32_t group, int32_t *nrhs, int64_t *rows) {
    return (mx_status)guard([&]() {
        arg_check(batch && nrhs && rows && group >= 0 && group < (int32_t)batch->plans.size(), "bad argument");
        *nrhs = batch->group_nrhs[group];
        *rows = batch->group_rows[group];
        return 0;
    });
}

mx_status mx_batch_layout(const mx_batch *batch, int32_t instance, int32_t *group, int64_t *row_offset) {
    return (mx_status)guard([&]() {
        arg_check(batch && group && row_offset && instance >= 0 && instance < (int32_t)batch->inst.size(),
                  "bad argument");
        *group = batch->inst[instance].group;
        *row_offset = batch->inst[instance].row_off;
        return 0;
    });
}

mx_status mx_batch_instance_guard(const mx_batch *batch, int32_t instance, double *x3_err, int32_t *fp32) {
    return (mx_status)guard([&]() {
        arg_check(batch && x3_err && fp32 && instance >= 0 && instance < (int32_t)batch->inst.size(), "bad argument");
        *x3_err = batch->x3_err[instance];
        *fp32 = batch->group_fp32[batch->inst[instance].group];
        return 0;
    });
}

mx_status mx_batch_plan(mx_batch *batch, int32_t group, mx_plan **plan) {
    return (mx_status)guard([&]() {
        arg_check(batch && plan && group >= 0 && group < (int32_t)batch->plans.size(), "bad argument");
        *plan = batch->plans[group];
        return 0;
    });
}

static int32_t batch_run(mx_batch *batch, int dir, mx_precision prec, const void *const *W, void *const *Y,
                         mx_io_type io, void *stream) {
    return guard([&]() {
        arg_check(batch && W && Y, "null argument");
        arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
        if (dir == DIR_ADJ && !batch->plans[0]->adjoint)
            throw MxError(MX_ERR_UNSUPPORTED, "batch was created without MX_PLAN_ADJOINT");
        bool prof = false;
        for (auto *P : batch->plans) prof = prof || P->profiling;
        if (batch->use_graph && !prof) {
            batch_eval_graph(batch, dir, prec, W, Y, io, (hipStream_t)stream);
        } else {
            batch_size_work(batch, dir, prec, io, Y);
            batch_eval(batch, dir, prec, W, Y, io, (hipStream_t)stream);
            if (batch_sweep_on(batch, dir, prec)) batch_note_eval(batch, dir, prec);
        }
        return 0;
    });
}

mx_status mx_batch_set_sweep(mx_batch *batch, int32_t direction, int32_t x3_variant) {
    return (mx_status)guard([&]() {
        arg_check(batch && (direction == 0 || direction == 1), "bad argument");
        arg_check(x3_variant == 0 || sweep_variant_ok(x3_variant),
                  "sweep tile: 0 (off), MX_X3_128x128_S2, MX_X3_128x64_S2, MX_X3_128x32_S2 or MX_X3_128x32_K8");
        if (direction == DIR_ADJ && !batch->plans[0]->adjoint)
            throw MxError(MX_ERR_UNSUPPORTED, "batch was created without MX_PLAN_ADJOINT");
        if (x3_variant && batch->sweep_var[direction] != x3_variant) build_batch_sweep(batch, direction, x3_variant);
        batch->sweep_var[direction] = x3_variant;
        batch_drop_graphs(batch);
        return 0;
    });
}

mx_status mx_batch_get_sweep(const mx_batch *batch, int32_t direction, int32_t *x3_variant) {
    return (mx_status)guard([&]() {
        arg_check(batch && x3_variant && (direction == 0 || direction == 1), "bad argument");
        *x3_variant = batch->sweep_var[direction];
        return 0;
    });
}

mx_status mx_batch_sweep_status(const mx_batch *batch, void *stream, int32_t *error) {
    return (mx_status)guard([&]() {
        arg_check(batch && error, "null argument");
        *error = 0;
        HIPCHK(hipStreamSynchronize((hipStream_t)stream));
        for (const auto &bs : batch->bsweep)
            if (bs.ctl) {
                int32_t e = 0;
                HIPCHK(hipMemcpy(&e, bs.ctl + SWEEP_ERR, sizeof(int32_t), hipMemcpyDeviceToHost));
                *error = std::max(*error, e);
            }
        return 0;
    });
}

mx_status mx_batch_eval(mx_batch *batch, mx_precision prec, const void *const *W, void *const *Y, mx_io_type io,
                        void *stream) {
    return (mx_status)batch_run(batch, DIR_FWD, prec, W, Y, io, stream);
}

mx_status mx_batch_eval_adjoint(mx_batch *batch, mx_precision prec, const void *const *dY, void *const *dW,
                                mx_io_type io, void *stream) {
    return (mx_status)batch_run(batch, DIR_ADJ, prec, dY, dW, io, stream);
}

mx_status mx_batch_backward(mx_batch *batch, mx_precision prec, const void *const *W, const void *const *dY,
                            void *const *dW, mx_io_type io, void *const *dD, void *const *dB, void *const *dVT,
                            void *stream) {
    return (mx_status)guard([&]() {
        arg_check(batch && W && dY && dW, "null argument");
        if (!batch->plans[0]->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "batch was created without MX_PLAN_ADJOINT");
        const hipStream_t st = (hipStream_t)stream;
        for (size_t g = 0; g < batch->plans.size(); ++g)
            backward(batch->plans[g], group_prec(batch, (int)g, prec), W[g], dY[g], dW[g], batch->group_nrhs[g], io,
                     nullptr, nullptr,
                     dD ? dD[g] : nullptr, dB ? dB[g] : nullptr, dVT ? dVT[g] : nullptr, st);
        return 0;
    });
}

mx_status mx_batch_group_blocks(const mx_batch *batch, int32_t group, int64_t *D_len, int64_t *B_len,
                                int64_t *VT_len) {
    return (mx_status)guard([&]() {
        arg_check(batch && D_len && B_len && VT_len && group >= 0 && group < (int32_t)batch->plans.size(),
                  "bad argument");
        const mx_plan *P = batch->plans[group];
        *D_len = P->D_len;
        *B_len = P->B_len;
        *VT_len = P->VT_len;
        return 0;
    });
}

mx_status mx_batch_block_layout(const mx_batch *batch, int32_t instance, int64_t *d_off, int64_t *b_off,
                                int64_t *vt_off) {
    return (mx_status)guard([&]() {
        arg_check(batch && d_off && b_off && vt_off && instance >= 0 && instance < (int32_t)batch->inst.size(),
                  "bad argument");
        const BatchInst &bi = batch->inst[instance];
        *d_off = bi.d_off;
        *b_off = bi.b_off;
        *vt_off = bi.vt_off;
        return 0;
    });
}

mx_status mx_batch_set_graph(mx_batch *batch, int32_t enable) {
    return (mx_status)guard([&]() {
        arg_check(batch != nullptr, "null batch");
        batch->use_graph = enable ? 1 : 0;
        if (!enable) batch_drop_graphs(batch);
        return 0;
    });
}

// With BF16x3: time the whole batch with its groups' own launches and with
// the batched sweep at each tile width, and keep the fastest (a choice
// replaces the groups' own launches only when more than 2% faster).
static void autotune_batch_sweep(mx_batch *B, mx_io_type io, int reps, hipStream_t st) {
    bool any = false;
    for (size_t g = 0; g < B->plans.size(); ++g)
        any = any || (!B->group_fp32[g] && (B->plans[g]->precisions & MX_PLAN_BF16X3) && B->plans[g]->sweep[0].ok);
    if (!any) return;
    const int ng = (int)B->plans.size();
    const size_t elem = io == MX_IO_F64 ? sizeof(double) : sizeof(float);
    std::vector<void *> W(ng, nullptr), Y(ng, nullptr);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        for (void *p : W)
            if (p) (void)hipFree(p);
        for (void *p : Y)
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    try {
        for (int g = 0; g < ng; ++g) {
            const int64_t nel = std::max<int64_t>(B->group_rows[g], 1) * B->group_nrhs[g];
            HIPCHK(hipMalloc(&W[g], (size_t)nel * elem));
            HIPCHK(hipMalloc(&Y[g], (size_t)nel * elem));
            launch_fill_uniform(W[g], nel, io == MX_IO_F64, st);
        }
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const int ndir = B->plans[0]->adjoint ? 2 : 1;
        for (int dir = 0; dir < ndir; ++dir) {
            auto time_batch = [&]() {
                batch_size_work(B, dir, MX_BF16X3, io, Y.data());
                batch_eval(B, dir, MX_BF16X3, W.data(), Y.data(), io, st);   // warm-up
                HIPCHK(hipEventRecord(e0, st));
                for (int r = 0; r < reps; ++r) batch_eval(B, dir, MX_BF16X3, W.data(), Y.data(), io, st);
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                return ms / reps;
            };
            B->sweep_var[dir] = 0;
            float best = time_batch();
            int pick = 0;
            for (int v : {X3_V128x128_S2, X3_V128x64_S2, X3_V128x32_S2, X3_V128x32_K8}) {
                if (B->bsweep[dir].var != v) build_batch_sweep(B, dir, v);
                B->sweep_var[dir] = v;
                const float ms = time_batch();
                if (ms < 0.98f * best) {
                    best = ms;
                    pick = v;
                }
            }
            if (pick && B->bsweep[dir].var != pick) build_batch_sweep(B, dir, pick);
            B->sweep_var[dir] = pick;
            batch_note_eval(B, dir, MX_BF16X3);
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

mx_status mx_batch_autotune(mx_batch *batch, mx_precision prec, mx_io_type io, int32_t reps, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(batch != nullptr, "null batch");
        if (prec == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 evaluation needs FP64 I/O");
        for (size_t g = 0; g < batch->plans.size(); ++g)
            autotune(batch->plans[g], group_prec(batch, (int)g, prec), batch->group_nrhs[g], io, reps,
                     (hipStream_t)stream, nullptr);
        if (prec == MX_BF16X3) autotune_batch_sweep(batch, io, std::max(reps, 1), (hipStream_t)stream);
        batch_drop_graphs(batch);
        return 0;
    });
}

mx_status mx_batch_flops(const mx_batch *batch, mx_precision prec, double *total) {
    return (mx_status)guard([&]() {
        arg_check(batch && total, "null argument");
        *total = 0.0;
        for (size_t g = 0; g < batch->plans.size(); ++g) {
            const mx_plan *P = batch->plans[g];
            const int kd = group_prec(batch, (int)g, prec) == MX_BF16X3 ? 1 : 0;
            // algorithmic: the coarse schedule's composite work is not counted
            const int sc = sched_coarse(P->sched[DIR_FWD][kd]) ? SCHED_FUSED : P->sched[DIR_FWD][kd];
            for (double f : P->prog[DIR_FWD][kd][sc].flops) *total += f * batch->group_nrhs[g];
        }
        return 0;
    });
}

mx_status mx_batch_set_points(mx_batch *batch, int32_t instance, const double *X_leaf, int32_t d, double h) {
    return (mx_status)guard([&]() {
        arg_check(batch && X_leaf && d >= 1 && h > 0 && instance >= 0 && instance < (int32_t)batch->inst.size(),
                  "bad argument");
        BatchInst &bi = batch->inst[instance];
        if (bi.X_leaf) HIPCHK(hipFree(bi.X_leaf));
        bi.X_leaf = nullptr;
        bi.X_leaf = upload(X_leaf, (size_t)bi.N * d);
        bi.pts_d = d;
        bi.pts_h = h;
        return 0;
    });
}

mx_status mx_batch_dense_rows(mx_batch *batch, int32_t instance, const int32_t *rows, int32_t nrows, const double *W,
                              double *out, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(batch && instance >= 0 && instance < (int32_t)batch->inst.size(), "bad argument");
        BatchInst &bi = batch->inst[instance];
        if (!bi.X_leaf) throw MxError(MX_ERR_STATE, "mx_batch_set_points has not been called for this instance");
        dense_rows_core(bi.X_leaf, bi.N, bi.pts_d, bi.pts_h, bi.nleaf, bi.wrow.data(), bi.Ddim.data(), rows, nrows,
                        W, batch->group_nrhs[bi.group], out, (hipStream_t)stream);
        return 0;
    });
}

}  // extern "C"
