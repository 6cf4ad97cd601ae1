// Plans: structure validation, per-destination task lists, device uploads and
// the launch sequence of one evaluation.  C-ABI in include/matrox_hgemm.h.
//
// Evaluation order (Y = K~ W, semantics of symGen/HGEMM_gen.h:117-238):
//   UP        one launch per tree level, leaves first:
//             w_v = VT_v W_leaf  |  w_v = VT_v[:, :s_l] w_l + VT_v[:, s_l:] w_r
//   FAR+DOWN  one launch per level, root's children first; each task owns one
//             node c and sums its far couplings and its parent's transfer:
//             u_c = beta*u_c + sum_{far (c,y)} B_cy w_y + VT_p[:, part(c)]^T u_p
//   NEAR+LEAF one launch; each task owns one leaf block of Y:
//             Y_x = beta*Y_x + sum_{near (x,y)} D_xy W_y + VT_x^T u_x
// The reference runs NEAR first, then UP, FAR and DOWN with beta = 1 updates
// (HGEMM_gen.h:125-236); fusing the transfers into the coupling and near
// tasks gives the same sums with each output block written once per launch.
#include "../../../include/matrox_hgemm.h"
#include "mx_bf16x3.h"
#include "mx_grad.h"
#include "mx_seg_gemm.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mx {

static thread_local std::string g_err;

struct MxError : std::runtime_error {
    mx_status code;
    MxError(mx_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            throw MxError(MX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));             \
    } while (0)

static void arg_check(bool ok, const std::string &msg) {
    if (!ok) throw MxError(MX_ERR_ARG, msg);
}

struct Launch {
    int tile_offset;  // into the plan's tile array
    int ntiles;
};

// The task lists of one evaluation direction (forward Y = K~ W or adjoint
// dW = K~^T dY) for one kernel kind (0: segmented FP64/FP32 GEMM, 1: BF16x3).
struct Program {
    Task *tasks = nullptr;
    Seg *segs = nullptr;
    int2 *tiles = nullptr;
    std::vector<Launch> launches;
    std::vector<double> flops;   // algorithmic flops per nrhs column of each launch
    bool built = false;
    void release() {
        if (tasks) (void)hipFree(tasks);
        if (segs) (void)hipFree(segs);
        if (tiles) (void)hipFree(tiles);
        tasks = nullptr;
        segs = nullptr;
        tiles = nullptr;
    }
};

enum { DIR_FWD = 0, DIR_ADJ = 1 };

// Schedules of the coupling and downward phases:
//  SCHED_FUSED: one launch per level, u_c = sum_far B w + VT_p[:, part]^T u_p;
//  SCHED_SPLIT: one launch for the couplings of every level (they only need
//      the upward results), then one launch per level adding the transfer
//      from the parent (beta = 1).  Long coupling lists at the top of the
//      tree (the adjoint's column-grouped lists) then run beside everything
//      else instead of in a nearly empty launch.
enum { SCHED_FUSED = 0, SCHED_SPLIT = 1, NSCHED = 2 };

}  // namespace mx

struct mx_plan {
    int device = 0;
    int nleaf = 0, numnodes = 0, is_hss = 0;
    int64_t N = 0, skel_rows = 0;
    std::vector<int32_t> wrow, Ddim, skelrow, slen;

    // task lists: prog[direction][kind][schedule]; each is UP levels, the
    // coupling and downward launches, NEAR+LEAF (kind 1: 128-row tiles and
    // packed split A operands)
    mx::Program prog[2][2][mx::NSCHED];
    int sched[2][2] = {{mx::SCHED_FUSED, mx::SCHED_FUSED}, {mx::SCHED_SPLIT, mx::SCHED_SPLIT}};
    bool adjoint = false;
    std::vector<int64_t> wprow, sprow;  // padded panel row offsets (W leaves, skeleton nodes)
    int64_t wp_rows = 0, sp_rows = 0;
    int x3_variant = mx::X3_V128x128_S2;   // BF16x3 kernel variant (tile, pipeline depth)
    // per-launch BF16x3 variants [direction][schedule] chosen by the
    // autotuner (empty: x3_variant for every launch)
    std::vector<int> lvar[2][mx::NSCHED];
    mx::X3Rows *d_wrows = nullptr;
    mx::X3Rows *d_srows = nullptr;      // skeleton blocks (slen > 0) for panel export
    int n_srows = 0;
    int panels_dir = -1, panels_nrhs = 0;   // the evaluation the skeleton panels hold
    uint16_t *panels = nullptr;         // W, wskel, uskel split panels (hi, lo)
    size_t panel_bytes = 0;

    // stored blocks per precision
    double *D64 = nullptr, *B64 = nullptr, *VT64 = nullptr;
    float *D32 = nullptr, *B32 = nullptr, *VT32 = nullptr;
    uint16_t *xhi[3] = {nullptr, nullptr, nullptr};   // packed split D, B, VT operands
    uint16_t *xlo[3] = {nullptr, nullptr, nullptr};
    int64_t x3_elems[3] = {0, 0, 0};
    int64_t D_len = 0, B_len = 0, VT_len = 0;
    uint32_t precisions = 0;

    // gradient products of the stored blocks (plans built with MX_PLAN_ADJOINT)
    mx::GTask *g_tasks = nullptr;
    mx::GSeg *g_segs = nullptr;
    int4 *g_tiles = nullptr;
    int g_ntiles = 0;
    double g_flops = 0.0;               // algorithmic flops per nrhs column
    // packing lists of the split-BF16 operands, kept to repack new blocks
    mx::X3Block *d_xblk[3] = {nullptr, nullptr, nullptr};
    int64_t n_xblk[3] = {0, 0, 0};

    // work buffers: skeletons of the forward (w^, u^) and of the adjoint
    // (y^ = dL/du^, g^ = dL/dw^) evaluations
    void *wskel = nullptr, *uskel = nullptr;
    void *wskel_adj = nullptr, *uskel_adj = nullptr;
    size_t skel_bytes = 0, skel_adj_bytes = 0;
    void *io_w = nullptr, *io_y = nullptr, *io_g = nullptr;   // FP32 staging when io is FP64
    size_t io_bytes = 0, io_g_bytes = 0;
    int last_prec = -1;


    // points for the dense reference rows
    double *X_leaf = nullptr;
    int pts_d = 0;
    double pts_h = 0.0;

    // captured evaluations (hipGraph) keyed by their arguments
    struct GraphEntry {
        int dir, prec, io, nrhs;
        const void *W;
        void *Y;
        hipGraphExec_t exec;
        uint64_t used;
    };
    int use_graph = 0;
    std::vector<GraphEntry> graphs;
    uint64_t graph_clock = 0;

    // launch timing
    int profiling = 0;
    int keep_prof = 0;                // set while one call records several evaluations
    std::vector<hipEvent_t> ev;       // 2 per launch
    std::vector<int> ev_kind;
    int ev_used = 0;

    ~mx_plan() {
        for (auto &pd : prog)
            for (auto &pk : pd)
                for (auto &ps : pk) ps.release();
        void *ptrs[] = {D64, B64, VT64, D32, B32,
                        VT32, xhi[0], xhi[1], xhi[2], xlo[0], xlo[1], xlo[2], wskel, uskel, io_w, io_y, X_leaf,
                        d_wrows, d_srows, panels, g_tasks, g_segs, g_tiles, d_xblk[0], d_xblk[1], d_xblk[2],
                        wskel_adj, uskel_adj, io_g};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    }
};

namespace mx {

// Host description of the task lists of one evaluation.
struct TaskLists {
    std::vector<Task> tasks;
    std::vector<Seg> segs;
    std::vector<int2> tiles;
    std::vector<Launch> launches;
    std::vector<int> kinds;       // per launch: 0 = FP32/FP64 seg kernel, 1 = BF16x3 kernel
    std::vector<int> task_kind;   // per task
    std::vector<double> flops;    // per launch, algorithmic flops per nrhs column
    int dummy = -1;               // an empty task (no rows, no segments) padding launches
};

static void compute_launch_flops(TaskLists &L) {
    L.flops.assign(L.launches.size(), 0.0);
    std::vector<int> seen(L.tasks.size(), -1);
    for (size_t li = 0; li < L.launches.size(); ++li) {
        for (int t = L.launches[li].tile_offset; t < L.launches[li].tile_offset + L.launches[li].ntiles; ++t) {
            int task = L.tiles[t].x;
            if (seen[task] == (int)li) continue;   // count each task once per launch
            seen[task] = (int)li;
            const Task &tk = L.tasks[task];
            for (int sg = tk.seg_begin; sg < tk.seg_end; ++sg) L.flops[li] += 2.0 * tk.m * (double)L.segs[sg].k;
        }
    }
}

struct Topology {
    std::vector<int> parent, depth;
    std::vector<std::vector<int>> levels;  // node ids by depth (root excluded)
    std::vector<int> node_of_leaf;
};

static Topology topology(const mx_structure &s) {
    Topology t;
    const int nn = s.numnodes;
    t.parent.assign(nn, -1);
    t.depth.assign(nn, -1);
    t.node_of_leaf.assign(s.nleaf, -1);
    for (int v = 0; v < nn; ++v) {
        int l = s.lchildren[v], r = s.rchildren[v];
        arg_check((l == -1) == (r == -1), "node " + std::to_string(v) + " has exactly one child");
        if (l != -1) {
            arg_check(l > 0 && l < nn && r > 0 && r < nn && l != r, "child index out of range");
            arg_check(t.parent[l] == -1 && t.parent[r] == -1, "node with two parents");
            t.parent[l] = v;
            t.parent[r] = v;
        } else {
            int lf = s.lm[v];
            arg_check(lf >= 0 && lf < s.nleaf, "leaf node " + std::to_string(v) + " has no leaf index");
            arg_check(t.node_of_leaf[lf] == -1, "leaf index used twice");
            t.node_of_leaf[lf] = v;
        }
    }
    arg_check(t.parent[0] == -1, "node 0 must be the root");
    // depth by BFS from the root
    std::vector<int> q{0};
    t.depth[0] = 0;
    for (size_t i = 0; i < q.size(); ++i) {
        int v = q[i];
        if (s.lchildren[v] != -1) {
            for (int c : {s.lchildren[v], s.rchildren[v]}) {
                t.depth[c] = t.depth[v] + 1;
                q.push_back(c);
            }
        }
    }
    arg_check((int)q.size() == nn, "tree does not reach every node from the root");
    int maxd = 0;
    for (int v = 0; v < nn; ++v) maxd = std::max(maxd, t.depth[v]);
    t.levels.assign(maxd + 1, {});
    for (int v = 1; v < nn; ++v) t.levels[t.depth[v]].push_back(v);
    for (int l = 0; l < s.nleaf; ++l) arg_check(t.node_of_leaf[l] != -1, "leaf index without node");
    return t;
}

static int cols_of(const mx_structure &s, int v) {
    if (s.lchildren[v] == -1) return s.Ddim[s.lm[v]];
    return s.slen[s.lchildren[v]] + s.slen[s.rchildren[v]];
}

static void validate(const mx_structure &s, const Topology &t) {
    const int nn = s.numnodes;
    for (int l = 0; l < s.nleaf; ++l) arg_check(s.Ddim[l] >= 0, "negative leaf size");
    for (int v = 0; v < nn; ++v) arg_check(s.slen[v] >= 0, "negative skeleton length");
    for (int v = 1; v < nn; ++v) {
        int64_t need = (int64_t)s.VTptr[v] + (int64_t)s.slen[v] * cols_of(s, v);
        arg_check(s.slen[v] == 0 || (s.VTptr[v] >= 0 && need <= s.VT_len),
                  "VT block of node " + std::to_string(v) + " exceeds VT");
    }
    for (int64_t k = 0; k < s.ncount; ++k) {
        int x = s.is_hss ? (int)k : s.npairx[k], y = s.is_hss ? (int)k : s.npairy[k];
        arg_check(x >= 0 && x < s.nleaf && y >= 0 && y < s.nleaf, "near pair leaf out of range");
        int64_t need = (int64_t)s.Dptr[k] + (int64_t)s.Ddim[x] * s.Ddim[y];
        arg_check(need <= s.D_len, "near block " + std::to_string(k) + " exceeds D");
    }
    for (int64_t k = 0; k < s.fcount; ++k) {
        int x, y;
        if (s.is_hss) {
            x = (int)k + 1;
            arg_check(x < nn, "HSS far count exceeds numnodes-1");
            int p = t.parent[x];
            y = s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
        } else {
            x = s.fpairx[k];
            y = s.fpairy[k];
        }
        arg_check(x > 0 && x < nn && y > 0 && y < nn, "far pair node out of range");
        int64_t need = (int64_t)s.Bptr[k] + (int64_t)s.slen[x] * s.slen[y];
        arg_check(need <= s.B_len, "far block " + std::to_string(k) + " exceeds B");
    }
    for (int l = 0; l < s.nleaf; ++l) arg_check(s.wrow[l] >= 0, "negative leaf row offset");
    for (int v = 0; v < nn; ++v) arg_check(s.skelrow[v] >= 0, "negative skeleton row offset");
}

static int add_tiles(TaskLists &L, int task, int m) {
    const int bm = L.task_kind[task] == 1 ? X3_BM : SG_BM;
    int n = (m + bm - 1) / bm;
    for (int i = 0; i < n; ++i) L.tiles.push_back(make_int2(task, i));
    return n;
}

// Order of the tiles of every launch.  xcd_remap hands XCD x the x-th
// contiguous eighth of a launch's work items (the column tiles of a tile are
// adjacent).  The tiles, in structural order (tree order: neighbouring
// leaves and nodes share W and skeleton blocks, which then stay in one XCD's
// L2), are cut into eight contiguous groups of about equal work; each group
// is sorted longest first (so long lists, such as the adjoint's
// column-grouped couplings at the top of the tree, start early) and padded
// with empty tiles to the size of the largest, so that group x is exactly
// the eighth xcd_remap gives XCD x.
static void order_tiles(TaskLists &L, int bk) {
    if (L.dummy < 0) {
        Task tk{};
        tk.c_buf = VEC_USKEL;
        L.dummy = (int)L.tasks.size();
        L.tasks.push_back(tk);
        L.task_kind.push_back(L.task_kind.empty() ? 0 : L.task_kind.back());
    }
    std::vector<int2> out;
    std::vector<Launch> lns;
    for (const Launch &ln : L.launches) {
        const int T = ln.ntiles;
        std::vector<std::pair<int64_t, int2>> v;
        v.reserve(T);
        int64_t total = 0;
        for (int i = 0; i < T; ++i) {
            int2 tl = L.tiles[ln.tile_offset + i];
            const Task &tk = L.tasks[tl.x];
            int64_t cost = 2;   // prologue and epilogue
            for (int sg = tk.seg_begin; sg < tk.seg_end; ++sg) cost += (L.segs[sg].k + bk - 1) / bk;
            v.push_back({cost, tl});
            total += cost;
        }
        const int ng = 8;
        std::vector<int> cut(ng + 1, T);
        cut[0] = 0;
        int64_t acc = 0;
        int g = 1;
        for (int i = 0; i < T && g < ng; ++i) {
            acc += v[i].first;
            while (g < ng && acc * ng >= total * g) cut[g++] = i + 1;
        }
        int gmax = 0;
        for (int q = 0; q < ng; ++q) gmax = std::max(gmax, cut[q + 1] - cut[q]);
        Launch nl{(int)out.size(), gmax * ng};
        for (int q = 0; q < ng; ++q) {
            std::stable_sort(v.begin() + cut[q], v.begin() + cut[q + 1],
                             [](const std::pair<int64_t, int2> &a, const std::pair<int64_t, int2> &b) {
                                 return a.first > b.first;
                             });
            for (int i = cut[q]; i < cut[q + 1]; ++i) out.push_back(v[i].second);
            for (int i = cut[q + 1] - cut[q]; i < gmax; ++i) out.push_back(make_int2(L.dummy, 0));
        }
        lns.push_back(nl);
    }
    L.tiles.swap(out);
    L.launches.swap(lns);
}

// A task with more than maxseg segments keeps its first maxseg; the rest go
// to continuation tasks (beta = 1, same output block and row tiles), the
// k-th continuation in a launch of its own inserted right after the task's
// launch, so the parts accumulate in order without write conflicts.  Used
// for the BF16x3 kernel, which stages a task's segment list in LDS.
static void split_long_tasks(TaskLists &L, int maxseg) {
    std::vector<int2> tiles;
    std::vector<Launch> launches;
    std::vector<int> kinds;
    std::vector<int> seen(L.tasks.size(), -1);
    for (size_t li = 0; li < L.launches.size(); ++li) {
        const Launch ln = L.launches[li];
        // continuation tasks of this launch, by depth
        std::vector<std::vector<int>> cont;
        for (int t = ln.tile_offset; t < ln.tile_offset + ln.ntiles; ++t) {
            const int task = L.tiles[t].x;
            if (seen[task] == (int)li) continue;
            seen[task] = (int)li;
            const Task base = L.tasks[task];
            const int nseg = base.seg_end - base.seg_begin;
            if (nseg <= maxseg) continue;
            L.tasks[task].seg_end = base.seg_begin + maxseg;
            for (int k = 1; k * maxseg < nseg; ++k) {
                Task c = base;
                c.seg_begin = base.seg_begin + k * maxseg;
                c.seg_end = std::min(base.seg_end, c.seg_begin + maxseg);
                c.beta = 1;
                if ((int)cont.size() < k) cont.resize(k);
                cont[k - 1].push_back((int)L.tasks.size());
                L.tasks.push_back(c);
                L.task_kind.push_back(L.task_kind[task]);
            }
        }
        Launch nl{(int)tiles.size(), ln.ntiles};
        for (int t = ln.tile_offset; t < ln.tile_offset + ln.ntiles; ++t) tiles.push_back(L.tiles[t]);
        launches.push_back(nl);
        kinds.push_back(L.kinds[li]);
        for (auto &group : cont) {
            Launch cl{(int)tiles.size(), 0};
            for (int ct : group) {
                const int bm = L.task_kind[ct] == 1 ? X3_BM : SG_BM;
                for (int i = 0; i < (L.tasks[ct].m + bm - 1) / bm; ++i) tiles.push_back(make_int2(ct, i));
            }
            cl.ntiles = (int)tiles.size() - cl.tile_offset;
            launches.push_back(cl);
            kinds.push_back(L.kinds[li]);
        }
        seen.resize(L.tasks.size(), -1);
    }
    L.tiles.swap(tiles);
    L.launches.swap(launches);
    L.kinds.swap(kinds);
}

// Build the forward task lists for one kernel kind (0: segmented MFMA GEMM
// in FP64/FP32 with 64-row tiles, 1: split-BF16x3 with 128-row tiles).
// beta_u / beta_y select the accumulation rule of the coupling and leaf
// tasks (0 for mx_eval, the reference's rules for the drop-in entry points).
// wrow_in / skel_in give the row offsets of W blocks and skeleton blocks in
// the buffers the kernel reads (reference layout for kind 0, split panels for
// kind 1); Y always uses the reference layout.
static TaskLists build_lists(const mx_structure &s, const Topology &t, int beta_u, int beta_y, int kind, int dir,
                             int sched, const int64_t *wrow_in, const int64_t *skel_in) {
    TaskLists L;
    const int nn = s.numnodes;
    // far pair k couples row node fx(k) with column node fy(k) through B_k
    auto fx = [&](int64_t k) { return s.is_hss ? (int)k + 1 : s.fpairx[k]; };
    auto fy = [&](int64_t k) {
        if (!s.is_hss) return s.fpairy[k];
        int x = (int)k + 1, p = t.parent[x];
        return s.lchildren[p] == x ? s.rchildren[p] : s.lchildren[p];
    };
    auto nx = [&](int64_t k) { return s.is_hss ? (int)k : s.npairx[k]; };
    auto ny = [&](int64_t k) { return s.is_hss ? (int)k : s.npairy[k]; };
    // pairs grouped by destination: the row node (forward) or the column
    // node (adjoint, where the block acts transposed)
    std::vector<std::vector<int64_t>> far_of(nn), near_of(s.nleaf);
    for (int64_t k = 0; k < s.fcount; ++k) far_of[dir == DIR_FWD ? fx(k) : fy(k)].push_back(k);
    for (int64_t k = 0; k < s.ncount; ++k) near_of[dir == DIR_FWD ? nx(k) : ny(k)].push_back(k);

    auto begin_launch = [&]() {
        L.launches.push_back(Launch{(int)L.tiles.size(), 0});
        L.kinds.push_back(kind);
    };
    auto end_launch = [&]() {
        L.launches.back().ntiles = (int)L.tiles.size() - L.launches.back().tile_offset;
        if (L.launches.back().ntiles == 0) {
            L.launches.pop_back();
            L.kinds.pop_back();
        }
    };
    auto new_task = [&](int64_t c_row, int m, int c_buf, int beta) {
        Task tk{};
        tk.c_row = c_row;
        tk.m = m;
        tk.c_buf = c_buf;
        tk.seg_begin = (int)L.segs.size();
        tk.seg_end = tk.seg_begin;
        tk.beta = beta;
        L.tasks.push_back(tk);
        L.task_kind.push_back(kind);
        return (int)L.tasks.size() - 1;
    };
    auto add_seg = [&](int task, int64_t a_off, int a_buf, int a_trans, int lda, int k, int b_buf, int64_t b_row,
                       int b_ld) {
        if (k <= 0) return;
        Seg sg{};
        sg.a_off = a_off;
        sg.a_buf = a_buf;
        sg.a_trans = a_trans;
        sg.lda = lda;
        sg.k = k;
        sg.b_buf = b_buf;
        sg.b_row = b_row;
        sg.b_ld = b_ld;
        L.segs.push_back(sg);
        L.tasks[task].seg_end = (int)L.segs.size();
    };
    const int maxd = (int)t.levels.size() - 1;
    // UP: deepest level first
    for (int d = maxd; d >= 1; --d) {
        begin_launch();
        for (int v : t.levels[d]) {
            int sv = s.slen[v];
            if (sv == 0) continue;
            int tk = new_task(skel_in[v], sv, VEC_WSKEL, 0);
            if (s.lchildren[v] == -1) {
                int lf = s.lm[v];
                add_seg(tk, s.VTptr[v], MAT_VT, 0, sv, s.Ddim[lf], VEC_W, wrow_in[lf], s.Ddim[lf]);
            } else {
                int l = s.lchildren[v], r = s.rchildren[v];
                add_seg(tk, s.VTptr[v], MAT_VT, 0, sv, s.slen[l], VEC_WSKEL, skel_in[l], s.slen[l]);
                add_seg(tk, (int64_t)s.VTptr[v] + (int64_t)sv * s.slen[l], MAT_VT, 0, sv, s.slen[r], VEC_WSKEL,
                        skel_in[r], s.slen[r]);
            }
            add_tiles(L, tk, sv);
        }
        end_launch();
    }
    auto add_far = [&](int tk, int c) {
        const int sc = s.slen[c];
        for (int64_t k : far_of[c]) {
            if (dir == DIR_FWD) {
                int y = fy(k);
                add_seg(tk, (int64_t)s.Bptr[k], MAT_B, 0, sc, s.slen[y], VEC_WSKEL, skel_in[y], s.slen[y]);
            } else {   // B_k^T, B_k stored s_x x s_y
                int x = fx(k);
                add_seg(tk, (int64_t)s.Bptr[k], MAT_B, 1, s.slen[x], s.slen[x], VEC_WSKEL, skel_in[x], s.slen[x]);
            }
        }
    };
    auto add_down = [&](int tk, int c) {   // VT_p[:, part(c)]^T u_p; false when there is none
        int p = t.parent[c];
        if (p <= 0 || s.slen[p] == 0) return false;
        int sp = s.slen[p];
        int64_t off = s.VTptr[p];
        if (s.rchildren[p] == c) off += (int64_t)sp * s.slen[s.lchildren[p]];
        add_seg(tk, off, MAT_VT, 1, sp, sp, VEC_USKEL, skel_in[p], sp);
        return true;
    };
    if (sched == SCHED_FUSED) {
        // FAR + DOWN, root's children first: u_c = sum_far B w_y + VT_p[:, part(c)]^T u_p
        for (int d = 1; d <= maxd; ++d) {
            begin_launch();
            for (int c : t.levels[d]) {
                int sc = s.slen[c];
                if (sc == 0) continue;
                int tk = new_task(skel_in[c], sc, VEC_USKEL, beta_u);
                add_far(tk, c);
                add_down(tk, c);
                add_tiles(L, tk, sc);
            }
            end_launch();
        }
    } else {
        // FAR for every level at once, then DOWN level by level (beta = 1)
        begin_launch();
        for (int d = 1; d <= maxd; ++d)
            for (int c : t.levels[d]) {
                int sc = s.slen[c];
                if (sc == 0) continue;
                int tk = new_task(skel_in[c], sc, VEC_USKEL, beta_u);
                add_far(tk, c);
                add_tiles(L, tk, sc);
            }
        end_launch();
        for (int d = 2; d <= maxd; ++d) {
            begin_launch();
            for (int c : t.levels[d]) {
                int sc = s.slen[c];
                if (sc == 0) continue;
                int tk = new_task(skel_in[c], sc, VEC_USKEL, 1);
                if (add_down(tk, c)) add_tiles(L, tk, sc);
                else L.tasks.pop_back();
            }
            end_launch();
        }
    }
    // NEAR + LEAF: Y_x = sum_near D W_y + VT_x^T u_x
    begin_launch();
    for (int lf = 0; lf < s.nleaf; ++lf) {
        if (s.Ddim[lf] == 0) continue;
        int tk = new_task(s.wrow[lf], s.Ddim[lf], VEC_Y, beta_y);
        for (int64_t k : near_of[lf]) {
            if (dir == DIR_FWD) {
                int y = ny(k);
                add_seg(tk, (int64_t)s.Dptr[k], MAT_D, 0, s.Ddim[lf], s.Ddim[y], VEC_W, wrow_in[y], s.Ddim[y]);
            } else {   // D_k^T, D_k stored Dim_x x Dim_y
                int x = nx(k);
                add_seg(tk, (int64_t)s.Dptr[k], MAT_D, 1, s.Ddim[x], s.Ddim[x], VEC_W, wrow_in[x], s.Ddim[x]);
            }
        }
        int v = t.node_of_leaf[lf];
        int sv = s.slen[v];
        if (v > 0 && sv > 0) add_seg(tk, s.VTptr[v], MAT_VT, 1, sv, sv, VEC_USKEL, skel_in[v], sv);
        add_tiles(L, tk, s.Ddim[lf]);
    }
    end_launch();
    order_tiles(L, kind == 1 ? X3_BK : SG_BK);
    return L;
}

// Gradient products of every stored block (see mx_grad.h): one task per
// near block, far block, leaf VT and half of an internal VT, tiled 64 x 64.
struct GradLists {
    std::vector<GTask> tasks;
    std::vector<GSeg> segs;
    std::vector<int4> tiles;
    double flops = 0.0;   // per nrhs column
};

static GradLists build_grad_lists(const mx_structure &s, const Topology &t) {
    GradLists G;
    auto task = [&](int64_t c_off, int m, int n, int ldc, int c_buf) {
        GTask k{};
        k.c_off = c_off;
        k.m = m;
        k.n = n;
        k.ldc = ldc;
        k.c_buf = c_buf;
        k.seg_begin = k.seg_end = (int)G.segs.size();
        G.tasks.push_back(k);
        return (int)G.tasks.size() - 1;
    };
    auto seg = [&](int tk, int x_buf, int64_t x_row, int x_ld, int z_buf, int64_t z_row, int z_ld) {
        GSeg g{};
        g.x_buf = x_buf;
        g.x_row = x_row;
        g.x_ld = x_ld;
        g.z_buf = z_buf;
        g.z_row = z_row;
        g.z_ld = z_ld;
        G.segs.push_back(g);
        G.tasks[tk].seg_end = (int)G.segs.size();
        G.flops += 2.0 * x_ld * z_ld;
    };
    auto tiles = [&](int tk) {
        const GTask &k = G.tasks[tk];
        for (int i = 0; i < (k.m + 63) / 64; ++i)
            for (int j = 0; j < (k.n + 63) / 64; ++j) G.tiles.push_back(make_int4(tk, i, j, 0));
    };
    for (int64_t k = 0; k < s.ncount; ++k) {
        int x = s.is_hss ? (int)k : s.npairx[k], y = s.is_hss ? (int)k : s.npairy[k];
        int mx_ = s.Ddim[x], ny_ = s.Ddim[y];
        if (mx_ == 0 || ny_ == 0) continue;
        int tk = task((int64_t)s.Dptr[k], mx_, ny_, mx_, MAT_D);
        seg(tk, GB_DY, s.wrow[x], mx_, GB_W, s.wrow[y], ny_);
        tiles(tk);
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
        int sx = s.slen[x], sy = s.slen[y];
        if (sx == 0 || sy == 0) continue;
        int tk = task((int64_t)s.Bptr[k], sx, sy, sx, MAT_B);
        seg(tk, GB_YHAT, s.skelrow[x], sx, GB_WHAT, s.skelrow[y], sy);
        tiles(tk);
    }
    for (int v = 1; v < s.numnodes; ++v) {
        int sv = s.slen[v];
        if (sv == 0) continue;
        if (s.lchildren[v] == -1) {
            int lf = s.lm[v], dm = s.Ddim[lf];
            if (dm == 0) continue;
            int tk = task(s.VTptr[v], sv, dm, sv, MAT_VT);
            seg(tk, GB_GHAT, s.skelrow[v], sv, GB_W, s.wrow[lf], dm);
            seg(tk, GB_UHAT, s.skelrow[v], sv, GB_DY, s.wrow[lf], dm);
            tiles(tk);
        } else {
            int64_t off = s.VTptr[v];
            for (int c : {s.lchildren[v], s.rchildren[v]}) {
                int sc = s.slen[c];
                if (sc > 0) {
                    int tk = task(off, sv, sc, sv, MAT_VT);
                    seg(tk, GB_GHAT, s.skelrow[v], sv, GB_WHAT, s.skelrow[c], sc);
                    seg(tk, GB_UHAT, s.skelrow[v], sv, GB_YHAT, s.skelrow[c], sc);
                    tiles(tk);
                }
                off += (int64_t)sv * sc;
            }
        }
    }
    return G;
}

template <typename T>
static T *upload(const T *h, size_t n) {
    T *d = nullptr;
    HIPCHK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
    if (h && n > 0) HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

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

static int cvt_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 8192);
}

// ---- kernel blocks generated on the device --------------------------------
//
// D_k = K(points of leaf x, points of leaf y) and B_k = K(skeleton points of
// x, skeleton points of y) computed where they are needed instead of being
// uploaded, for matrices whose FP64 blocks do not fit the host (the
// reference computes them on the host, matroxTest-V1/compression.h:256-360,
// with the same norm expansion |a|^2 + |b|^2 - 2 a.b).

struct GenInfo {
    const double *X = nullptr;        // host, n x d row-major, original point order
    int64_t n = 0;
    int d = 0;
    double h = 1.0;
    const int32_t *lids = nullptr;    // host, N: leaf-order row -> point id
    const int32_t *skels = nullptr;   // host: skeleton point ids, node v at skoff[v]
    const int64_t *skoff = nullptr;   // host, numnodes
    int64_t nskels = 0;
};

struct alignas(16) GenBlock {
    int64_t out;    // element offset of the block (column-major na x nb)
    int64_t a, b;   // offsets of its row and column point ids in the index array
    int32_t na, nb;
};

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

// Points, norms and block lists of one matrix whose D and B are generated:
// near blocks index the leaf-order point ids (lids), coupling blocks the
// skeleton point ids.
struct GenState {
    double *Xd = nullptr, *nrm = nullptr;
    int32_t *idx[2] = {nullptr, nullptr};   // 0: lids, 1: skeleton ids
    int d = 0;
    double kscal = 0.0;
    std::vector<GenBlock> blk[2];           // 0: near (D), 1: coupling (B)
    std::map<int64_t, int> at[2];           // block offset -> index in blk
    GenState() = default;
    GenState(const GenState &) = delete;
    GenState &operator=(const GenState &) = delete;
    ~GenState() {
        for (void *p : {(void *)Xd, (void *)nrm, (void *)idx[0], (void *)idx[1]})
            if (p) (void)hipFree(p);
    }

    void init(const mx_structure &s, const Topology &t, const GenInfo &gen) {
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

    // w: 0 = D, 1 = B, in their stored layout
    template <typename T>
    void emit(int w, T *out) const {
        const std::vector<GenBlock> &b = blk[w];
        for (size_t b0 = 0; b0 < b.size(); b0 += 65535) {
            const size_t nb = std::min<size_t>(65535, b.size() - b0);
            GenBlock *db = upload(b.data() + b0, nb);
            gen_blocks<T><<<dim3(64, (unsigned)nb), 256>>>(Xd, nrm, d, kscal, idx[w], db, out);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipDeviceSynchronize();
            (void)hipFree(db);
            if (e != hipSuccess) throw MxError(MX_ERR_HIP, std::string("block generation: ") + hipGetErrorString(e));
        }
    }

    // packed split-BF16 operands of buffer w for the packing list xb
    void emit_x3(int w, const std::vector<X3Block> &xb, uint16_t *hi, uint16_t *lo) const {
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
};

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

static mx_plan *create_plan(const mx_structure &s, uint32_t precisions, int beta_u, int beta_y,
                            const GenInfo *gen = nullptr) {
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
    if (const char *e = getenv("MX_X3_VARIANT")) {
        int v = atoi(e);
        if (v >= 1 && v <= X3_NVARIANT) P->x3_variant = v;
    }

    std::vector<int64_t> wrow64(s.wrow, s.wrow + s.nleaf), skel64(s.skelrow, s.skelrow + s.numnodes);
    const int ndir = P->adjoint ? 2 : 1;
    for (int dir = 0; dir < ndir; ++dir)
        for (int sc = 0; sc < NSCHED; ++sc) {
            TaskLists L = build_lists(s, t, beta_u, beta_y, 0, dir, sc, wrow64.data(), skel64.data());
            compute_launch_flops(L);
            Program &pg = P->prog[dir][0][sc];
            upload_lists(L, &pg.tasks, &pg.segs, &pg.tiles);
            pg.launches = L.launches;
            pg.flops = L.flops;
            pg.built = true;
        }
    std::vector<X3Block> xblk[3];   // blocks to pack per stored buffer (D, B, VT)
    if (precisions & MX_PLAN_BF16X3) {
        // split panels: padded row offsets of W leaf blocks and skeleton blocks
        P->wprow.assign(s.nleaf, 0);
        P->sprow.assign(s.numnodes, 0);
        int64_t r = 0;
        for (int l = 0; l < s.nleaf; ++l) {
            P->wprow[l] = r;
            r += x3_pad_rows(s.Ddim[l]);
        }
        P->wp_rows = r;
        r = 0;
        for (int v = 0; v < s.numnodes; ++v) {
            P->sprow[v] = r;
            r += x3_pad_rows(s.slen[v]);
        }
        P->sp_rows = r;
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
        for (int dsc = 0; dsc < ndir * NSCHED; ++dsc) {
            const int dir = dsc / NSCHED, sc = dsc % NSCHED;
            TaskLists X = build_lists(s, t, beta_u, beta_y, 1, dir, sc, P->wprow.data(), P->sprow.data());
            split_long_tasks(X, X3_MAXSEG);
            for (size_t ti = 0; ti < X.tasks.size(); ++ti) {
                if (X.tasks[ti].seg_end - X.tasks[ti].seg_begin > X3_MAXSEG)
                    throw MxError(MX_ERR_ARG, "internal: a BF16x3 task longer than its segment window");
                for (int si = X.tasks[ti].seg_begin; si < X.tasks[ti].seg_end; ++si) {
                    Seg &sg = X.segs[si];
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
            compute_launch_flops(X);
            Program &pg = P->prog[dir][1][sc];
            upload_lists(X, &pg.tasks, &pg.segs, &pg.tiles);
            pg.launches = X.launches;
            pg.flops = X.flops;
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
        P->g_flops = G.flops;
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
            G->emit<double>(0, P->D64);
            G->emit<double>(1, P->B64);
        }
    }
    P->VT64 = upload(s.VT, (size_t)s.VT_len);
    if (precisions & MX_PLAN_FP32) {
        P->VT32 = upload<float>(nullptr, (size_t)s.VT_len);
        cvt_f64_f32<<<cvt_grid(s.VT_len), 256>>>(P->VT64, P->VT32, s.VT_len);
        P->D32 = upload<float>(nullptr, (size_t)s.D_len);
        P->B32 = upload<float>(nullptr, (size_t)s.B_len);
        if (gen) {
            G->emit<float>(0, P->D32);
            G->emit<float>(1, P->B32);
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
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    if (!(precisions & MX_PLAN_FP64)) {   // FP64 blocks were only a source
        for (double **p : {&P->D64, &P->B64, &P->VT64}) {
            if (*p) HIPCHK(hipFree(*p));
            *p = nullptr;
        }
    }
    return P.release();
}

static void ensure_adj_work(mx_plan *P, int nrhs, size_t elem) {
    size_t sb = (size_t)std::max<int64_t>(P->skel_rows, 1) * nrhs * elem;
    if (sb > P->skel_adj_bytes) {
        if (P->wskel_adj) HIPCHK(hipFree(P->wskel_adj));
        if (P->uskel_adj) HIPCHK(hipFree(P->uskel_adj));
        P->wskel_adj = P->uskel_adj = nullptr;
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

static void ensure_work(mx_plan *P, int nrhs, size_t elem, bool need_io) {
    size_t sb = (size_t)std::max<int64_t>(P->skel_rows, 1) * nrhs * elem;
    if (sb > P->skel_bytes) {
        if (P->wskel) HIPCHK(hipFree(P->wskel));
        if (P->uskel) HIPCHK(hipFree(P->uskel));
        P->wskel = P->uskel = nullptr;
        HIPCHK(hipMalloc(&P->wskel, sb));
        HIPCHK(hipMalloc(&P->uskel, sb));
        P->skel_bytes = sb;
    }
    if (need_io) {
        size_t ib = (size_t)std::max<int64_t>(P->N, 1) * nrhs * sizeof(float);
        if (ib > P->io_bytes) {
            if (P->io_w) HIPCHK(hipFree(P->io_w));
            if (P->io_y) HIPCHK(hipFree(P->io_y));
            P->io_w = P->io_y = nullptr;
            HIPCHK(hipMalloc(&P->io_w, ib));
            HIPCHK(hipMalloc(&P->io_y, ib));
            P->io_bytes = ib;
        }
    }
}

static void ensure_panels(mx_plan *P, int nrhs) {
    size_t bytes = ((size_t)2 * P->wp_rows + (size_t)4 * P->sp_rows) * nrhs * sizeof(uint16_t);
    if (bytes > P->panel_bytes) {
        if (P->panels) HIPCHK(hipFree(P->panels));
        P->panels = nullptr;
        HIPCHK(hipMalloc(&P->panels, bytes));
        P->panel_bytes = bytes;
    }
}

template <typename T>
static void run_seg_launch(const Launch &ln, const Task *tasks, const Seg *segs, const int2 *tiles, int nrhs,
                           const Bufs<T> &b, hipStream_t st) {
    int nnt = (nrhs + SG_BN - 1) / SG_BN;
    int64_t total = (int64_t)ln.ntiles * nnt;
    if (total <= 0) return;
    if (total > INT32_MAX) throw MxError(MX_ERR_ARG, "launch too large");
    hipLaunchKernelGGL(seg_gemm_kernel<T>, dim3((unsigned)total), dim3(SG_THREADS), 0, st, tasks, segs,
                       tiles + ln.tile_offset, ln.ntiles, nrhs, nnt, b);
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

static void eval(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
                 hipStream_t st) {
    arg_check(W && Y, "null W or Y");
    arg_check(nrhs >= 1, "nrhs must be >= 1");
    arg_check(io == MX_IO_F64 || io == MX_IO_F32, "bad io type");
    int dev;
    HIPCHK(hipGetDevice(&dev));
    if (dev != P->device) throw MxError(MX_ERR_STATE, "plan was created on another device");
    if (!(P->precisions & (1u << prec))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the plan");
    if (dir == DIR_ADJ && !P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
    const Program &pg0 = P->prog[dir][0][P->sched[dir][0]];
    const Program &pg1 = P->prog[dir][1][P->sched[dir][1]];
    const int64_t nel = P->N * (int64_t)nrhs;
    if (prec != MX_BF16X3 && dir == DIR_ADJ) ensure_adj_work(P, nrhs, prec == MX_FP64 ? sizeof(double) : sizeof(float));
    void *wsk = dir == DIR_ADJ ? P->wskel_adj : P->wskel;
    void *usk = dir == DIR_ADJ ? P->uskel_adj : P->uskel;
    if (!P->keep_prof) P->ev_used = 0;
    if (prec == MX_FP64) {
        if (io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 evaluation needs FP64 W and Y");
        ensure_work(P, nrhs, sizeof(double), false);
        if (dir == DIR_FWD) wsk = P->wskel, usk = P->uskel;
        Bufs<double> b;
        b.mat[MAT_D] = P->D64;
        b.mat[MAT_B] = P->B64;
        b.mat[MAT_VT] = P->VT64;
        b.vec[VEC_W] = const_cast<double *>(static_cast<const double *>(W));
        b.vec[VEC_WSKEL] = (double *)wsk;
        b.vec[VEC_USKEL] = (double *)usk;
        b.vec[VEC_Y] = (double *)Y;
        for (size_t i = 0; i < pg0.launches.size(); ++i) {
            prof_mark(P, st, MX_KIND_SEG_F64, true);
            run_seg_launch<double>(pg0.launches[i], pg0.tasks, pg0.segs, pg0.tiles, nrhs, b, st);
            prof_mark(P, st, MX_KIND_SEG_F64, false);
        }
    } else if (prec == MX_FP32) {
        ensure_work(P, nrhs, sizeof(float), io == MX_IO_F64);
        if (dir == DIR_FWD) wsk = P->wskel, usk = P->uskel;
        const float *Wf;
        float *Yf;
        if (io == MX_IO_F64) {
            prof_mark(P, st, MX_KIND_CONVERT, true);
            cvt_f64_f32<<<cvt_grid(nel), 256, 0, st>>>((const double *)W, (float *)P->io_w, nel);
            prof_mark(P, st, MX_KIND_CONVERT, false);
            Wf = (const float *)P->io_w;
            Yf = (float *)P->io_y;
        } else {
            Wf = (const float *)W;
            Yf = (float *)Y;
        }
        Bufs<float> b;
        b.mat[MAT_D] = P->D32;
        b.mat[MAT_B] = P->B32;
        b.mat[MAT_VT] = P->VT32;
        b.vec[VEC_W] = const_cast<float *>(Wf);
        b.vec[VEC_WSKEL] = (float *)wsk;
        b.vec[VEC_USKEL] = (float *)usk;
        b.vec[VEC_Y] = Yf;
        for (size_t i = 0; i < pg0.launches.size(); ++i) {
            prof_mark(P, st, MX_KIND_SEG_F32, true);
            run_seg_launch<float>(pg0.launches[i], pg0.tasks, pg0.segs, pg0.tiles, nrhs, b, st);
            prof_mark(P, st, MX_KIND_SEG_F32, false);
        }
        if (io == MX_IO_F64) {
            prof_mark(P, st, MX_KIND_CONVERT, true);
            cvt_f32_f64<<<cvt_grid(nel), 256, 0, st>>>(Yf, (double *)Y, nel);
            prof_mark(P, st, MX_KIND_CONVERT, false);
        }
    } else {
        ensure_panels(P, nrhs);
        const size_t wp = (size_t)P->wp_rows * nrhs, sp = (size_t)P->sp_rows * nrhs;
        X3Args xa{};
        xa.nrhs = nrhs;
        const std::vector<int> &lv = P->lvar[dir][P->sched[dir][1]];
        xa.y_f64 = io == MX_IO_F64;
        xa.Y = Y;
        xa.tasks = pg1.tasks;
        xa.segs = pg1.segs;
        for (int w = 0; w < 3; ++w) {
            xa.ahi[w] = P->xhi[w];
            xa.alo[w] = P->xlo[w];
        }
        uint16_t *pw = P->panels;
        xa.phi[VEC_W] = pw;
        xa.plo[VEC_W] = pw + wp;
        xa.phi[VEC_WSKEL] = pw + 2 * wp;
        xa.plo[VEC_WSKEL] = pw + 2 * wp + sp;
        xa.phi[VEC_USKEL] = pw + 2 * wp + 2 * sp;
        xa.plo[VEC_USKEL] = pw + 2 * wp + 3 * sp;
        prof_mark(P, st, MX_KIND_CONVERT, true);
        launch_x3_split_w(W, io == MX_IO_F64, P->d_wrows, P->nleaf, nrhs, xa.phi[VEC_W], xa.plo[VEC_W], st);
        prof_mark(P, st, MX_KIND_CONVERT, false);
        for (size_t i = 0; i < pg1.launches.size(); ++i) {
            const int v = lv.size() == pg1.launches.size() ? lv[i] : P->x3_variant;
            const int bn = x3_variant_bn(v);
            xa.n_ntiles = (nrhs + bn - 1) / bn;
            xa.tiles = pg1.tiles + pg1.launches[i].tile_offset;
            xa.ntiles = pg1.launches[i].ntiles;
            prof_mark(P, st, MX_KIND_BF16X3, true);
            launch_x3(xa, v, st);
            prof_mark(P, st, MX_KIND_BF16X3, false);
        }
        P->panels_dir = dir;
        P->panels_nrhs = nrhs;
    }
    HIPCHK(hipGetLastError());
    if (dir == DIR_FWD) P->last_prec = prec;
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

// Backward pass of Y = K~ W for the upstream gradient dY: dW = K~^T dY by the
// adjoint evaluation, then (when requested) the block gradients from the
// forward skeletons (w^, u^: caller copies, or the plan's own buffers) and the
// adjoint ones (y^, g^).  FP64 or FP32 compute; outputs in the I/O type.
static void backward(mx_plan *P, mx_precision prec, const void *W, const void *dY, void *dW, int nrhs, mx_io_type io,
                     const void *wskel, const void *uskel, void *dD, void *dB, void *dVT, hipStream_t st) {
    arg_check(W && dY && dW, "null W, dY or dW");
    arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
    arg_check(io == MX_IO_F64 || io == MX_IO_F32, "bad io type");
    if (prec == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 backward needs FP64 I/O");
    if (!P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
    if (!(P->precisions & (1u << prec))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the plan");
    const bool grads = dD || dB || dVT;
    const bool x3 = prec == MX_BF16X3;   // BF16x3 evaluations, FP32 gradient products
    const size_t elem = prec == MX_FP64 ? sizeof(double) : sizeof(float);
    if (grads && (!wskel || !uskel)) {
        if (P->last_prec != prec) throw MxError(MX_ERR_STATE, "no forward evaluation in this precision to take skeletons from");
        if (x3) {   // the forward skeletons are split panels: FP32 copies before the adjoint reuses the panels
            ensure_work(P, nrhs, sizeof(float), false);
            export_panels(P, DIR_FWD, nrhs, (float *)P->wskel, (float *)P->uskel, st);
        } else if ((size_t)P->skel_rows * nrhs * elem > P->skel_bytes) {
            throw MxError(MX_ERR_STATE, "the last forward evaluation had fewer columns");
        }
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
    const int kind = prec == MX_FP64 ? MX_KIND_GRAD_F64 : MX_KIND_GRAD_F32;
    if (x3) {   // adjoint skeletons to FP32; W and dY to FP32 when the I/O is FP64
        ensure_adj_work(P, nrhs, sizeof(float));
        export_panels(P, DIR_ADJ, nrhs, (float *)P->wskel_adj, (float *)P->uskel_adj, st);
        if (io == MX_IO_F64) {
            ensure_work(P, nrhs, sizeof(float), true);
            cvt_f64_f32<<<cvt_grid(nel), 256, 0, st>>>((const double *)dY, (float *)P->io_w, nel);
        }
    }
    if (prec == MX_FP64) {
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
    if (P->precisions & MX_PLAN_BF16X3) pack_all(P, src, st);
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

static void drop_graphs(mx_plan *P) {
    for (auto &g : P->graphs) HIPCHK(hipGraphExecDestroy(g.exec));
    P->graphs.clear();
}

static void set_schedule(mx_plan *P, int dir, mx_precision prec, int sched, int variant) {
    arg_check(dir == DIR_FWD || dir == DIR_ADJ, "bad direction");
    arg_check(sched == SCHED_FUSED || sched == SCHED_SPLIT, "bad schedule");
    arg_check(variant >= 0 && variant <= X3_NVARIANT, "bad BF16x3 kernel variant (0 keeps)");
    if (dir == DIR_ADJ && !P->adjoint) throw MxError(MX_ERR_UNSUPPORTED, "plan was created without MX_PLAN_ADJOINT");
    const int kd = prec == MX_BF16X3 ? 1 : 0;
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
static void autotune(mx_plan *P, mx_precision prec, int nrhs, mx_io_type io, int reps, hipStream_t st,
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
                    const int sc = si2 == 0 ? cur_sched[dir] : 1 - cur_sched[dir];   // current first
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
        if (prec == MX_BF16X3) {
            // per launch: the deep top-of-tree launches and the wide leaf
            // launches favour different variants
            P->profiling = 1;
            for (int dir = 0; dir < ndir; ++dir) {
                const int sc = P->sched[dir][1];
                const size_t nl = P->prog[dir][1][sc].launches.size();
                std::vector<int> &lv = P->lvar[dir][sc];
                std::vector<int> pick(nl, P->x3_variant);
                std::vector<float> best(nl, 0.f);
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
            }
            P->profiling = 0;
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

// Evaluation through a captured hipGraph: the launch sequence of one
// (precision, I/O type, nrhs, W, Y) is captured once and replayed; work
// buffers are sized before the capture and the sequence holds no memsets.
// Up to four graphs are kept per plan (least recently used is dropped).
static void eval_graph(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
                       hipStream_t st) {
    for (auto &g : P->graphs) {
        if (g.dir == dir && g.prec == prec && g.io == io && g.nrhs == nrhs && g.W == W && g.Y == Y) {
            g.used = ++P->graph_clock;
            HIPCHK(hipGraphLaunch(g.exec, st));
            if (dir == DIR_FWD) P->last_prec = prec;
            return;
        }
    }
    if (prec == MX_BF16X3) {
        ensure_panels(P, nrhs);
    } else {
        ensure_work(P, nrhs, prec == MX_FP64 ? sizeof(double) : sizeof(float), io == MX_IO_F64);
        if (dir == DIR_ADJ) ensure_adj_work(P, nrhs, prec == MX_FP64 ? sizeof(double) : sizeof(float));
    }
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
    if (P->graphs.size() >= 4) {
        auto lru = std::min_element(P->graphs.begin(), P->graphs.end(),
                                    [](const mx_plan::GraphEntry &a, const mx_plan::GraphEntry &b) {
                                        return a.used < b.used;
                                    });
        HIPCHK(hipGraphExecDestroy(lru->exec));
        P->graphs.erase(lru);
    }
    P->graphs.push_back(mx_plan::GraphEntry{dir, prec, io, nrhs, W, Y, exec, ++P->graph_clock});
    HIPCHK(hipGraphLaunch(exec, st));
    P->last_prec = prec;
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
static void dense_rows_core(const double *X_leaf, int64_t N, int d, double h, int nleaf, const int32_t *wrow,
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
        Launch ln{0, (int)L.tiles.size()};
        run_seg_launch<double>(ln, dt, ds, dtl, nrhs, b, st);
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

// ---- drop-in entry points -------------------------------------------------

struct CompatDims {
    bool set = false;
    int32_t nleaf = 0, nl = 0, fl = 0, cdepth = 0;
};
static CompatDims g_dims;

static std::vector<int64_t> visit_blocks(const int32_t *set, const int32_t *blocks, int32_t rows) {
    std::vector<int64_t> ks;
    for (int i = 0; i < rows; ++i)
        for (int j = set[i]; j < set[i + 1]; ++j)
            for (int k = blocks[j]; k < blocks[j + 1]; ++k) ks.push_back(k);
    return ks;
}

static int32_t compat_eval(bool hss, double *D, double *B, double *VT, uint64_t *Dptr, uint64_t *Bptr,
                           int32_t *VTptr, int32_t *lc, int32_t *rc, double *mrhs, double *apres, int32_t nrhs,
                           int32_t *Ddim, int32_t *wptr, int32_t *uptr, double *wskel, int32_t *wso,
                           double *uskel, int32_t *uso, int32_t *lm, int32_t *slen, int32_t *nblockSet,
                           int32_t *nblocks, int32_t *npairx, int32_t *npairy, int32_t *fblockSet,
                           int32_t *fblocks, int32_t *fpairx, int32_t *fpairy) {
    if (!g_dims.set) throw MxError(MX_ERR_STATE, "call mx_hgemm_configure before HGEMM");
    arg_check(nrhs >= 1, "nrhs must be >= 1");
    arg_check(D && B && VT && Dptr && Bptr && VTptr && lc && rc && mrhs && apres && Ddim && wptr && uptr &&
                  wskel && wso && uskel && uso && lm && slen,
              "null argument");
    const int nleaf = g_dims.nleaf, nn = 2 * nleaf - 1;
    mx_structure s{};
    s.nleaf = nleaf;
    s.numnodes = nn;
    s.is_hss = hss ? 1 : 0;
    s.lchildren = lc;
    s.rchildren = rc;
    s.lm = lm;
    s.Ddim = Ddim;
    s.slen = slen;
    s.VTptr = VTptr;
    std::vector<int32_t> wrow(nleaf), skelrow(nn);
    for (int l = 0; l < nleaf; ++l) {
        arg_check(wptr[l] % nrhs == 0 && uptr[l] == wptr[l], "wptr/uptr must be equal multiples of nrhs");
        wrow[l] = wptr[l] / nrhs;
    }
    for (int v = 0; v < nn; ++v) {
        arg_check(wso[v] % nrhs == 0 && uso[v] == wso[v], "skeleton offsets must be equal multiples of nrhs");
        skelrow[v] = wso[v] / nrhs;
    }
    s.wrow = wrow.data();
    s.skelrow = skelrow.data();
    std::vector<int32_t> px, py, fx, fy;
    std::vector<uint64_t> dp, bp;
    if (hss) {
        s.ncount = nleaf;
        s.fcount = nn - 1;
        s.Dptr = Dptr;
        s.Bptr = Bptr;
    } else {
        arg_check(nblockSet && nblocks && npairx && npairy && fblockSet && fblocks && fpairx && fpairy,
                  "null blocking array");
        for (int64_t k : visit_blocks(nblockSet, nblocks, g_dims.nl)) {
            px.push_back(npairx[k]);
            py.push_back(npairy[k]);
            dp.push_back(Dptr[k]);
        }
        for (int64_t k : visit_blocks(fblockSet, fblocks, g_dims.fl)) {
            fx.push_back(fpairx[k]);
            fy.push_back(fpairy[k]);
            bp.push_back(Bptr[k]);
        }
        s.ncount = (int64_t)px.size();
        s.fcount = (int64_t)fx.size();
        s.npairx = px.data();
        s.npairy = py.data();
        s.fpairx = fx.data();
        s.fpairy = fy.data();
        s.Dptr = dp.data();
        s.Bptr = bp.data();
    }
    // buffer extents implied by the offsets
    int64_t dl = 0, bl = 0, vl = 0;
    for (int64_t k = 0; k < s.ncount; ++k) {
        int x = hss ? (int)k : s.npairx[k], y = hss ? (int)k : s.npairy[k];
        arg_check(x >= 0 && x < nleaf && y >= 0 && y < nleaf, "near pair leaf out of range");
        dl = std::max<int64_t>(dl, (int64_t)s.Dptr[k] + (int64_t)Ddim[x] * Ddim[y]);
    }
    for (int64_t k = 0; k < s.fcount; ++k) {
        int x, y;
        if (hss) {
            x = (int)k + 1;
            y = (x & 1) ? x + 1 : x - 1;
        } else {
            x = s.fpairx[k];
            y = s.fpairy[k];
        }
        arg_check(x > 0 && x < nn && y > 0 && y < nn, "far pair node out of range");
        bl = std::max<int64_t>(bl, (int64_t)s.Bptr[k] + (int64_t)slen[x] * slen[y]);
    }
    for (int v = 1; v < nn; ++v) {
        arg_check((lc[v] == -1) == (rc[v] == -1), "node with one child");
        if (lc[v] == -1) arg_check(lm[v] >= 0 && lm[v] < nleaf, "leaf map out of range");
        else arg_check(lc[v] > 0 && lc[v] < nn && rc[v] > 0 && rc[v] < nn, "child out of range");
        int cols = lc[v] == -1 ? Ddim[lm[v]] : slen[lc[v]] + slen[rc[v]];
        vl = std::max<int64_t>(vl, (int64_t)VTptr[v] + (int64_t)slen[v] * cols);
    }
    s.D = D;
    s.B = B;
    s.VT = VT;
    s.D_len = dl;
    s.B_len = bl;
    s.VT_len = vl;
    int beta = hss ? 0 : 1;
    std::unique_ptr<mx_plan> P(create_plan(s, MX_PLAN_FP64, beta, beta));
    int64_t N = P->N, SK = P->skel_rows;
    double *dW = upload(mrhs, (size_t)(N * nrhs));
    double *dY = upload(apres, (size_t)(N * nrhs));
    ensure_work(P.get(), nrhs, sizeof(double), false);
    HIPCHK(hipMemcpy(P->uskel, uskel, (size_t)(SK * nrhs) * sizeof(double), hipMemcpyHostToDevice));
    try {
        eval(P.get(), DIR_FWD, MX_FP64, dW, dY, nrhs, MX_IO_F64, nullptr);
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(apres, dY, (size_t)(N * nrhs) * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(wskel, P->wskel, (size_t)(SK * nrhs) * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(uskel, P->uskel, (size_t)(SK * nrhs) * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
        (void)hipFree(dW);
        (void)hipFree(dY);
        throw;
    }
    HIPCHK(hipFree(dW));
    HIPCHK(hipFree(dY));
    return 0;
}

template <typename F>
static int32_t guard(F &&f) {
    try {
        return (int32_t)f();
    } catch (const MxError &e) {
        g_err = e.what();
        return (int32_t)e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return (int32_t)MX_ERR_ARG;
    }
}

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
    std::vector<int64_t> group_rows;
    std::vector<hipStream_t> streams;      // one per group (when there are several)
    std::vector<hipEvent_t> events;        // fork + one join event per group
    std::vector<mx::BatchInst> inst;
    int use_graph = 0;
    struct GraphEntry {
        int prec, io;
        std::vector<const void *> W;
        std::vector<void *> Y;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<GraphEntry> graphs;
    uint64_t clock = 0;
    ~mx_batch() {
        for (auto *p : plans) delete p;
        for (auto st : streams) (void)hipStreamDestroy(st);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        for (auto &i : inst)
            if (i.X_leaf) (void)hipFree(i.X_leaf);
    }
};

namespace mx {

static mx_batch *create_batch(const mx_structure *const *in, const int32_t *nrhs, int count, uint32_t precisions) {
    arg_check(in && nrhs && count >= 1, "bad batch arguments");
    std::unique_ptr<mx_batch> B(new mx_batch());
    std::map<int, std::vector<int>> by_nrhs;
    for (int i = 0; i < count; ++i) {
        arg_check(in[i] != nullptr && nrhs[i] >= 1, "instance " + std::to_string(i) + ": null structure or bad nrhs");
        by_nrhs[nrhs[i]].push_back(i);
    }
    B->inst.resize(count);
    for (auto &kv : by_nrhs) {
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
        mx_structure v = G.view();
        B->plans.push_back(create_plan(v, precisions, 0, 0));
        B->group_nrhs.push_back(kv.first);
        B->group_rows.push_back(G.N);
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

// all groups, forked from and joined back to st
static void batch_eval(mx_batch *B, mx_precision prec, const void *const *W, void *const *Y, mx_io_type io,
                       hipStream_t st) {
    const int ng = (int)B->plans.size();
    for (int g = 0; g < ng; ++g) arg_check(W[g] && Y[g], "null W or Y of group " + std::to_string(g));
    if (ng == 1) {
        eval(B->plans[0], DIR_FWD, prec, W[0], Y[0], B->group_nrhs[0], io, st);
        return;
    }
    HIPCHK(hipEventRecord(B->events[0], st));
    for (int g = 0; g < ng; ++g) {
        HIPCHK(hipStreamWaitEvent(B->streams[g], B->events[0], 0));
        eval(B->plans[g], DIR_FWD, prec, W[g], Y[g], B->group_nrhs[g], io, B->streams[g]);
        HIPCHK(hipEventRecord(B->events[g + 1], B->streams[g]));
    }
    for (int g = 0; g < ng; ++g) HIPCHK(hipStreamWaitEvent(st, B->events[g + 1], 0));
}

static void batch_size_work(mx_batch *B, mx_precision prec, mx_io_type io) {
    for (size_t g = 0; g < B->plans.size(); ++g) {
        mx_plan *P = B->plans[g];
        if (!(P->precisions & (1u << prec))) throw MxError(MX_ERR_UNSUPPORTED, "precision not built into the batch");
        if (prec == MX_BF16X3) ensure_panels(P, B->group_nrhs[g]);
        else ensure_work(P, B->group_nrhs[g], prec == MX_FP64 ? sizeof(double) : sizeof(float), io == MX_IO_F64);
    }
}

static void batch_eval_graph(mx_batch *B, mx_precision prec, const void *const *W, void *const *Y, mx_io_type io,
                             hipStream_t st) {
    const int ng = (int)B->plans.size();
    for (auto &g : B->graphs) {
        bool same = g.prec == prec && g.io == io;
        for (int i = 0; same && i < ng; ++i) same = g.W[i] == W[i] && g.Y[i] == Y[i];
        if (same) {
            g.used = ++B->clock;
            HIPCHK(hipGraphLaunch(g.exec, st));
            return;
        }
    }
    batch_size_work(B, prec, io);
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
            batch_eval(B, prec, W, Y, io, cs);
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
    if (B->graphs.size() >= 4) {
        auto lru = std::min_element(B->graphs.begin(), B->graphs.end(),
                                    [](const mx_batch::GraphEntry &a, const mx_batch::GraphEntry &b) {
                                        return a.used < b.used;
                                    });
        HIPCHK(hipGraphExecDestroy(lru->exec));
        B->graphs.erase(lru);
    }
    mx_batch::GraphEntry e;
    e.prec = prec;
    e.io = io;
    e.W.assign(W, W + ng);
    e.Y.assign(Y, Y + ng);
    e.exec = exec;
    e.used = ++B->clock;
    B->graphs.push_back(e);
    HIPCHK(hipGraphLaunch(exec, st));
}

static void batch_drop_graphs(mx_batch *B) {
    for (auto &g : B->graphs) HIPCHK(hipGraphExecDestroy(g.exec));
    B->graphs.clear();
}

}  // namespace mx

using namespace mx;

extern "C" {

mx_status mx_hgemm_configure(int32_t nleaf, int32_t nl, int32_t fl, int32_t cdepth) {
    return (mx_status)guard([&]() {
        arg_check(nleaf >= 1 && nl >= 0 && fl >= 0 && cdepth >= 0, "bad dimensions");
        g_dims.nleaf = nleaf;
        g_dims.nl = nl;
        g_dims.fl = fl;
        g_dims.cdepth = cdepth;
        g_dims.set = true;
        return 0;
    });
}

int32_t HGEMM(double *D, double *B, double *VT, uint64_t *Dptr, uint64_t *Bptr, int32_t *VTptr,
              int32_t *lchildren, int32_t *rchildren, int32_t *levelset, int32_t *idx, double *mrhs,
              double *apres, int32_t nrhs, int32_t *Ddim, int32_t *wptr, int32_t *uptr, double *wskel,
              int32_t *wskeloffset, double *uskel, int32_t *uskeloffset, int32_t *lm, int32_t *slen,
              int32_t *nblockSet, int32_t *nblocks, int32_t *npairx, int32_t *npairy, int32_t *fblockSet,
              int32_t *fblocks, int32_t *fpairx, int32_t *fpairy, int32_t *wpart, int32_t *clevelset) {
    (void)levelset;
    (void)idx;
    (void)wpart;
    (void)clevelset;
    return guard([&]() {
        return compat_eval(false, D, B, VT, Dptr, Bptr, VTptr, lchildren, rchildren, mrhs, apres, nrhs, Ddim,
                           wptr, uptr, wskel, wskeloffset, uskel, uskeloffset, lm, slen, nblockSet, nblocks,
                           npairx, npairy, fblockSet, fblocks, fpairx, fpairy);
    });
}

int32_t HGEMM_HSS(double *D, double *B, double *VT, uint64_t *Dptr, uint64_t *Bptr, int32_t *VTptr,
                  int32_t *lchildren, int32_t *rchildren, int32_t *levelset, int32_t *idx, double *mrhs,
                  double *apres, int32_t nrhs, int32_t *Ddim, int32_t *wptr, int32_t *uptr, double *wskel,
                  int32_t *wskeloffset, double *uskel, int32_t *uskeloffset, int32_t *lm, int32_t *slen,
                  int32_t *wpart, int32_t *clevelset) {
    (void)levelset;
    (void)idx;
    (void)wpart;
    (void)clevelset;
    return guard([&]() {
        return compat_eval(true, D, B, VT, Dptr, Bptr, VTptr, lchildren, rchildren, mrhs, apres, nrhs, Ddim, wptr,
                           uptr, wskel, wskeloffset, uskel, uskeloffset, lm, slen, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr);
    });
}

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
        size_t elem = plan->last_prec == MX_FP64 ? sizeof(double) : sizeof(float);
        size_t bytes = (size_t)plan->skel_rows * nrhs * elem;
        arg_check(bytes <= plan->skel_bytes, "nrhs exceeds the last evaluation's buffers");
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

mx_status mx_plan_get_schedule(const mx_plan *plan, int32_t direction, mx_precision prec, int32_t *schedule,
                               int32_t *x3_variant) {
    return (mx_status)guard([&]() {
        arg_check(plan && schedule && x3_variant && (direction == 0 || direction == 1), "bad argument");
        *schedule = plan->sched[direction][prec == MX_BF16X3 ? 1 : 0];
        *x3_variant = plan->x3_variant;
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
            const Program &pg = plan->prog[DIR_FWD][1][plan->sched[DIR_FWD][1]];
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

mx_status mx_batch_create(mx_batch **batch, const mx_structure *const *instances, const int32_t *nrhs, int32_t count,
                          uint32_t precisions) {
    return (mx_status)guard([&]() {
        arg_check(batch != nullptr, "null argument");
        *batch = nullptr;
        arg_check((precisions & MX_PLAN_ADJOINT) == 0, "batches are forward only");
        *batch = create_batch(instances, nrhs, count, precisions);
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

mx_status mx_batch_plan(mx_batch *batch, int32_t group, mx_plan **plan) {
    return (mx_status)guard([&]() {
        arg_check(batch && plan && group >= 0 && group < (int32_t)batch->plans.size(), "bad argument");
        *plan = batch->plans[group];
        return 0;
    });
}

mx_status mx_batch_eval(mx_batch *batch, mx_precision prec, const void *const *W, void *const *Y, mx_io_type io,
                        void *stream) {
    return (mx_status)guard([&]() {
        arg_check(batch && W && Y, "null argument");
        arg_check(prec == MX_FP64 || prec == MX_FP32 || prec == MX_BF16X3, "bad precision");
        bool prof = false;
        for (auto *P : batch->plans) prof = prof || P->profiling;
        if (batch->use_graph && !prof) batch_eval_graph(batch, prec, W, Y, io, (hipStream_t)stream);
        else batch_eval(batch, prec, W, Y, io, (hipStream_t)stream);
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

mx_status mx_batch_autotune(mx_batch *batch, mx_precision prec, mx_io_type io, int32_t reps, void *stream) {
    return (mx_status)guard([&]() {
        arg_check(batch != nullptr, "null batch");
        if (prec == MX_FP64 && io != MX_IO_F64) throw MxError(MX_ERR_UNSUPPORTED, "FP64 evaluation needs FP64 I/O");
        for (size_t g = 0; g < batch->plans.size(); ++g)
            autotune(batch->plans[g], prec, batch->group_nrhs[g], io, reps, (hipStream_t)stream, nullptr);
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
            const int kd = prec == MX_BF16X3 ? 1 : 0;
            for (double f : P->prog[DIR_FWD][kd][P->sched[DIR_FWD][kd]].flops) *total += f * batch->group_nrhs[g];
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

const char *mx_last_error(void) { return g_err.c_str(); }

}  // extern "C"
