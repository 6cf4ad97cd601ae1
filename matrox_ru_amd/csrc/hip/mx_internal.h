// Internal declarations shared by the evaluator's translation units:
//   mx_lists.hip     structure checks and per-destination task lists
//   mx_generate.hip  kernel blocks computed on the device from the points
//   mx_plan.hip      plans, evaluation, backward, autotuner, C-ABI
//   mx_dropin.hip    the drop-in HGEMM / HGEMM_HSS entry points
//   mx_batch.hip     many independent instances
#pragma once
#include "../../../include/matrox_hgemm.h"
#include "mx_bf16x3.h"
#include "mx_f64p.h"
#include "mx_grad.h"
#include "mx_seg_gemm.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mx {

extern thread_local std::string g_err;

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

inline void arg_check(bool ok, const std::string &msg) {
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
    // BF16x3 programs: the same launches cut into tiles of 2 X3_BM rows
    int2 *tiles2 = nullptr;
    std::vector<Launch> launches2;
    std::vector<double> flops;   // algorithmic flops per nrhs column of each launch
    int last_write[4] = {-1, -1, -1, -1};   // per VecBuf: the last launch writing it (-1: none)
    // FP64 programs of plans with FP64: the segments over packed A operands
    // (f64p_kernel, on tiles2) and the longest task of each launch
    Seg *segs_p = nullptr;
    std::vector<int> maxseg;
    bool built = false;
    void release() {
        if (tasks) (void)hipFree(tasks);
        if (segs) (void)hipFree(segs);
        if (segs_p) (void)hipFree(segs_p);
        segs_p = nullptr;
        if (tiles) (void)hipFree(tiles);
        if (tiles2) (void)hipFree(tiles2);
        tasks = nullptr;
        segs = nullptr;
        tiles = nullptr;
        tiles2 = nullptr;
    }
};

// A composite A operand of the coarse schedule: op(A1) (m x kmid) times
// op(A2) (kmid x n), both stored blocks (MAT_B or MAT_VT; op = transposed
// when *_trans), formed in FP64 when the plan is built.  In host lists a
// segment over it has a_buf = MAT_COMP and a_off = its index in comps.
constexpr int32_t MAT_COMP = 3;
struct Composite {
    int64_t a1_off, a2_off;
    int32_t a1_buf, a1_trans, a1_lda, a2_buf, a2_trans, a2_lda;
    int32_t m, kmid, n, pad;
};

// A composite as formed on the device: its FP64 product goes to c_off (m x
// n, column-major) of a scratch buffer, packed from there.
struct CompDesc {
    Composite c;
    int64_t c_off;
};

enum { DIR_FWD = 0, DIR_ADJ = 1 };

// Schedules of the coupling and downward phases:
//  SCHED_FUSED: one launch per level, u_c = sum_far B w + VT_p[:, part]^T u_p;
//  SCHED_SPLIT: one launch for the couplings of every level (they only need
//      the upward results), then one launch per level adding the transfer
//      from the parent (beta = 1).  Long coupling lists at the top of the
//      tree (the adjoint's column-grouped lists) then run beside everything
//      else instead of in a nearly empty launch.
// coupling/downward schedule (bit 0: split) and NEAR+LEAF balancing (bit 1);
//  SCHED_COARSE, SCHED_COARSE_SPLIT (BF16x3 only): the fused or the split
//      lists with runs of consecutive launches merged where the later ones
//      read the earlier ones' skeletons through few, small products
//      (coarsen in mx_lists.hip).
enum {
    SCHED_FUSED = 0,
    SCHED_SPLIT = 1,
    SCHED_FUSED_BAL = 2,
    SCHED_SPLIT_BAL = 3,
    SCHED_COARSE = 4,
    SCHED_COARSE_SPLIT = 5,
    NSCHED = 6
};
inline bool sched_coarse(int sc) { return sc == SCHED_COARSE || sc == SCHED_COARSE_SPLIT; }

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
    // dependency-driven sweeps (BF16x3) per direction: device copies of the
    // SweepLists, whether evaluations use them (instead of the level
    // launches of sched[dir][1]) and their tile width
    struct Sweep {
        mx::Task *tasks = nullptr;
        mx::Seg *segs = nullptr;
        mx::SweepTask *st = nullptr;
        mx::SweepWait *w = nullptr;
        mx::SweepTile *q = nullptr;
        int2 *qinfo = nullptr;
        std::vector<int2> qinfo_h;
        // host copies (batched sweeps merge the lists of several plans)
        std::vector<mx::SweepTile> h_tiles;
        std::vector<mx::SweepTask> h_st;
        std::vector<mx::SweepWait> h_w;
        int nblocks = 0;   // counter blocks
        bool ok = false;
        void release() {
            void *ptrs[] = {tasks, segs, st, w, q, qinfo};
            for (void *p : ptrs)
                if (p) (void)hipFree(p);
            tasks = nullptr;
            segs = nullptr;
            st = nullptr;
            w = nullptr;
            q = nullptr;
            qinfo = nullptr;
        }
    } sweep[2];
    int sweep_on[2] = {0, 0};
    int sweep_var[2] = {mx::X3_V128x128_S2, mx::X3_V128x128_S2};
    int sweep_grid[4] = {0, 0, 0, 0};   // resident workgroups per sweep tile (128, 64, 32 columns, split-K), 0: not queried
    int32_t *sweep_last_ctl = nullptr;   // control block of the last sweep enqueued (mx_plan_sweep_status)
    // per-launch BF16x3 variants [direction][schedule] chosen by the
    // autotuner (empty: x3_variant for every launch)
    std::vector<int> lvar[2][mx::NSCHED];
    // per-launch variant of the FP64 / FP32 kernel (tile size, prefetch
    // depth) [direction][0: FP64, 1: FP32][schedule] chosen by the autotuner
    // (empty: 64 x 64 tiles, depth 1)
    std::vector<int> lsg[2][2][mx::NSCHED];
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
    // FP64 A operands packed in 16 x 16 fragments (plans with FP64), and
    // their packing lists
    double *f64p[3] = {nullptr, nullptr, nullptr};
    float *f32p[3] = {nullptr, nullptr, nullptr};   // the same fragments in FP32 (plans with FP32)
    int64_t f64p_elems[3] = {0, 0, 0};
    mx::X3Block *d_fblk[3] = {nullptr, nullptr, nullptr};
    int64_t n_fblk[3] = {0, 0, 0};
    int64_t D_len = 0, B_len = 0, VT_len = 0;
    uint32_t precisions = 0;

    // gradient products of the stored blocks (plans built with MX_PLAN_ADJOINT)
    mx::GTask *g_tasks = nullptr;
    mx::GSeg *g_segs = nullptr;
    int4 *g_tiles = nullptr;            // 64 x 64 tiles
    int g_ntiles = 0;
    int4 *g_tiles_x3 = nullptr;         // GRAD_X3_TILE tiles
    int g_ntiles_x3 = 0;
    int4 *g_tiles_x3w = nullptr;        // GRAD_X3_WTILE tiles
    int g_ntiles_x3w = 0;
    int g_x3_tile = 0;                  // BF16x3 gradient tile: 0 = by padded work, 128 or 256
    bool g_x3w_auto = false;            // the padded work favours the 256 x 256 tiles
    // split-BF16x3 gradient products: segments over K-fragment blocks, the
    // blocks to pack (leaf blocks of W / dY, skeleton blocks in the column
    // layout and as split panels), and the K-fragment buffers of the six
    // operands (hi, lo), sized for k_nrhs columns
    mx::KSeg *g_ksegs = nullptr;
    mx::KBlock *kb_leaf = nullptr, *kb_skel = nullptr, *kb_panel = nullptr;
    int n_kb_leaf = 0, n_kb_skel = 0;
    int64_t kfr_leaf = 0, kfr_skel = 0;   // 16-row groups of a leaf-blocked and a skeleton operand
    uint16_t *kfrag = nullptr;
    size_t kfrag_bytes = 0;
    double g_flops = 0.0;               // algorithmic flops per nrhs column
    // packing lists of the split-BF16 operands, kept to repack new blocks
    mx::X3Block *d_xblk[3] = {nullptr, nullptr, nullptr};
    int64_t n_xblk[3] = {0, 0, 0};
    // coarse schedule (BF16x3): its composite A operands (formed in FP64
    // from B64 / VT64, which the plan then keeps) and their packing list
    // (into the packed VT operands)
    mx::CompDesc *comp_desc = nullptr;
    mx::X3Block *comp_xblk = nullptr;
    int64_t n_comp = 0, comp_elems = 0, comp_tiles = 1;
    std::vector<int64_t> comp_gen;   // descriptor ranges of the generations (a nested composite after its factor)

    // work buffers: skeletons of the forward (w^, u^) and of the adjoint
    // (y^ = dL/du^, g^ = dL/dw^) evaluations
    void *wskel = nullptr, *uskel = nullptr;
    void *wskel_adj = nullptr, *uskel_adj = nullptr;
    size_t skel_bytes = 0, skel_adj_bytes = 0;
    void *io_w = nullptr, *io_y = nullptr, *io_g = nullptr;   // FP32 staging when io is FP64
    size_t io_bytes = 0, io_g_bytes = 0;
    // the last forward evaluation (precision, nrhs) whose skeletons the work
    // buffers hold; -1 / 0 when none, or when its buffers were reallocated
    int last_prec = -1, last_nrhs = 0;
    // bumped whenever a buffer an evaluation reads or writes (wskel, uskel,
    // their adjoint copies, io_w / io_y, panels) is freed and allocated again;
    // graphs captured under an older generation hold freed addresses
    uint64_t buf_gen = 0;


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
        uint64_t gen;   // buf_gen at capture
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
        for (auto &sw : sweep) sw.release();
        for (auto &pd : prog)
            for (auto &pk : pd)
                for (auto &ps : pk) ps.release();
        void *ptrs[] = {D64, B64, VT64, D32, B32,
                        VT32, xhi[0], xhi[1], xhi[2], xlo[0], xlo[1], xlo[2], wskel, uskel, io_w, io_y, X_leaf,
                        d_wrows, d_srows, panels, g_tasks, g_segs, g_tiles, g_tiles_x3, g_tiles_x3w, d_xblk[0], d_xblk[1], d_xblk[2],
                        wskel_adj, uskel_adj, io_g, g_ksegs, kb_leaf, kb_skel, kb_panel, kfrag,
                        f64p[0], f64p[1], f64p[2], f32p[0], f32p[1], f32p[2], d_fblk[0], d_fblk[1], d_fblk[2],
                        comp_desc, comp_xblk};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    }
};

namespace mx {

struct TaskLists {
    std::vector<Composite> comps;   // coarse schedule: the composite A operands
    std::vector<Task> tasks;
    std::vector<Seg> segs;
    std::vector<int2> tiles;
    std::vector<Launch> launches;
    std::vector<int> kinds;       // per launch: 0 = FP32/FP64 seg kernel, 1 = BF16x3 kernel
    std::vector<int> task_kind;   // per task
    std::vector<double> flops;    // per launch, algorithmic flops per nrhs column
    int last_write[4] = {-1, -1, -1, -1};   // per VecBuf: the last launch writing it
    int dummy = -1;               // an empty task (no rows, no segments) padding launches
    std::vector<int2> tiles2;     // tiles of twice the rows (pair_tiles)
    std::vector<Launch> launches2;
};

struct Topology {
    std::vector<int> parent, depth;
    std::vector<std::vector<int>> levels;  // node ids by depth (root excluded)
    std::vector<int> node_of_leaf;
};

struct GradLists {
    std::vector<GTask> tasks;
    std::vector<GSeg> segs;
    std::vector<int4> tiles;      // 64 x 64 tiles (FP64 / FP32 products)
    std::vector<int4> tiles_x3;   // GRAD_X3_TILE tiles (split-BF16x3 products)
    std::vector<int4> tiles_x3w;  // GRAD_X3_WTILE tiles (the same products, 256 x 256)
    double area_x3 = 0.0, area_x3w = 0.0;   // tile area x segments of each tiling (padded work)
    double flops = 0.0;   // per nrhs column
};

// ---- task lists (mx_lists.hip) ----------------------------------------------

void compute_launch_flops(TaskLists &L);
Topology topology(const mx_structure &s);
// Bounds of every offset in the structure (throws MX_ERR_ARG).
void validate(const mx_structure &s, const Topology &t);
void split_long_tasks(TaskLists &L, int maxseg);
// tiles2 / launches2: every launch's tiles merged in row pairs, in the same
// order and XCD groups
void pair_tiles(TaskLists &L);
// kind 0: FP64/FP32 lists (64-row tiles), kind 1: BF16x3 (128-row tiles).
// The tiles of big launches are grouped by shared B operands for kind 1, or
// with group set; otherwise kept in tree order (order_tiles).
TaskLists build_lists(const mx_structure &s, const Topology &t, int beta_u, int beta_y, int kind, int dir, int sched,
                      const int64_t *wrow_in, const int64_t *skel_in, bool group = false);
GradLists build_grad_lists(const mx_structure &s, const Topology &t);

// The dependency-driven sweep of one direction (protocol in mx_bf16x3.h),
// built from the BF16x3 fused-schedule lists: the same tasks with each
// task's u^-reading segments moved last, its early and late waits, and the
// launches' XCD groups as queues.
struct SweepLists {
    std::vector<Task> tasks;
    std::vector<Seg> segs;
    std::vector<SweepTask> stasks;
    std::vector<SweepWait> waits;
    std::vector<SweepTile> tiles;   // queue 0, then queue 1, ...
    std::vector<int2> qinfo;        // per queue: first tile, tiles
    int nblocks = 0;                // counter blocks
    bool ok = false;                // false: the lists have a block with several writers
};
// X: fused-schedule BF16x3 lists after order_tiles / split_long_tasks;
// sprow: first panel row of every skeleton node.
SweepLists build_sweep(const TaskLists &X, int numnodes, const std::vector<int64_t> &sprow);
// Ints of a sweep control block for nrhs columns (the widest tiling).
int64_t sweep_ctl_ints(int numnodes, int nrhs);

// ---- device buffers ----------------------------------------------------------

template <typename T>
T *upload(const T *h, size_t n) {
    T *d = nullptr;
    HIPCHK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
    if (h && n > 0) HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

inline int cvt_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 8192);
}

// ---- generated blocks (mx_generate.hip) ---------------------------------------

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
    ~GenState();
    void init(const mx_structure &s, const Topology &t, const GenInfo &gen);
    // w: 0 = D, 1 = B, in their stored layout
    void emit(int w, double *out) const;
    void emit(int w, float *out) const;
    // packed split-BF16 operands of buffer w for the packing list xb
    void emit_x3(int w, const std::vector<X3Block> &xb, uint16_t *hi, uint16_t *lo) const;
};

// ---- plan core (mx_plan.hip) -----------------------------------------------------

mx_plan *create_plan(const mx_structure &s, uint32_t precisions, int beta_u, int beta_y,
                     const GenInfo *gen = nullptr);
void eval(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io, hipStream_t st);
// dW = K~^T dY and (for non-null outputs) the block gradients, from the
// forward skeletons wskel/uskel (null: the plan's own, of its last forward
// evaluation in this precision)
void backward(mx_plan *P, mx_precision prec, const void *W, const void *dY, void *dW, int nrhs, mx_io_type io,
              const void *wskel, const void *uskel, void *dD, void *dB, void *dVT, hipStream_t st);
// Work buffers of one evaluation: the FP64/FP32 skeletons of its direction,
// the FP32 staging of FP64 I/O (FP32 compute), the split panels (BF16x3).
struct Work {
    void *wsk = nullptr, *usk = nullptr;
    void *io_w = nullptr, *io_y = nullptr;
    uint16_t *panels = nullptr;
    int32_t *sweep_ctl = nullptr;   // BF16x3: control block of the sweep (sweep_ctl_ints)
    bool prof = false;   // record launch events
    // FP64/FP32: recorded on the stream once the last launch writing wskel
    // (resp. uskel) is enqueued, so copies of them can start early
    hipEvent_t after_wskel = nullptr, after_uskel = nullptr;
};
// Argument and state checks of an evaluation (throws).
void eval_checks(const mx_plan *P, int dir, mx_precision prec, const void *W, const void *Y, int nrhs, mx_io_type io);
// The launch sequence of one evaluation on the work buffers wk (no
// allocation, no synchronisation).
void eval_core(mx_plan *P, int dir, mx_precision prec, const void *W, void *Y, int nrhs, mx_io_type io,
               hipStream_t st, const Work &wk);
// Size the plan's own work buffers for one evaluation (may reallocate).
void size_work(mx_plan *P, int dir, mx_precision prec, mx_io_type io, int nrhs);
void ensure_work(mx_plan *P, int nrhs, size_t elem, bool need_io);   // forward skeletons (+ staging)
void ensure_io(mx_plan *P, int nrhs);                                 // FP32 staging of FP64 W and Y
void ensure_adj_work(mx_plan *P, int nrhs, size_t elem);   // skeletons of the adjoint evaluation
void ensure_panels(mx_plan *P, int nrhs);
// Packed operands and split-panel pointers of P's BF16x3 kernels for nrhs
// columns on the panel buffer `panels` (sets nrhs, ahi, alo, phi, plo).
void x3_args_of(const mx_plan *P, int nrhs, uint16_t *panels, X3Args &xa);
// Bytes of the split panels for nrhs columns, and of panels plus the sweep
// control block after them (at offset round256(panel bytes)).
size_t panel_bytes_of(const mx_plan *P, int nrhs);
size_t panel_ctl_bytes_of(const mx_plan *P, int nrhs);
// The sweep's workgroup count for an evaluation of nrhs columns (queried
// once per tile width).
int sweep_grid_for(mx_plan *P, int dir, int nrhs);
// Record what an evaluation (eager, captured or replayed) leaves in the work
// buffers: the skeletons of direction dir at nrhs columns.
void note_eval(mx_plan *P, int dir, mx_precision prec, int nrhs);
void autotune(mx_plan *P, mx_precision prec, int nrhs, mx_io_type io, int reps, hipStream_t st, float *best_ms);
// n unit-variance pseudo-random values (splitmix64 of the index), FP64 or FP32
void launch_fill_uniform(void *p, int64_t n, bool f64, hipStream_t st);
void dense_rows_core(const double *X_leaf, int64_t N, int d, double h, int nleaf, const int32_t *wrow,
                     const int32_t *Ddim, const int32_t *rows, int nrows, const double *W, int nrhs, double *out,
                     hipStream_t st);

// C-ABI error channel: run f, turn exceptions into status codes and keep the
// message for mx_last_error.
template <typename F>
int32_t guard(F &&f) {
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

}  // namespace mx
