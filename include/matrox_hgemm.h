/*
 * matrox_hgemm.h — C-ABI of the gfx950 HIP evaluator of Y = K~ W.
 *
 * Two layers:
 *
 * 1. Drop-in entry points with the exact signatures of the functions MatRox
 *    generates (HGEMM for H2-b, symGen/HGEMM_gen.h:117-124; the HSS form,
 *    symGen/HGEMM_v1.h:117-124).  All buffers are caller-owned HOST memory in
 *    the reference's CDS layout; the call copies them to the GPU, evaluates in
 *    FP64 and copies apres/wskel/uskel back, with the reference's
 *    accumulation rules (H2-b: apres and uskel accumulate; HSS: apres and the
 *    far part of uskel are overwritten).  The generated code bakes the loop
 *    bounds nl, fl, cdepth and the leaf count into its body; here they are set
 *    once by mx_hgemm_configure(), the equivalent of code generation.
 *
 * 2. A plan API for device-resident evaluation: mx_plan_create() uploads the
 *    structure and blocks once (in every precision requested), mx_eval() runs
 *    one evaluation on device buffers on a HIP stream.  W and Y use the
 *    leaf-blocked layout of the reference's TW/apres (matroxTest-V1/hcds.h:
 *    463-485): leaf l's rows start at row wrow[l], stored Dim_l x nrhs
 *    column-major at element offset wrow[l]*nrhs.
 *
 * Every function returns an mx_status (0 = success).  mx_last_error() gives a
 * message for the last failure on the calling thread.
 */
#ifndef MATROX_HGEMM_H
#define MATROX_HGEMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mx_status {
    MX_OK = 0,
    MX_ERR_ARG = -1,        /* null pointer, bad size, inconsistent arrays   */
    MX_ERR_HIP = -2,        /* HIP runtime failure                           */
    MX_ERR_STATE = -3,      /* call order (e.g. HGEMM before configure)      */
    MX_ERR_UNSUPPORTED = -4 /* precision or layout not built into the plan   */
} mx_status;

/* Arithmetic of an evaluation. */
typedef enum mx_precision {
    MX_FP64 = 0,   /* FP64 MFMA, the drop-in path                              */
    MX_FP32 = 1,   /* FP32 blocks and vectors, FP32 MFMA                        */
    MX_BF16X3 = 2  /* near and far blocks as split-BF16 x3 products on MFMA
                      with FP32 accumulation; tree transfers in FP32           */
} mx_precision;

/* Element type of W and Y in mx_eval. */
typedef enum mx_io_type { MX_IO_F64 = 0, MX_IO_F32 = 1 } mx_io_type;

/* ---- layer 1: drop-in entry points (host memory, FP64) ------------------- */

/* Loop bounds the generated HGEMM bakes in: number of leaves, near
 * row-blocks nl, far row-blocks fl and coarsened level groups cdepth
 * (symGen/HGEMM_gen.h:126,142,180,197).  Process-wide; call before HGEMM. */
mx_status mx_hgemm_configure(int32_t nleaf, int32_t nl, int32_t fl, int32_t cdepth);

/* Whether HGEMM / HGEMM_HSS keep their device plan (the uploaded structure
 * and blocks) between calls; no counterpart in the reference, whose
 * generated code reads the host arrays directly.  A call reuses the kept
 * plan when it passes the same D, B and VT pointers and structure arrays of
 * the same contents (always compared in full: they are small); W, apres,
 * uskel and nrhs may change.
 *   MX_HGEMM_CACHE_OFF     build and free a plan in every call;
 *   MX_HGEMM_CACHE_VERIFY  (default) keep it, and check every call, by a
 *                          hash of every byte of D, B and VT computed on
 *                          the host threads while the device evaluates,
 *                          that the blocks are unchanged (a changed block
 *                          rebuilds the plan and evaluates again);
 *   MX_HGEMM_CACHE_TRUST   keep it without reading D, B and VT: the caller
 *                          promises that blocks at the same addresses hold
 *                          the same values (or calls mx_hgemm_cache again,
 *                          which drops the kept plan).
 * The kept plan holds the FP64 blocks on the device until it is dropped. */
typedef enum mx_hgemm_cache_mode {
    MX_HGEMM_CACHE_OFF = 0,
    MX_HGEMM_CACHE_VERIFY = 1,
    MX_HGEMM_CACHE_TRUST = 2
} mx_hgemm_cache_mode;
mx_status mx_hgemm_cache(int32_t mode);

/* H2-b evaluation; replaces symGen/HGEMM_gen.h:117-238.  Argument meaning,
 * layout and accumulation are the reference's; returns 0 on success or a
 * negative mx_status (the reference always returns 0). */
int32_t HGEMM(double *D, double *B, double *VT, uint64_t *Dptr, uint64_t *Bptr, int32_t *VTptr,
              int32_t *lchildren, int32_t *rchildren, int32_t *levelset, int32_t *idx, double *mrhs,
              double *apres, int32_t nrhs, int32_t *Ddim, int32_t *wptr, int32_t *uptr, double *wskel,
              int32_t *wskeloffset, double *uskel, int32_t *uskeloffset, int32_t *lm, int32_t *slen,
              int32_t *nblockSet, int32_t *nblocks, int32_t *npairx, int32_t *npairy, int32_t *fblockSet,
              int32_t *fblocks, int32_t *fpairx, int32_t *fpairy, int32_t *wpart, int32_t *clevelset);

/* HSS evaluation; replaces symGen/HGEMM_v1.h:117-405 (the reference names it
 * HGEMM too and selects it with #define HSS; a build links this symbol under
 * that name, see the integration notes).  Near: diagonal leaf blocks Dptr[i];
 * far: node i with its sibling through B[Bptr[i-1]]. */
int32_t HGEMM_HSS(double *D, double *B, double *VT, uint64_t *Dptr, uint64_t *Bptr, int32_t *VTptr,
                  int32_t *lchildren, int32_t *rchildren, int32_t *levelset, int32_t *idx, double *mrhs,
                  double *apres, int32_t nrhs, int32_t *Ddim, int32_t *wptr, int32_t *uptr, double *wskel,
                  int32_t *wskeloffset, double *uskel, int32_t *uskeloffset, int32_t *lm, int32_t *slen,
                  int32_t *wpart, int32_t *clevelset);

/* ---- layer 2: plans on device memory ------------------------------------- */

/* The compressed matrix in the reference layout (host arrays; copied by
 * mx_plan_create).  Pair arrays follow the packing order of D and B.  For an
 * HSS structure (is_hss = 1) near pair k is leaf k and far pair k couples
 * node k+1 with its sibling, the order HGEMM_v1.h indexes.  Row offsets are
 * in rows: wptr = wrow*nrhs and wskeloffset = skelrow*nrhs in the
 * reference's element units (hcds.h:362-413). */
typedef struct mx_structure {
    int32_t nleaf, numnodes, is_hss;
    int64_t ncount, fcount;
    const int32_t *lchildren, *rchildren, *lm, *Ddim, *slen;
    const int32_t *npairx, *npairy;  /* leaf indices, ncount entries */
    const int32_t *fpairx, *fpairy;  /* node ids, fcount entries     */
    const uint64_t *Dptr, *Bptr;
    const int32_t *VTptr;
    const int32_t *wrow;             /* nleaf row offsets of TW/Y blocks      */
    const int32_t *skelrow;          /* numnodes row offsets of wskel/uskel   */
    const double *D, *B, *VT;
    int64_t D_len, B_len, VT_len;    /* element counts of D, B, VT            */
} mx_structure;

typedef struct mx_plan mx_plan;

/* Bits of `precisions` in mx_plan_create.  MX_PLAN_ADJOINT also builds the
 * transposed task lists (and, with MX_PLAN_BF16X3, the packed transposed D
 * and B) used by mx_eval_adjoint. */
#define MX_PLAN_FP64 (1u << MX_FP64)
#define MX_PLAN_FP32 (1u << MX_FP32)
#define MX_PLAN_BF16X3 (1u << MX_BF16X3)
#define MX_PLAN_ADJOINT (1u << 8)

/* Validate the structure, build the per-destination task lists and upload
 * blocks for each precision in `precisions` to the current HIP device. */
mx_status mx_plan_create(mx_plan **plan, const mx_structure *s, uint32_t precisions);

/* As mx_plan_create, but the near and coupling blocks are computed on the
 * device from the points instead of read from s->D and s->B (which may be
 * NULL; D_len, B_len, Dptr and Bptr still give the layout): D_k is the
 * Gaussian kernel K(x_i, x_j) = exp(-|x_i - x_j|^2 / (2 h^2)) between the
 * points of its two leaves, B_k between the skeleton points of its two nodes
 * (the reference's Fsubmatrix, matroxTest-V1/compression.h:256-360).  X: host
 * n x d row-major points in original order; lids: the N leaf-order point ids;
 * skels / skelsoffset: skeleton point ids of node v at skels[skelsoffset[v]]
 * (nskels entries).  For matrices whose FP64 blocks do not fit the host. */
mx_status mx_plan_create_generated(mx_plan **plan, const mx_structure *s, uint32_t precisions, const double *X,
                                   int64_t n, int32_t d, double h, const int32_t *lids, const int32_t *skels,
                                   const int64_t *skelsoffset, int64_t nskels);
mx_status mx_plan_destroy(mx_plan *plan);

/* Y = K~ W on device buffers (leaf-blocked, N x nrhs), enqueued on `stream`
 * (a hipStream_t; NULL = default stream).  Y is overwritten.  The plan's
 * skeleton work buffers are (re)allocated when nrhs grows; evaluations on one
 * plan must not overlap. */
mx_status mx_eval(mx_plan *plan, mx_precision prec, const void *W, void *Y, int32_t nrhs, mx_io_type io,
                  void *stream);

/* Size of the static work area of mx_eval_noalloc, per GPU (bytes). */
#define MX_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)

/* mx_eval without any device allocation: no counterpart in the reference
 * (whose generated HGEMM takes caller-owned work buffers); for callers that
 * cannot allocate device memory — code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc.  Same arguments, kernels and
 * results as mx_eval (forward direction), but the work buffers (skeletons,
 * the FP32 staging of FP64 I/O, the split panels) come from one static
 * device array of MX_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, instead of the plan's own buffers.  Needs, in bytes, about
 *     FP64:    16 * skel_rows * nrhs
 *     FP32:     8 * skel_rows * nrhs  (+ 8 * N * nrhs with FP64 I/O)
 *     BF16x3:   4 * (padded N + 2 * padded skel_rows) * nrhs
 * (each part rounded up to 256 bytes); a larger evaluation returns
 * MX_ERR_ARG with the size it needs and does nothing.  Every call on one
 * GPU uses the same array, which keeps the last call's contents: calls that
 * overlap in time on one GPU (other streams, other threads) race.  The
 * plan's own work buffers and skeleton state are not touched, so
 * mx_backward and mx_plan_copy_skeletons do not see these evaluations.
 * Launch timing (mx_plan_set_profiling) does not record them. */
mx_status mx_eval_noalloc(mx_plan *plan, mx_precision prec, const void *W, void *Y, int32_t nrhs, mx_io_type io,
                          void *stream);

/* Adjoint evaluation dW = K~^T dY (same layouts and rules as mx_eval): the
 * upward pass of dY, couplings grouped by their column node through B^T,
 * downward pass, near blocks grouped by their column leaf through D^T.  This
 * is the reference evaluation applied to the transposed structure (pair
 * lists swapped, D and B transposed, the same VT).  Needs MX_PLAN_ADJOINT. */
mx_status mx_eval_adjoint(mx_plan *plan, mx_precision prec, const void *dY, void *dW, int32_t nrhs, mx_io_type io,
                          void *stream);

/* Backward pass of Y = K~ W (no counterpart in the reference, which has no
 * derivatives): dW = K~^T dY as mx_eval_adjoint, and, for each non-NULL
 * dD / dB / dVT (device arrays of D_len / B_len / VT_len elements of the I/O
 * type, every element written), the gradients of the stored blocks,
 *     dD_k = dY_x W_y^T,  dB_k = y^_x w^_y^T,
 *     leaf dVT_v = g^_v W_v^T + u^_v dY_v^T,
 *     internal dVT_v = [g^_v w^_l^T + u^_v y^_l^T,  g^_v w^_r^T + u^_v y^_r^T],
 * where w^, u^ are the skeletons of the forward evaluation of W and
 * y^ = V^T-upward pass of dY, g^ = dL/dw^ those of the adjoint evaluation.
 * wskel / uskel: the forward skeletons as mx_plan_copy_skeletons writes them
 * (double for MX_FP64, float for MX_FP32 and MX_BF16X3); NULL takes the
 * plan's own, valid when the plan's last forward evaluation was of this W in
 * this precision.  prec is MX_FP64 (FP64 I/O), MX_FP32 or MX_BF16X3 (either
 * I/O type; with MX_BF16X3 the evaluations and the block products are
 * split-BF16x3, each operand split once into a plan work buffer).  With
 * block gradients requested, dW must not alias W or dY (MX_ERR_ARG): the
 * products read them after dW is written.  Needs MX_PLAN_ADJOINT. */
mx_status mx_backward(mx_plan *plan, mx_precision prec, const void *W, const void *dY, void *dW, int32_t nrhs,
                      mx_io_type io, const void *wskel, const void *uskel, void *dD, void *dB, void *dVT,
                      void *stream);

/* Replace the stored blocks of a plan (same structure and lengths) from
 * device FP64 arrays, re-deriving every precision built into it; NULL keeps
 * that buffer.  No counterpart in the reference (its blocks are fixed once
 * compressed); used to differentiate through the blocks. */
mx_status mx_plan_set_blocks(mx_plan *plan, const double *D, const double *B, const double *VT, void *stream);

/* Algorithmic flops of the block gradients of mx_backward with nrhs columns
 * (2*m*n*nrhs per product). */
mx_status mx_plan_grad_flops(const mx_plan *plan, int32_t nrhs, double *flops);

/* Tile of the split-BF16x3 block-gradient products of mx_backward: 128
 * (128 x 128, two workgroups per CU), 256 (256 x 256, one workgroup of 8
 * waves per CU, half the operand bytes per flop), or 0 (default: 256 when
 * its padded work is within 15% of the 128 tiles', i.e. blocks of about 256
 * rows and columns).  mx_plan_get_grad_tile gives the tile in use.  No
 * counterpart in the reference. */
mx_status mx_plan_set_grad_tile(mx_plan *plan, int32_t tile);
mx_status mx_plan_get_grad_tile(const mx_plan *plan, int32_t *tile);

/* Device pointers of the skeleton buffers of the last evaluation
 * (wskel = V^T W per node, uskel = coupled + transferred values per node),
 * laid out at skelrow[v]*nrhs with leading dimension slen[v]; element type is
 * double for MX_FP64 and float for MX_FP32.  After an MX_BF16X3 evaluation
 * the skeletons exist only as split panels and this returns
 * MX_ERR_UNSUPPORTED. */
mx_status mx_plan_skeletons(mx_plan *plan, const void **wskel, const void **uskel);

/* Copy the skeleton buffers of the last evaluation (skel_rows x nrhs
 * elements each, layout as above) into caller device buffers on `stream`.
 * After an MX_BF16X3 evaluation the copies are FP32 (hi + lo of the split
 * panels). */
mx_status mx_plan_copy_skeletons(mx_plan *plan, void *wskel_dst, void *uskel_dst, int32_t nrhs, void *stream);

/* Replay evaluations from captured hipGraphs: with enable != 0, the first
 * mx_eval with a given (precision, io, nrhs, W, Y) captures its launch
 * sequence and later calls with the same arguments replay it (a few graphs
 * are kept per plan).  While profiling is on, evaluations run eagerly.
 * enable == 0 drops the captured graphs. */
mx_status mx_plan_set_graph(mx_plan *plan, int32_t enable);

/* Schedules of the coupling and downward phases (per direction and kernel
 * kind): MX_SCHED_FUSED runs one launch per tree level computing
 * u_c = sum_far B w + VT_p^T u_p; MX_SCHED_SPLIT runs the couplings of all
 * levels in one launch, then the parent transfers level by level.  The
 * _BALANCED forms also move the leading near blocks of leaves with long near
 * lists (the adjoint's column lists) out of the final NEAR+LEAF launch into
 * the upper-level UP launches.  All give the same result up to summation
 * order.  Defaults: fused forward, split adjoint.  x3_variant: BF16x3 kernel
 * tile and pipeline depth (one per plan, 0 keeps the current): */
enum { MX_SCHED_FUSED = 0, MX_SCHED_SPLIT = 1, MX_SCHED_FUSED_BALANCED = 2, MX_SCHED_SPLIT_BALANCED = 3 };
/* MX_SCHED_SWEEP (MX_BF16X3 only): one persistent launch per evaluation runs
 * every tile of the fused schedule, each as soon as the skeleton blocks it
 * reads are written (per-block counters; a tile reads its parent's transfer
 * last, so the couplings and near blocks run while the top-down chain is in
 * progress).  x3_variant picks its tile width: MX_X3_128x128_S2,
 * MX_X3_128x64_S2 or MX_X3_128x32_S2.  The autotuner picks it when it is the
 * faster; any other schedule turns it off.  No counterpart in the reference,
 * whose sweeps synchronise per level group (symGen/HGEMM_gen.h:142-236). */
enum { MX_SCHED_SWEEP = 4 };
/* MX_SCHED_COARSE (MX_BF16X3 only; plans of stored blocks): the fused
 * schedule with runs of consecutive launches merged into one.  A segment
 * A_c x_P over a skeleton block x_P = sum_s A_s B_s written earlier in the
 * same run becomes the segments (A_c A_s) B_s, over composite blocks A_c A_s
 * (A_s itself a composite when P was rewritten so) formed in FP64 when the
 * plan is built (and again by mx_plan_set_blocks).  A launch joins the run
 * when that adds at most 4.5 MFLOP per right-hand side column and does not
 * lengthen the run's longest task by more than its own longest task plus 4
 * chunks of 32.  Fewer launches for the latency-bound top of the tree,
 * every skeleton block still written.  The plan keeps FP64 copies of B and
 * VT for it.  Like the reference's coarsened level groups
 * (symGen/HGEMM_gen.h:142-178, 196-236), with the group boundaries removed
 * by re-association instead of barriers.  MX_SCHED_COARSE_SPLIT: the same
 * over the MX_SCHED_SPLIT lists (whose per-level transfers accumulate, so
 * only the UP launches and the couplings merge). */
enum { MX_SCHED_COARSE = 5, MX_SCHED_COARSE_SPLIT = 6 };

/* Error word of the last sweep this plan enqueued (after synchronising
 * `stream`): 0 = none or clean; 1 = a wait for a producer, 2 = claiming
 * a work item, exceeded one second, after which the launch drained without
 * waiting (its outputs are then incomplete).  No counterpart in the
 * reference. */
mx_status mx_plan_sweep_status(const mx_plan *plan, void *stream, int32_t *error);

/* Host-only: the BF16x3 evaluation's task lists of one direction and
 * schedule (MX_SCHED_FUSED ... MX_SCHED_SPLIT_BALANCED, MX_SCHED_COARSE or
 * MX_SCHED_COARSE_SPLIT)
 * of a structure, in terms of the stored blocks and the reference layouts
 * (no device work, for checks of the schedules).  With launches == NULL it
 * returns the counts.  launches: nlaunch + 1 offsets into ltasks, the tasks
 * of each launch in order; tasks: 6 per task (c_row, m, c_buf, seg_begin,
 * seg_end, beta); segs: 8 per segment (a_off, b_row, a_buf, a_trans, lda,
 * k, b_buf, b_ld), a_buf 3 naming composite a_off; comps: 11 per composite
 * (a1_off, a2_off, a1_buf, a1_trans, a1_lda, a2_buf, a2_trans, a2_lda, m,
 * kmid, n).  Buffers: a_buf 0/1/2 = D/B/VT, c_buf/b_buf 0/1/2/3 = W, w^,
 * u^, Y; a block of rows r at row offset c_row (b_row) is r x nrhs
 * column-major at element c_row * nrhs.  No counterpart in the reference. */
mx_status mx_lists_describe(const mx_structure *s, int32_t direction, int32_t schedule, int32_t *nlaunch,
                            int32_t *nltask, int32_t *ntask, int32_t *nseg, int32_t *ncomp, int64_t *launches,
                            int64_t *ltasks, int64_t *tasks, int64_t *segs, int64_t *comps);

/* Host-only: the sweep's work lists of one direction of a structure (no
 * device work, for checks of the scheduling protocol).  With tiles == NULL
 * it returns the counts.  tiles: 12 int32 per tile (task, row tile, rank,
 * 0, the last producer tile per queue x 8), queues back to back; qinfo: per
 * queue its first tile and tile count (16); tasks: 8 per task (early waits
 * [b, e), late waits [b, e), index of the first u^ segment or -1, counter
 * block signalled or -1, 0, 0); waits: 2 per wait (counter block, target
 * count of finished (tile, wave) writes per column tile).  Counter block v
 * is the w^ block of node v, numnodes + v its u^ block.  No counterpart in
 * the reference. */
mx_status mx_sweep_describe(const mx_structure *s, int32_t direction, int32_t *ntiles, int32_t *ntasks,
                            int32_t *nwaits, int32_t *tiles, int32_t *qinfo, int32_t *tasks, int32_t *waits);
enum { MX_DIR_FORWARD = 0, MX_DIR_ADJOINT = 1 };
enum {
    MX_X3_128x128_S2 = 1, /* 4 waves, 2 stages, two workgroups per CU (default) */
    MX_X3_128x128_S4 = 2, /* 4 waves, 4 stages */
    MX_X3_128x256_S3 = 3, /* 8 waves, 3 stages */
    MX_X3_128x128_REG = 4, /* 4 waves, fragments staged through VGPRs */
    MX_X3_128x64_S2 = 5,   /* 4 waves of 64 x 32, for small nrhs */
    MX_X3_128x32_S2 = 6,   /* 4 waves of 64 x 16, for small nrhs */
    MX_X3_128x128_W8S3 = 7, /* 8 waves of 64 x 32, 3 stages */
    MX_X3_128x128_W8S4 = 8, /* 8 waves of 64 x 32, 4 stages */
    MX_X3_256x128_W8S3 = 9, /* 8 waves of 64 x 64, 3 stages, 256-row tiles */
    MX_X3_128x32_S6 = 10,   /* 4 waves of 64 x 16, 6 stages (latency-bound launches) */
    MX_X3_128x64_S5 = 11,   /* 4 waves of 64 x 32, 5 stages */
    MX_X3_128x32_K8 = 12,   /* K split over 8 waves (latency-bound launches) */
    MX_X3_256x256_W8S2 = 13 /* 8 waves of 64 x 128, 2 stages, 256-row tiles */
};
mx_status mx_plan_set_schedule(mx_plan *plan, int32_t direction, mx_precision prec, int32_t schedule,
                               int32_t x3_variant);
mx_status mx_plan_get_schedule(const mx_plan *plan, int32_t direction, mx_precision prec, int32_t *schedule,
                               int32_t *x3_variant);

/* The BF16x3 kernel variant of each launch of one direction (in launch
 * order, under the direction's current schedule). */
mx_status mx_plan_launch_variants(const mx_plan *plan, int32_t direction, int32_t *variants, int32_t cap,
                                  int32_t *count);

/* Fix the BF16x3 kernel variant of each launch of one direction under its
 * current schedule (count = number of launches, as mx_plan_launch_variants
 * reports; count = 0 returns to the plan-wide variant).  No counterpart in the
 * reference; lets a profiled run replay the autotuner's choices. */
mx_status mx_plan_set_launch_variants(mx_plan *plan, int32_t direction, const int32_t *variants, int32_t count);

/* The same for the FP64 / FP32 kernel (prec MX_FP64 or MX_FP32): per launch
 * 1 = 64 x 64 tiles with one chunk of loads in flight, 2 = 64 x 64 with two,
 * 3 = 128 x 128 tiles (FP32 only), 4 / 5 = 128 x 128 tiles on A operands
 * packed when the plan is built (FP64: two / three LDS stages; FP32: two /
 * four), 6 / 7 = the same with 128 x 256 tiles (each stored block read
 * half as often); a launch with a task of more than 256 segments runs variant 1
 * instead, and so does every launch of a plan whose packed copies did not
 * fit in device memory beside its blocks, with 16 GiB + 10% of the device
 * to spare, when it was created.  The autotuner picks them launch by
 * launch.  No counterpart in the reference. */
mx_status mx_plan_seg_variants(const mx_plan *plan, int32_t direction, mx_precision prec, int32_t *variants,
                               int32_t cap, int32_t *count);
mx_status mx_plan_set_seg_variants(mx_plan *plan, int32_t direction, mx_precision prec, const int32_t *variants,
                                   int32_t count);

/* Autotuner: times every schedule (and, for MX_BF16X3, every kernel
 * variant, first for the whole evaluation, then launch by launch, for the
 * schedule picked and for the coarse schedules) of each direction built into the
 * plan with at least `reps` evaluations of nrhs columns (more for short
 * evaluations: about 2 ms per measurement, at most 20) on scratch buffers
 * it allocates and frees, and keeps the fastest.  best_ms (optional, 2
 * entries) receives the chosen per-direction times.  Synchronises
 * `stream`; drops captured graphs. */
mx_status mx_plan_autotune(mx_plan *plan, mx_precision prec, int32_t nrhs, mx_io_type io, int32_t reps,
                           float *best_ms, void *stream);

/* ---- per-result checks and measurement ------------------------------------ */

/* Attach the points (host, N x d row-major, in LEAF order: row j is point
 * lids[j]) and the Gaussian bandwidth h, enabling mx_dense_rows. */
mx_status mx_plan_set_points(mx_plan *plan, const double *X_leaf, int32_t d, double h);

/* Dense FP64 reference rows: out (device, nrows x nrhs column-major) =
 * K(rows, :) W with K(i,j) = exp(-|x_i - x_j|^2 / (2 h^2)) evaluated directly
 * (matroxTest-V1/Util.h:182-215, 303-308).  rows are host leaf-order row
 * indices; W is the FP64 leaf-blocked input on device. */
mx_status mx_dense_rows(mx_plan *plan, const int32_t *rows, int32_t nrows, const double *W, int32_t nrhs,
                        double *out, void *stream);

/* Launch timing: when enabled, every kernel launch of mx_eval, mx_eval_adjoint
 * and mx_backward is bracketed
 * by HIP events on the evaluation stream.  mx_plan_launch_times waits for the
 * last evaluation and returns, per launch, its duration in ms and its kernel
 * kind (MX_KIND_*); *count is the number of launches (<= cap filled).
 * MX_KIND_SEG_F64 / MX_KIND_SEG_F32 are every FP64 / FP32 evaluation launch,
 * on the 64 x 64 kernel or on the packed-operand kernels alike. */
enum {
    MX_KIND_SEG_F64 = 0,
    MX_KIND_SEG_F32 = 1,
    MX_KIND_BF16X3 = 2,
    MX_KIND_CONVERT = 3,
    MX_KIND_GRAD_F64 = 4,
    MX_KIND_GRAD_F32 = 5,
    MX_KIND_GRAD_X3 = 6
};
mx_status mx_plan_set_profiling(mx_plan *plan, int32_t enable);
mx_status mx_plan_launch_times(mx_plan *plan, float *ms, int32_t *kind, int32_t cap, int32_t *count);

/* Algorithmic flops of one evaluation with nrhs columns (the count of
 * matroxTest-V1/compression.h:1079-1130), split by kernel kind of the given
 * precision: flops[MX_KIND_*] (4 entries); returns the total in *total. */
mx_status mx_plan_flops(const mx_plan *plan, mx_precision prec, int32_t nrhs, double *flops, double *total);

/* Algorithmic flops of each launch of one evaluation (in launch order, the
 * order mx_plan_launch_times reports kernel launches after the conversion).
 * Under MX_SCHED_COARSE: the flops each launch performs, composite products
 * included (mx_plan_flops still gives the algorithmic total). */
mx_status mx_plan_launch_flops(const mx_plan *plan, mx_precision prec, int32_t nrhs, double *flops, int32_t cap,
                               int32_t *count);

/* Number of rows N and of skeleton rows sum(slen) of a plan. */
mx_status mx_plan_dims(const mx_plan *plan, int64_t *N, int64_t *skel_rows);

/* ---- many independent instances ------------------------------------------ */

/* A batch evaluates `count` independent compressed matrices, instance i with
 * nrhs[i] columns (no counterpart in the reference, which evaluates one
 * matrix per process).  Instances with equal nrhs form a group: their trees
 * are joined under rank-0 nodes into one tree whose plan covers a tree level
 * of every member in each launch; groups run concurrently on streams forked
 * from the caller's.  A group's W and Y are single device buffers of
 * rows x nrhs elements in which instance i occupies the leaf-blocked rows
 * [row_offset, row_offset + N_i) (element offset row_offset * nrhs).
 * precisions as for mx_plan_create; with MX_PLAN_ADJOINT every group can
 * also be evaluated transposed (mx_batch_eval_adjoint). */
typedef struct mx_batch mx_batch;
mx_status mx_batch_create(mx_batch **batch, const mx_structure *const *instances, const int32_t *nrhs, int32_t count,
                          uint32_t precisions);

/* Split-BF16x3 accuracy guard of a batch.  When `precisions` holds both
 * MX_PLAN_BF16X3 and MX_PLAN_FP32, mx_batch_create estimates every
 * instance's relative error in BF16x3: each group is evaluated once in
 * BF16x3 and in FP64 (else FP32) on a unit-variance pseudo-random W with its
 * nrhs, forward and, with MX_PLAN_ADJOINT, adjoint.  Instances whose
 * estimate exceeds MX_BATCH_X3_GUARD are put in groups of their own that
 * evaluate in FP32 whenever BF16x3 is requested (their plans hold no split
 * blocks), so that every instance stays within 1e-5 of FP64.  The error of
 * the split products grows with the depth of the tree and the weight of the
 * far field: small, smooth instances can exceed 1e-5.  Add
 * MX_BATCH_NO_X3_GUARD to `precisions` to switch the guard off. */
#define MX_BATCH_X3_GUARD 8e-6
#define MX_BATCH_NO_X3_GUARD (1u << 9)
/* The guard's estimate for an instance (-1 when no estimate was made) and
 * whether its group evaluates in FP32 when BF16x3 is requested. */
mx_status mx_batch_instance_guard(const mx_batch *batch, int32_t instance, double *x3_err, int32_t *fp32);
mx_status mx_batch_destroy(mx_batch *batch);
mx_status mx_batch_groups(const mx_batch *batch, int32_t *ngroups);
mx_status mx_batch_group_info(const mx_batch *batch, int32_t group, int32_t *nrhs, int64_t *rows);
mx_status mx_batch_layout(const mx_batch *batch, int32_t instance, int32_t *group, int64_t *row_offset);
/* The plan of a group (owned by the batch): schedules, profiling, flops. */
mx_status mx_batch_plan(mx_batch *batch, int32_t group, mx_plan **plan);
/* Y[g] = K~_g W[g] for every group g (arrays of ngroups device pointers). */
mx_status mx_batch_eval(mx_batch *batch, mx_precision prec, const void *const *W, void *const *Y, mx_io_type io,
                        void *stream);
/* dW[g] = K~_g^T dY[g] for every group (instance by instance, the adjoint of
 * mx_batch_eval in the same layout); needs MX_PLAN_ADJOINT.  No counterpart
 * in the reference. */
mx_status mx_batch_eval_adjoint(mx_batch *batch, mx_precision prec, const void *const *dY, void *const *dW,
                                mx_io_type io, void *stream);
/* Backward pass of every group (mx_backward on each group's plan): dW[g] =
 * K~_g^T dY[g] and, where dD / dB / dVT (arrays of ngroups device pointers,
 * or NULL) give a buffer, the gradients of the group's stored blocks, from
 * the skeletons of the batch's last forward evaluation in this precision
 * (of the same W).  A group's block buffers hold its instances' blocks one
 * after another (mx_batch_group_blocks gives their lengths,
 * mx_batch_block_layout an instance's first elements), in each instance's
 * own CDS layout.  Needs MX_PLAN_ADJOINT.  No counterpart in the
 * reference. */
mx_status mx_batch_backward(mx_batch *batch, mx_precision prec, const void *const *W, const void *const *dY,
                            void *const *dW, mx_io_type io, void *const *dD, void *const *dB, void *const *dVT,
                            void *stream);
mx_status mx_batch_group_blocks(const mx_batch *batch, int32_t group, int64_t *D_len, int64_t *B_len,
                                int64_t *VT_len);
mx_status mx_batch_block_layout(const mx_batch *batch, int32_t instance, int64_t *d_off, int64_t *b_off,
                                int64_t *vt_off);
/* Capture the whole batch (all groups and their fork/join) as one hipGraph
 * per (direction, precision, io, W, Y) and replay it. */
mx_status mx_batch_set_graph(mx_batch *batch, int32_t enable);
/* mx_plan_autotune on every group with its own nrhs; with MX_BF16X3, also
 * the batched sweep below (on or off, and its tile width) for the whole
 * batch. */
mx_status mx_batch_autotune(mx_batch *batch, mx_precision prec, mx_io_type io, int32_t reps, void *stream);

/* Batched sweep: every BF16x3 group of the batch (all nrhs) evaluated by one
 * dependency-driven launch (MX_SCHED_SWEEP's protocol over the union of
 * the groups' work lists), after each group's W split; FP32-routed groups
 * run beside it.  x3_variant = tile width (MX_X3_128x128_S2,
 * MX_X3_128x64_S2 or MX_X3_128x32_S2), 0 = off (every group runs its own
 * plan's launches, concurrently).  Per direction; drops captured graphs.
 * No counterpart in the reference, which evaluates one matrix per
 * process (matroxTest-V1/main.cpp:22-115). */
mx_status mx_batch_set_sweep(mx_batch *batch, int32_t direction, int32_t x3_variant);
mx_status mx_batch_get_sweep(const mx_batch *batch, int32_t direction, int32_t *x3_variant);
/* Error word of the batched sweeps (as mx_plan_sweep_status). */
mx_status mx_batch_sweep_status(const mx_batch *batch, void *stream, int32_t *error);
mx_status mx_batch_flops(const mx_batch *batch, mx_precision prec, double *total);
/* Per-instance dense check rows, as mx_plan_set_points / mx_dense_rows; W is
 * the instance's FP64 block inside its group buffer, out is nrows x nrhs. */
mx_status mx_batch_set_points(mx_batch *batch, int32_t instance, const double *X_leaf, int32_t d, double h);
mx_status mx_batch_dense_rows(mx_batch *batch, int32_t instance, const int32_t *rows, int32_t nrows, const double *W,
                              double *out, void *stream);

const char *mx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MATROX_HGEMM_H */
