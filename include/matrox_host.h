/*
 * matrox_host.h — host-side structure builder for the H-matrix evaluator.
 *
 * Builds, on the CPU and in FP64, everything the evaluation phase consumes:
 * the cluster tree, the near/far interaction lists, the coarsened level sets,
 * the near/far blocking, the kNN-sampled interpolative decomposition (ID) and
 * the packed "CDS" arrays.  The array set and its index semantics are those of
 * the MatRox runtime (matroxTest-V1/HTree.h:17-74 struct HTree, and
 * matroxTest-V1/hcds.h:12-34 struct hcds); each array is fetched by the name
 * the reference's generated HGEMM gives its argument (symGen/HGEMM_gen.h:121-124).
 *
 * All functions are plain C: pointers and sizes, no C++ or torch types.
 * Errors are reported through return codes (0 = success) and mxh_last_error().
 */
#ifndef MATROX_HOST_H
#define MATROX_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel function of the compressed matrix.  Only the Gaussian kernel
 * exp(-|x-y|^2 / (2 h^2)) is on the evaluated path
 * (matroxTest-V1/compression.h:256-312, Internal::KS_GAUSSIAN). */
enum { MXH_KERNEL_GAUSSIAN = 0 };

/* Build parameters.  Every constant the reference hard-codes is a field here:
 * leaf size m (sympiler/HTree.cpp:34), kNN k (HTree.cpp:35), rank cap
 * (compression.h:449), ID tolerance (compression.h:469), near budget
 * (HTree.cpp:46-48), bandwidth h (main.cpp:68,84), near/far block sizes
 * (sympiler/DSInspector.cpp:179-181), level aggregation parlevel(2,2)
 * (DSInspector.cpp:92) and the bin count of BalanceCoarLevelSet, which the
 * reference takes from omp_get_max_threads() (compression.h:1011). */
typedef struct mxh_config {
    int32_t  leaf_size;      /* m: a node with more points is split          */
    int32_t  max_rank;       /* skeleton rank cap                            */
    double   tol;            /* relative ID tolerance |R_ss/R_00| < tol      */
    double   budget;         /* near budget tau; 0 gives an HSS structure    */
    double   h;              /* Gaussian bandwidth                           */
    int32_t  kernel;         /* MXH_KERNEL_*                                  */
    int32_t  knn_k;          /* neighbours per point for sampling            */
    int32_t  near_block;     /* leaves per near row-block (2)                */
    int32_t  far_block;      /* nodes per far row-block (4)                  */
    int32_t  agg_levels;     /* levels merged per coarsened group (2)        */
    int32_t  bottom_levels;  /* levels in the bottom coarsened group (2)     */
    int32_t  nbins;          /* partitions per coarsened group (12)          */
    int32_t  nthreads;       /* OpenMP threads for the build, 0 = default    */
    uint64_t seed;           /* seed of the sampling random streams          */
    int32_t  skip_blocks;    /* 1: lay out D and B (Dptr, Bptr, sizes in the
                                scalars "D_len", "B_len") without computing
                                them, for evaluators that generate them      */
    int32_t  all_rows;       /* 1: factor every node's ID on all n points
                                instead of the sampled rows (a control for
                                the sampling's accuracy; O(n) rows a node)   */
} mxh_config;

/* Fill a config with the reference's defaults (m=256, rank 256, tol 1e-5,
 * budget 0.03, h=5, k=32, blocks 2/4, parlevel(2,2), 12 bins). */
void mxh_config_default(mxh_config *cfg);

typedef struct mxh_handle mxh_handle;

/* Build tree, lists, sampling, compression and CDS arrays from points X
 * (n x d, row-major, as read by sympiler/HTree.cpp:51-63).  The handle keeps
 * its own copy of X.  Returns NULL on failure (see mxh_last_error()). */
mxh_handle *mxh_build(const double *X, int32_t n, int32_t d, const mxh_config *cfg);

/* Build only the structure (tree, lists, coarsening, blocking) and the kNN
 * table; no compression and no D/B/VT blocks.  Used for structure-only tests. */
mxh_handle *mxh_build_structure(const double *X, int32_t n, int32_t d, const mxh_config *cfg);

void mxh_free(mxh_handle *h);

/* Scalar properties: "n", "d", "nleaf", "numnodes", "depth", "cdepth", "nl",
 * "fl", "ncount", "fcount", "is_hss", "N_total", "skel_rows", "D_len",
 * "B_len", "VT_len".  Returns -1 if unknown. */
int64_t mxh_get_scalar(const mxh_handle *h, const char *name);

/* Named arrays (pointer stays valid until mxh_free).  int32 arrays:
 *   VTptr lchildren rchildren levelset idx_level idx(=cidx) Ddim lm slen
 *   nblockSet nblocks npairx npairy fblockSet fblocks fpairx fpairy
 *   wpart clevelset lids lidsoffset lidslen leaf proj_column
 *   skels skelsoffset flat_nx flat_ny flat_fx flat_fy
 *   id_rows   (per node: rows of the block its ID factored)
 *   wrow urow skelrow   (row offsets: multiply by nrhs for wptr/uptr/wskeloffset)
 * uint64 arrays: Dptr Bptr
 * double arrays: D B VT X
 *   id_ratio  (per node: |R_ss / R_00| of the first column the ID left out,
 *              with s its rank; 0 when it kept every column)
 * For an HSS structure (budget giving one near leaf per leaf) D/B and the
 * pair arrays are in the flat order the HSS evaluator indexes (Dptr[leaf],
 * Bptr[node-1]) and the blocked arrays describe that order as one row-block.
 * Returns 0 and sets ptr and len, or a negative code if the name is unknown. */
int mxh_get_i32(const mxh_handle *h, const char *name, const int32_t **ptr, int64_t *len);
int mxh_get_u64(const mxh_handle *h, const char *name, const uint64_t **ptr, int64_t *len);
int mxh_get_f64(const mxh_handle *h, const char *name, const double **ptr, int64_t *len);

/* Leaf-blocked permutation of a column-major n x nrhs matrix W into TW
 * (matroxTest-V1/hcds.h:463-485) and its inverse (TW -> W order). */
int mxh_transform(const mxh_handle *h, const double *W, int32_t nrhs, double *TW);
int mxh_untransform(const mxh_handle *h, const double *TW, int32_t nrhs, double *W);

/* Synthetic inputs (splitmix64 + Box-Muller, identical on every platform).
 * Mixture of `ncomp` Gaussians in d dimensions: means ~ U[-mean_range, mean_range]^d;
 * per-component standard deviations: isotropic sigma ~ U[sig_lo, sig_hi] when
 * diag == 0, or one sigma per (component, dimension) ~ U[sig_lo, sig_hi] when
 * diag != 0.  Components are drawn uniformly per point.  When standardise != 0
 * every feature is z-scored.  X is n x d row-major. */
int mxh_gaussian_mixture(double *X, int64_t n, int32_t d, int32_t ncomp, double mean_range,
                         double sig_lo, double sig_hi, int32_t diag, int32_t standardise,
                         uint64_t seed);
/* i.i.d. N(0,1) values. */
int mxh_normal(double *out, int64_t count, uint64_t seed);
int mxh_normal_f32(float *out, int64_t count, uint64_t seed);

/* The reference's on-disk structure (one raw array per file: levelset.bin,
 * idx.bin, clevelset.bin, wpart.bin, cidx.bin, lchildren.bin, rchildren.bin,
 * nblockset.bin, nblocks.bin, nxval.bin, nyval.bin, fblockset.bin,
 * fblocks.bin, fxval.bin, fyval.bin, dim.bin, leafmap.bin, lids.bin,
 * lidsoffset.bin, lidslen.bin, NN.bin, distNN.bin), as written by
 * savetree2disk/SaveSampling (sympiler/HTree.cpp:234-249, 469-542) and read by
 * readHtree (matroxTest-V1/HTree.h:76-260).  mxh_save_bin writes the set of a
 * handle into an existing directory; mxh_load_bin reads a set (for example
 * one written by the reference), checks it, and compresses and packs it with
 * cfg over the points X; the saved pair order, level sets and partitions are
 * kept.  An HSS set has no blocking files. */
int mxh_save_bin(const mxh_handle *h, const char *dir);
mxh_handle *mxh_load_bin(const char *dir, const double *X, int32_t n, int32_t d, const mxh_config *cfg);

const char *mxh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MATROX_HOST_H */
