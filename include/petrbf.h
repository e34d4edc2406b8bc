/*
 * petrbf.h -- C ABI of the MI355X (gfx950) Gaussian-RBF interpolation library.
 *
 * Drop-in boundary for barbagroup/petrbf's solve path.  Each entry point names
 * the reference interface it replaces (paths relative to the reference tree).
 * Plain pointers and sizes only; no PETSc, MPI or torch types.
 *
 * Conventions
 *   - Every function returns a prbf_status (0 = PRBF_OK).  prbf_last_error()
 *     gives the message of the last failure on that context.  Null vectors,
 *     null coordinate arrays and out-of-range solver options are rejected
 *     with PRBF_ERR_ARG before anything is copied or launched; a failed host
 *     allocation inside a call returns PRBF_ERR_NOMEM.  No C++ exception
 *     crosses the boundary.
 *   - "sorted" arrays are in cluster order (particles counting-sorted by cell,
 *     2d/get_cluster.h:77-138), the order of the reference's MatShell vectors;
 *     "original" arrays are in the caller's order.
 *   - Host-pointer functions copy in, run on the context's stream and
 *     synchronise.  *_dev functions take device pointers, are ordered on the
 *     context's stream and return without synchronising.
 *   - The caller owns all memory it passes; the context owns its device
 *     memory.  One context per host thread (thread-compatible, not
 *     thread-safe), like the reference's BOTH context (2d/par.h:148-151).
 */
#ifndef PETRBF_H
#define PETRBF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct prbf_ctx prbf_ctx;

typedef enum {
  PRBF_OK = 0,
  PRBF_ERR_ARG = 1,      /* invalid argument or size */
  PRBF_ERR_HIP = 2,      /* HIP runtime error (device fault, launch failure) */
  PRBF_ERR_NOMEM = 3,    /* device allocation failed */
  PRBF_ERR_STATE = 4,    /* call out of order (e.g. solve before set_points) */
  PRBF_ERR_FACTOR = 5,   /* reserved: RASM blocks are factorised with regularised pivots
                            (prbf_solve_info.n_regularised), not rejected */
  PRBF_ERR_OUTSIDE = 6,  /* a point fell outside the cell grid */
  PRBF_ERR_LIMIT = 7     /* a block exceeds the kernel's supported size */
} prbf_status;

/* Cluster geometry.  The reference takes nsigma_box, sigma_buffer and
 * sigma_trunc as ints in units of sigma (2d/par.h:72-74,
 * 2d/rbf_interpolation.cxx:25-26); they are reals here with the same meaning:
 * cell side = nsigma_box*sigma + 1e-6 (2d/get_cluster.h:29), buffer box
 * half-width = sigma_buffer*sigma/2 + 1e-6 (2d/get_buffer.h:11), truncation
 * box half-width = sigma_trunc*sigma/2 + 1e-6 (2d/get_trunc.h:14). */
typedef struct {
  int dim;              /* 2 or 3 */
  double sigma;
  double nsigma_box;
  double sigma_buffer;
  double sigma_trunc;
} prbf_params;

typedef enum { PRBF_KSP_GMRES = 0, PRBF_KSP_FGMRES = 1 } prbf_ksp;
/* RASM_FP64: exact FP64 block solves (the reference's -sub_pc_type lu).
 * RASM_FP32: blocks of 32 or more unknowns factorised and solved in FP32 on
 * the matrix cores, smaller ones in FP64, all applied with FP64
 * accumulation; meant for FGMRES. */
/* LEARNED: the network of prbf_learned_attach() on the clusters it accepted,
 * exact FP32 blocks on the others; FGMRES only. */
typedef enum { PRBF_PC_NONE = 0, PRBF_PC_RASM_FP64 = 1, PRBF_PC_RASM_FP32 = 2, PRBF_PC_LEARNED = 3 } prbf_pc;
typedef enum { PRBF_TUNE_CHOL_WAVES = 1, PRBF_TUNE_CHOL_PANEL = 2, PRBF_TUNE_CHOL_ORDER = 3 } prbf_tune_key;
/* Values of PRBF_TUNE_CHOL_ORDER: right-looking (trailing updates) or
 * left-looking (each panel updated from the finished L rows). */
typedef enum { PRBF_CHOL_RIGHT = 1, PRBF_CHOL_LEFT = 2 } prbf_chol_order;

/* Solver options; prbf_solve_opts_default() gives the reference's run line
 * (-ksp_type gmres default restart 30, -pc_type asm -sub_pc_type lu,
 * 2d/Makefile:7) with rtol 1e-10 and max_it 1000. */
typedef struct {
  int ksp;          /* prbf_ksp: GMRES = left preconditioning (PETSc default) */
  int pc;           /* prbf_pc */
  int restart;      /* Krylov subspace size per cycle (30) */
  int max_it;       /* iteration cap (1000) */
  double rtol;      /* relative tolerance on the solver's residual norm */
  double abstol;    /* 1e-50 */
  double dtol;      /* 1e5 */
  int graph;        /* 1 (default): each Arnoldi step (preconditioner, matvec,
                       Gram-Schmidt, Givens update) runs as one captured HIP
                       graph, one host wait per iteration; 0: kernel by kernel.
                       Both give the same bits. */
} prbf_solve_opts;

typedef struct {
  int its;           /* iterations (KSPGetIterationNumber, 2d/rbf_interpolation.cxx:234) */
  int reason;        /* PETSc KSPConvergedReason numbering: 2 rtol, 3 atol, <0 diverged */
  double rnorm0;     /* initial residual norm of the convergence test */
  double rnorm;      /* final residual norm of the convergence test */
  double t_setup_s;  /* RASM setup wall time (0 if it was already set up) */
  double t_solve_s;  /* Krylov iterations wall time */
  double t_matvec_s; /* summed device time of the matvec kernels (0 with graph = 1) */
  double t_pc_s;     /* summed device time of the preconditioner applies (0 with graph = 1) */
  int n_matvec;
  int n_pc;
  int n_regularised; /* RASM blocks whose Cholesky met a pivot at or below tau times the
                        diagonal 1/(2 pi sigma^2) (tau = 2^-14 for FP32 blocks, 2^-40 for
                        FP64) and were factorised again as A_c + s I, s = 2^-12, 2^-8,
                        2^-4 (FP32) or 2^-36, 2^-28, 2^-20 (FP64) times the diagonal,
                        until no pivot did (after the last, the floor replaces such
                        pivots): numerically singular blocks of clustered points; 0 on
                        well-separated points */
  int n_network;     /* LEARNED: clusters served by the network at the end of the solve */
  int n_net_fallback; /* LEARNED: clusters the solve moved from the network to exact blocks */
} prbf_solve_info;

/* ---- lifetime -------------------------------------------------------- */
void prbf_solve_opts_default(prbf_solve_opts* o);
prbf_status prbf_create(prbf_ctx** ctx, const prbf_params* p);
prbf_status prbf_destroy(prbf_ctx* ctx);
const char* prbf_last_error(const prbf_ctx* ctx);
/* Use an existing HIP stream (hipStream_t) for all work of this context. */
prbf_status prbf_set_stream(prbf_ctx* ctx, void* stream);
prbf_status prbf_synchronize(prbf_ctx* ctx);

/* ---- kernel (a): cell list, sort, buffer and truncation lists --------
 * Replaces Get_cluster::get_cluster (2d/get_cluster.h:15-276) and the list
 * builders Get_buffer::get_buffer (2d/get_buffer.h:9-90) and
 * Get_trunc::get_trunc (2d/get_trunc.h:9-87), built once on the device.
 * Targets (i-set) and sources (j-set); xj == NULL makes the source set the
 * target set (the solve, 2d/rbf_interpolation.cxx:195-201).  The grid spans
 * the union bounding box (2d/vorticity_evaluation.cxx:59-70). */
prbf_status prbf_set_points(prbf_ctx* ctx, int64_t ni, const double* xi, const double* yi, const double* zi,
                            int64_t nj, const double* xj, const double* yj, const double* zj);
prbf_status prbf_set_points_dev(prbf_ctx* ctx, int64_t ni, const double* xi, const double* yi, const double* zi,
                                int64_t nj, const double* xj, const double* yj, const double* zj);

/* Adaptive cluster structure over one point set (the solve), for strongly
 * non-uniform particle sets.  No reference counterpart: it generalises
 * Get_cluster::get_cluster's uniform grid (2d/get_cluster.h:20-166) and the
 * list builders (2d/get_buffer.h:9-90, 2d/get_trunc.h:9-87).  The grid of
 * prbf_set_points is level 0; a cell holding more than `cap` particles splits
 * into 2^dim children of half its side (child 2 qx + qy, 3D 4 qz + 2 qx + qy,
 * q = coordinate >= the cell centre), recursively down to depth max_depth
 * (0..20); the leaves, in depth-first order with empty ones kept, are the
 * clusters, and the particles are sorted by leaf, then original index.  A
 * leaf at depth d has buffer half-width buffer_len 2^-d (blocks keep their
 * shape relative to the cluster) and truncation half-width
 * (trunc_len - box/2) + box 2^-(d+1) (the same cutoff beyond the cluster);
 * its lists are its own particles, then those of the level-0 cells within the
 * reference's reach of its cell, in the reference's scan order, except its
 * own, that pass the per-axis box test around the leaf centre.  With no split
 * everything equals prbf_set_points.  Every later call (matvec, RASM, solve,
 * evaluation) works on the leaves; prbf_get_cells and the CSR getters return
 * per-leaf arrays and info[0] the number of leaves. */
prbf_status prbf_set_points_adaptive(prbf_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                                     int cap, int max_depth);
prbf_status prbf_set_points_adaptive_dev(prbf_ctx* ctx, int64_t n, const double* x, const double* y,
                                         const double* z, int cap, int max_depth);
/* Per leaf of the adaptive structure: centre, depth, level-0 cell, buffer and
 * truncation half-widths (any pointer may be NULL; zc ignored in 2D). */
prbf_status prbf_get_leaves(prbf_ctx* ctx, double* xc, double* yc, double* zc, int32_t* depth, int32_t* cell0,
                            double* blen, double* tlen);

/* info[0..15]: ncell (clusters: grid cells, or leaves), nx, ny, nz, nb_buffer, nb_trunc, buffer_nnz, trunc_nnz,
 * max_buffer, max_core, max_trunc, ni, nj, and the wavefronts, panel width
 * and order (prbf_chol_order) the last block factorisation used (0 before
 * any).
 * geom[0..7]: xmin, ymin, zmin, box, buffer half-width, trunc half-width, the
 * device time (s) of the last setup's block-factorisation launches and their
 * algorithmic FLOPs, sum over blocks of n^3/3 + n^2 T + T^3 (Cholesky, the
 * backward substitution of the T core right-hand sides through all n rows and
 * their forward substitution, which starts at the core, ordered last). */
prbf_status prbf_get_info(prbf_ctx* ctx, int64_t* info, double* geom);
/* Sorted position -> original index (the reference's particle->ii / jj). */
prbf_status prbf_get_permutation(prbf_ctx* ctx, int32_t* perm_i, int32_t* perm_j);
/* Per-cell ranges in sorted order (cluster->ista/iend/jsta/jend). */
prbf_status prbf_get_cells(prbf_ctx* ctx, int32_t* ista, int32_t* iend, int32_t* jsta, int32_t* jend);
/* Buffer lists (cluster->idx, the PCASM `is`): CSR over cells, sorted i-set ids. */
prbf_status prbf_get_buffer_csr(prbf_ctx* ctx, int64_t* ptr, int32_t* idx);
/* Truncation lists (the sources get_trunc gathers): CSR over cells, sorted j-set ids. */
prbf_status prbf_get_trunc_csr(prbf_ctx* ctx, int64_t* ptr, int32_t* idx);
/* Sorted coordinates of the i-set (z may be NULL in 2D). */
prbf_status prbf_get_sorted_points(prbf_ctx* ctx, double* x, double* y, double* z);

/* ---- kernel (b): truncated Gaussian matvec ---------------------------
 * MatShell MULT callback mymatmult (2d/matmult.cxx:6-51) and the GPU seam
 * gpumatmult (2d/gpumatmult.cu:65-67): w = A~ g in sorted order, FP64.
 * nrhs vectors stored back to back (lengths nj in, ni out). */
prbf_status prbf_matvec(prbf_ctx* ctx, int nrhs, const double* g_sorted, double* w_sorted);
prbf_status prbf_matvec_dev(prbf_ctx* ctx, int nrhs, const double* g_sorted, double* w_sorted);

/* ---- kernel (c): evaluation at targets -------------------------------
 * vorticity_evaluation (2d/vorticity_evaluation.cxx:26-299 with
 * Get_vorticity, 2d/get_vorticity.h:9-27): w_i = sum_trunc g_j G(x_i - x_j),
 * both in the caller's original order. */
prbf_status prbf_evaluate(prbf_ctx* ctx, int nrhs, const double* gj, double* wi);
prbf_status prbf_evaluate_dev(prbf_ctx* ctx, int nrhs, const double* gj, double* wi);

/* Allocation-free evaluation: prbf_evaluate_dev with its sorted-order staging
 * (nrhs (nj + ni) doubles) in a static device array of 2^21 doubles (16 MiB,
 * one per GPU, held until the process ends) instead of per-call device
 * allocations; stream-ordered, no hipMalloc, hipFree or synchronisation, so
 * it may run under HIP graph capture.  No counterpart in the reference: it is
 * for callers that cannot allocate (graph capture, tight serving loops).
 * Larger inputs return PRBF_ERR_LIMIT.  All calls on one GPU share the array:
 * calls that overlap on one GPU (several contexts or streams) race. */
prbf_status prbf_evaluate_dev_noalloc(prbf_ctx* ctx, int nrhs, const double* gj, double* wi);

/* ---- RASM preconditioner ----------------------------------------------
 * MatShell GET_SUBMATRICES callback mysubmat (2d/matmult.cxx:53-104) + the
 * PETSc dense sub-factorisation, for the subdomains set up by
 * PCASMSetLocalSubdomains (2d/rbf_interpolation.cxx:213-229). */
prbf_status prbf_rasm_setup(prbf_ctx* ctx);
/* Same for a chosen prbf_pc (RASM_FP64 or RASM_FP32).  No reference
 * counterpart for FP32: it is the mixed-precision variant of the same PC. */
prbf_status prbf_rasm_setup_pc(prbf_ctx* ctx, int pc);
/* Force a launch configuration of the block factorisation instead of the
 * built-in autotuner: PRBF_TUNE_CHOL_WAVES (4, 6 or 8 wavefronts per
 * workgroup), PRBF_TUNE_CHOL_PANEL (1 to 4 16-wide tile columns per panel),
 * PRBF_TUNE_CHOL_ORDER (prbf_chol_order; right-looking unless set).
 * Right-looking: FP32 has every pair; FP64 has 4 or 8 wavefronts with 1, 2
 * or 4 columns.  Left-looking: 4 or 8 wavefronts with 1, 2 or 4 columns in
 * both precisions.  A forced configuration without a kernel, or whose LDS does not
 * fit, runs right-looking as 4 wavefronts with 1 column.  No reference
 * counterpart. */
prbf_status prbf_set_tuning(prbf_ctx* ctx, int key, int value);
/* PCApply_ASM (restrict): z = sum_c P_core A_c^{-1} R_buffer r, sorted order. */
prbf_status prbf_rasm_apply(prbf_ctx* ctx, const double* r_sorted, double* z_sorted);
prbf_status prbf_rasm_apply_dev(prbf_ctx* ctx, const double* r_sorted, double* z_sorted);
/* The dense untruncated RASM block of cluster c, A[i*n + j] = G(x_i - x_j)
 * over its buffer particles (FP64, row-major, n x n), and the particles'
 * sorted indices idx[0..n) in the order of the PCASM subdomain `is` (core
 * first, then the neighbour cells'): what the MatShell GET_SUBMATRICES
 * callback mysubmat builds per cluster (2d/matmult.cxx:53-104, subdomains
 * from 2d/rbf_interpolation.cxx:214-229).  A or idx may be NULL; n receives
 * the block size.  Needs prbf_set_points with one particle set. */
prbf_status prbf_rasm_get_block(prbf_ctx* ctx, int64_t c, double* A, int32_t* idx, int32_t* n);
/* Core rows of A_c^{-1} for cluster c: M[t*n + b], t < T core members,
 * b < n buffer slots; n and T returned.  Pass M == NULL to query the sizes. */
prbf_status prbf_rasm_get_core_rows(prbf_ctx* ctx, int64_t c, double* M, int32_t* n, int32_t* T);

/* ---- solve -------------------------------------------------------------
 * rbf_interpolation (2d/rbf_interpolation.cxx:25-310): finds gamma with
 * A~ gamma = w, KSPSolve with zero initial guess (:183-185,232).  w and
 * gamma in original order.  hist (optional) receives the residual history
 * of the convergence test (up to hist_cap entries). */
prbf_status prbf_solve(prbf_ctx* ctx, const prbf_solve_opts* opts, const double* w, double* gamma,
                       prbf_solve_info* info, double* hist, int hist_cap);
prbf_status prbf_solve_dev(prbf_ctx* ctx, const prbf_solve_opts* opts, const double* w, double* gamma,
                           prbf_solve_info* info, double* hist, int hist_cap);

/* The same solve for nrhs = 1..3 right-hand sides at once (w and gamma hold
 * them back to back, ni values each, original order; info receives nrhs
 * entries; hist, optional, nrhs rows of hist_cap).  The reference solves the
 * three components of the 3D vorticity field (3d/main.cxx:104) with one
 * RBFInterpolation call each (3d/wrapper.cxx:93-101); here each right-hand
 * side keeps its own Krylov basis, Givens rotations and convergence test (the
 * results equal nrhs single solves run kernel by kernel, bit for bit), while
 * every iteration runs one matvec over all of them (one exponential per
 * pair) and one preconditioner apply that streams each block's core rows
 * once.  Kernel by kernel whatever opts->graph says.  With PRBF_PC_LEARNED
 * (FGMRES) the network runs on each right-hand side and a network cluster
 * moves to its exact block at a restart when its core residual, summed over
 * the right-hand sides still running, fell by less than 1000x over the
 * cycle. */
prbf_status prbf_solve_multi(prbf_ctx* ctx, const prbf_solve_opts* opts, int nrhs, const double* w, double* gamma,
                             prbf_solve_info* info, double* hist, int hist_cap);
prbf_status prbf_solve_multi_dev(prbf_ctx* ctx, const prbf_solve_opts* opts, int nrhs, const double* w,
                                 double* gamma, prbf_solve_info* info, double* hist, int hist_cap);

/* ---- learned preconditioner (no counterpart in the reference) -----------
 * A small MLP (nin -> 256 -> 256 -> 256 -> nout, GELU, BF16 weights and
 * activations on the BF16 matrix cores, FP32 accumulation) maps the residual
 * on the lattice-slot patch around a cluster (scaled by its RMS) and the
 * particles' offsets from their slots to a correction on the cluster's core
 * particles.  2D: 8x8 patch around a cluster of 4x4 lattice spacings,
 * nin = 192 (64 residuals, 128 offsets), nout = 16.  3D: 5x5x5 patch around
 * a cluster of 3x3x3 spacings, nin = 512 (125 residuals, 375 offsets, 12
 * zeros), nout = 32 (27 core corrections; the other 5 outputs are trained
 * toward zero and unused).  The network is trained on the device (Adam, FP32
 * master weights) on pairs (patch, exact RASM core solve) harvested from
 * FGMRES runs with the exact preconditioner.  Device arrays; these calls
 * synchronise. */
typedef struct prbf_mlp prbf_mlp;
/* parameter count of the 2D network, and of the network of dimension dim
 * (2 or 3; -1 otherwise) */
int64_t prbf_mlp_num_params(void);
int64_t prbf_mlp_num_params_dim(int dim);
/* a 2D network; prbf_mlp_create_dim: of dimension 2 or 3 (Glorot-uniform
 * weights from a splitmix64 stream of seed, zero biases) */
prbf_status prbf_mlp_create(prbf_mlp** net, uint64_t seed);
prbf_status prbf_mlp_create_dim(prbf_mlp** net, uint64_t seed, int dim);
prbf_status prbf_mlp_destroy(prbf_mlp* net);
const char* prbf_mlp_last_error(const prbf_mlp* net);
/* params: host, prbf_mlp_num_params_dim(dim) floats (W0, b0, ..., W3, b3; W_l row-major in x out) */
prbf_status prbf_mlp_set_params(prbf_mlp* net, const float* params, double out_scale);
prbf_status prbf_mlp_get_params(prbf_mlp* net, float* params, double* out_scale);
/* Y (m x nout) = net(X (m x nin)), row-major device arrays */
prbf_status prbf_mlp_forward_dev(prbf_mlp* net, int64_t m, const float* X, float* Y);
/* mean-square loss against T (m x nout) and its gradient (device, num_params floats); m % 16 == 0 */
prbf_status prbf_mlp_loss_grad_dev(prbf_mlp* net, int64_t m, const float* X, const float* T, float* loss,
                                   float* grad);
/* `steps` Adam steps on minibatches of `batch` pairs drawn (seeded) from the
 * npairs rows of X, T; loss_hist (host, steps floats, nullable) */
prbf_status prbf_mlp_train_dev(prbf_mlp* net, int64_t npairs, const float* X, const float* T, int batch, int steps,
                               float lr, uint64_t seed, float* loss_hist);
/* number of clusters with a complete patch (the network's candidates) */
prbf_status prbf_learned_eligible(prbf_ctx* ctx, int64_t* n);
/* For the sorted vector r: features X (n x nin) of the eligible clusters, the
 * patch RMS s (n), and with z != NULL the targets T = z_core / (s out_scale)
 * (n x nout, zeros past the core); at most cap rows.  nin, nout: those of
 * the context's dimension. */
prbf_status prbf_learned_pairs_dev(prbf_ctx* ctx, const double* r, const double* z, double out_scale, float* X,
                                   float* T, double* s, int64_t cap, int64_t* n);
/* rho = |r - A z| / |r| on each eligible cluster's patch for the core values z */
prbf_status prbf_learned_rho_dev(prbf_ctx* ctx, const double* r, const double* z, double* rho, int64_t cap,
                                 int64_t* n);
/* Attach a trained network of the context's dimension: an eligible cluster
 * uses it when its patch residual reduction rho stays <= tau on two seeded
 * probe residuals; every other cluster falls back to its exact block,
 * factorised in FP32 here.  n_network receives the number of clusters on
 * the network. */
prbf_status prbf_learned_attach(prbf_ctx* ctx, prbf_mlp* net, double tau, int64_t* n_network);
/* Copies basis vector j of the last FGMRES cycle and its preconditioned
 * image into the device arrays v and z (sorted order; for harvesting pairs). */
prbf_status prbf_krylov_vectors_dev(prbf_ctx* ctx, int j, double* v, double* z);

/* ---- library entry of the 3D tree ---------------------------------------
 * RBFInterpolation (3d/wrapper.cxx:20-105), for dim 2 or 3: FP32 grid
 * (xg, yg, zg) with values wg in; if the grid is not the particle set
 * (|xg[0] - xp[0]| > 1e-6, 3d/wrapper.cxx:93) the field of the particles
 * (xp, yp, zp, strengths gp) is evaluated at the grid first; the grid
 * strengths solving A~ g = w are written over wg.  No context: errors are
 * reported by prbf_last_error(NULL). */
prbf_status prbf_rbf_interpolation_f32(const prbf_params* p, int64_t ng, const float* xg, const float* yg,
                                       const float* zg, float* wg, int64_t np, const float* xp, const float* yp,
                                       const float* zp, const float* gp, const prbf_solve_opts* opts, int* its);

#ifdef __cplusplus
}
#endif

#endif /* PETRBF_H */
