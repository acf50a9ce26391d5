/*
 * fmca_amd.h -- C ABI of the MI355X-native H2-matrix kernel action.
 *
 * Drop-in for FMCA's matrix-free H2 action (muchip/fmca).  Plain pointers and
 * sizes only; every matrix is column-major FP64 as an Eigen::MatrixXd is.
 * Points are dim x n (one point per column), right-hand sides n x nrhs.
 *
 * Host pointers are borrowed for the duration of a call only (FMCA's actor
 * keeps references to P/Peval, FMCA/src/Moments/NystromMoments.h:41; this
 * library copies what it needs).  Device memory is owned by the actor handle.
 * All functions return a status code instead of FMCA's assert()s.
 */
#ifndef FMCA_AMD_H_
#define FMCA_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fmca_h2_actor fmca_h2_actor; /* opaque */

typedef enum {
  FMCA_OK = 0,
  FMCA_ERR_INVALID_ARGUMENT = 1,
  FMCA_ERR_UNKNOWN_KERNEL = 2, /* FMCA: assert "desired kernel not implemented" */
  FMCA_ERR_NOT_SUPPORTED = 3,  /* dim > 3, degree > 9, ... */
  FMCA_ERR_NO_DEVICE = 4,
  FMCA_ERR_OUT_OF_MEMORY = 5,
  FMCA_ERR_DEVICE = 6,
  FMCA_ERR_NOT_CONVERGED = 7,
  FMCA_ERR_NONFINITE = 8
} fmca_status;

/* Arithmetic of the coupling and near-field kernels. */
enum {
  FMCA_PRECISION_FP64 = 0, /* plain form: every phase in FP64 */
  FMCA_PRECISION_FP32 = 1, /* precision form: FP32 kernel entries and contractions */
  /* precision form with the contractions of the kernel entries on BF16 matrix
   * cores: FP32 entries and right operands, each split into two BF16 terms
   * (hi = bf16(v), lo = bf16(v - hi)), products hi.hi + lo.hi + hi.lo with
   * FP32 accumulation -- the error class of FP32 entries at 3/16 of the FP32
   * matrix-core cycles.  Used where the FP32 form would use the matrix cores
   * for kernel values (>= 8 rhs); every other phase is the FP32 form's. */
  FMCA_PRECISION_FP32_BF16 = 2
};

enum { FMCA_ROWS = 0, FMCA_COLS = 1 }; /* row tree = Peval (targets), column tree = P */

/* Kernel description: FMCA::CovarianceKernel(ktype, l, c, nu)
 * (FMCA/src/CovarianceKernel/CovarianceKernel.h:50-52).  ktype is one of
 * EXPONENTIAL, MATERN32, MATERN52, MATERN72, MATERN92, GAUSSIAN,
 * INVMULTIQUADRIC, MULTIQUADRIC, BIHARMONIC2D, BIHARMONIC3D, TRIHARMONIC3D,
 * TPS2D (case-insensitive). */
typedef struct {
  const char *ktype;
  double l, c, nu;
} fmca_kernel_desc;

/* Replaces H2kernelActor(kernel, Peval, P, mpole_deg, eta)
 * (FMCA/src/H2Matrix/H2kernelActor.h:21-36): tensor-product Chebyshev
 * moments of degree `deg`, ClusterTree with CardinalityBisection, pattern
 * with CompareCluster at `eta`.  Peval: dim x m (targets), P: dim x n
 * (sources).  The plan is built on the host; device buffers are created on
 * the first action (or fmca_h2_upload), on the current HIP device. */
fmca_status fmca_h2_create(const fmca_kernel_desc *kernel, int dim, int64_t m,
                           const double *Peval, int64_t n, const double *P, int deg,
                           double eta, fmca_h2_actor **actor);

/* Interpolators of NystromMoments<Interpolator>. */
enum {
  FMCA_INTERP_TENSOR_PRODUCT = 0, /* TensorProductInterpolator: q = (deg+1)^dim */
  FMCA_INTERP_TOTAL_DEGREE = 1    /* TotalDegreeInterpolator: q = binom(deg+dim, dim) */
};

/* Replaces H2kernelActor<CovarianceKernel, NystromMoments<Interpolator>, ...>
 * (FMCA/src/H2Matrix/H2kernelActor.h:17-36) with the interpolator chosen at
 * run time.  FMCA_INTERP_TOTAL_DEGREE is TotalDegreeInterpolator
 * (FMCA/src/Interpolator/TotalDegreeInterpolator.h:27-48), the interpolator
 * of FMCA's own actor test (tests/test_H2kernelActor.cpp:20-24) and Python
 * binding (py/FMCA.cpp:37-38): Legendre moments, DeMarchi nodes, dense invV;
 * the tree's minimum cluster size is its q.  fmca_h2_create is this call
 * with FMCA_INTERP_TENSOR_PRODUCT. */
fmca_status fmca_h2_create_interp(const fmca_kernel_desc *kernel, int dim, int64_t m,
                                  const double *Peval, int64_t n, const double *P, int deg,
                                  double eta, int interpolator, fmca_h2_actor **actor);

/* Where fmca_h2_create_ex plans. */
enum { FMCA_PLAN_HOST = 0, FMCA_PLAN_DEVICE = 1 };

/* fmca_h2_create_interp with a choice of planner.  FMCA_PLAN_DEVICE runs on
 * the current HIP device what H2kernelActor's constructor runs on the CPU
 * (H2kernelActor.h:21-30):
 *  - the O(n) work of ClusterTreeInitializer<ClusterTree>::init for both trees
 *    (ClusterTreeInitializer_ClusterTree.h:24-132, CardinalityBisection in
 *    ClusterTreeSplitter.h:60-81): the points are copied there for the build
 *    and freed after it;
 *  - the level-by-level expansion of H2MatrixBase::computePattern
 *    (H2MatrixBase.h:160-191) with CompareCluster (Admissibility.h:95-115).
 * The device memory they use is transient and counted against
 * fmca_set_device_memory_limit.  The plan equals the host planner's bit for
 * bit.  Where coordinates are equal at a split point or inside a leaf, or a
 * coordinate is -0.0, that tree is built on the host; where a cluster pair's
 * admissibility lies within rounding of eta, the expansion runs on the host;
 * where the device memory (or the cap) has no room, that part also runs on
 * the host (fmca_h2_plan_report says which).  FMCA_ERR_DEVICE /
 * FMCA_ERR_NO_DEVICE if a device step fails otherwise. */
fmca_status fmca_h2_create_ex(const fmca_kernel_desc *kernel, int dim, int64_t m,
                              const double *Peval, int64_t n, const double *P, int deg,
                              double eta, int interpolator, int planner,
                              fmca_h2_actor **actor);

/* Seconds the create call spent building the trees and the block pattern,
 * and where (where[0] row tree, where[1] column tree, where[2] block
 * expansion): 0 host, 1 device, 2 device asked for, host used.  Null outputs
 * are skipped.  No reference counterpart (FMCA reports no setup times). */
fmca_status fmca_h2_plan_report(const fmca_h2_actor *actor, double *tree_seconds,
                                double *pattern_seconds, int *where);

/* H2Matrix::computeH2Matrix (H2MatrixBase.h:222-255): generate every coupling
 * matrix S_ts = K(xi_t, xi_s), every dense block K(P_t, P_s) and every leaf
 * basis V_t once and keep them in device memory, in the given precision, as
 * matrix-core operand tiles.  Afterwards every forward action of that
 * precision with the kernel value -- fmca_h2_action* in natural or cluster
 * order, the matvecs of fmca_h2_cg_device and the FMCA_KERNEL_VALUE forward
 * calls of fmca_h2_action_variant(s)_device (so the differentiable forward
 * too) -- streams them instead of generating entries, as operator*
 * (H2MatrixBase.h:107-109, matrix_vector_product_impl.h:17-98) applies the
 * stored blocks; transposed actions and the derivative variants still
 * generate entries.  One precision at a time; fmca_h2_set_kernel with other
 * kernel parameters, and fmca_h2_release_assembled, free them.  *bytes (if
 * not null): the device bytes held (about 4 B per entry in FP32, 8 in FP64,
 * entries padded to 16-row / 4-column tiles).  FMCA_ERR_NOT_SUPPORTED for
 * q > 128; FMCA_ERR_OUT_OF_MEMORY if they do not fit. */
fmca_status fmca_h2_assemble(fmca_h2_actor *actor, int precision, int64_t *bytes);

/* Frees the assembled blocks; later actions generate entries again. */
fmca_status fmca_h2_release_assembled(fmca_h2_actor *actor);

/* Replaces H2MatrixProjector<Derived>(h2mat).project(mat)
 * (FMCA/src/H2Matrix/H2MatrixProjector.h:21-75): projects the dense matrix M
 * (m x n, column-major host memory with leading dimension ld, natural order:
 * row i belongs to target point i of Peval, column j to source point j of P)
 * onto the actor's H2 format.  Near blocks are M's blocks; every far block
 * (t, s) gets the coupling matrix
 *   S_ts = (V_t V_t^T)^-1 (V_t M_ts V_s^T) (V_s V_s^T)^-1,
 * the Frobenius-orthogonal projection onto the cluster bases (FMCA computes
 * the inverse Gram matrices through its dual cluster bases,
 * compute_dual_cluster_bases_impl.h:25-57; here by Cholesky factors, FP64,
 * on the host).  The projected blocks are kept in device memory in the given
 * precision as fmca_h2_assemble keeps its blocks, and the forward actions of
 * that precision stream them until fmca_h2_release_assembled,
 * fmca_h2_assemble or fmca_h2_set_kernel with other parameters.
 * dual_check (if not null): FMCA's check of the dual bases
 * (compute_dual_cluster_bases_impl.h:66-86) on the row and the column tree,
 * max over clusters of ||V^T - V^T (V V^T)^-1 V V^T||_F / ||V^T||_F (FMCA
 * asserts < 5e-10).  FMCA_ERR_NOT_SUPPORTED for the total-degree
 * interpolator, q > 128 or a singular basis Gram matrix. */
fmca_status fmca_h2_project(fmca_h2_actor *actor, const double *M, int64_t ld, int precision,
                            double dual_check[2]);

/* Interpolation data, as the interpolator's idcs(), Xi() and invV() return
 * them (TotalDegreeInterpolator.h:50-55, TensorProductInterpolator.h:83-85):
 * the interpolator id, q, the q multi-indices in FMCA's order (q x dim,
 * row-major), the nodes Xi (dim x q, column-major) and invV (q x q,
 * column-major; the identity for the tensor product).  Null outputs are
 * skipped. */
fmca_status fmca_h2_interpolation(const fmca_h2_actor *actor, int *interpolator, int *q,
                                  int32_t *multi_indices, double *nodes, double *invV);

/* Destroys the actor and frees its device memory. */
fmca_status fmca_h2_destroy(fmca_h2_actor *actor);

/* Uploads the plan to the current device now instead of on the first action. */
fmca_status fmca_h2_upload(fmca_h2_actor *actor);

/* Replaces Matrix H2kernelActor::action(const Matrix &rhs) const
 * (H2kernelActor.h:38-80): out = K(Peval, P) * rhs in natural order.
 * rhs: n x nrhs (leading dimension ld_rhs >= n), out: m x nrhs (ld_out >= m),
 * both host memory.  Synchronous. */
fmca_status fmca_h2_action(fmca_h2_actor *actor, const double *rhs, int64_t ld_rhs,
                           int64_t nrhs, double *out, int64_t ld_out, int precision);

/* Replaces H2MatrixBase::action(mat_eval, rhs) (FMCA/src/H2Matrix/H2MatrixBase.h:113-152)
 * as called by the Python binding (py/FMCA.cpp:230-240): rhs in column-tree
 * order, result in row-tree order.  Host memory, synchronous. */
fmca_status fmca_h2_action_cluster_order(fmca_h2_actor *actor, const double *rhs,
                                         int64_t ld_rhs, int64_t nrhs, double *out,
                                         int64_t ld_out, int precision);

/* Device-memory form of the two calls above: d_rhs / d_out are device
 * pointers, the work is enqueued on `stream` (a hipStream_t, NULL = default
 * stream) and the call returns without synchronising.  cluster_order != 0
 * selects the H2MatrixBase::action ordering.  Every action of an actor uses
 * the actor's workspaces, so an action enqueued on another stream than the
 * actor's previous action first waits for that action (an event on the
 * previous stream); this holds for every device entry point below.  Work
 * captured into a HIP graph records no such event: while a stream captures
 * actions of an actor, no other stream may drive it. */
fmca_status fmca_h2_action_device(fmca_h2_actor *actor, const double *d_rhs, int64_t ld_rhs,
                                  int64_t nrhs, double *d_out, int64_t ld_out, int precision,
                                  int cluster_order, void *stream);

/* fmca_h2_action_device that allocates nothing.  No counterpart in FMCA: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * permuted rhs, moments and far-field sums live in a fixed 64 MiB device
 * array of the library (one per GPU, held until the process ends) instead of
 * the actor's workspaces: n * nrhs * b + (row + column clusters) * q * nrhs * b
 * bytes (b = 8 FP64, 4 FP32; each part rounded up to 256 B) must fit, else
 * FMCA_ERR_NOT_SUPPORTED.  The plan must already be on the device
 * (fmca_h2_upload or an earlier action), else FMCA_ERR_INVALID_ARGUMENT.  No
 * profiling events are recorded.  Every call on one GPU shares that GPU's
 * array, whatever the actor: calls that overlap on one GPU (other streams,
 * other threads, other actors) race, and the caller must order them. */
fmca_status fmca_h2_action_device_noalloc(fmca_h2_actor *actor, const double *d_rhs,
                                          int64_t ld_rhs, int64_t nrhs, double *d_out,
                                          int64_t ld_out, int precision, int cluster_order,
                                          void *stream);

/* Transposed action out = K(Peval, P)^T * rhs: rhs m x nrhs, out n x nrhs,
 * device memory (the adjoint needed by the differentiation form; FMCA has no
 * counterpart). */
fmca_status fmca_h2_action_transpose_device(fmca_h2_actor *actor, const double *d_rhs,
                                            int64_t ld_rhs, int64_t nrhs, double *d_out,
                                            int64_t ld_out, int precision, void *stream);

/* Options of the iterative solve of (K + ridge I) x = b, anchored on FMCA's
 * pCG (FMCA/src/util/PCG.h:15-44: identity preconditioner here, stop at
 * ||b - (K + ridge I) x|| / ||b|| <= tol, tol = 1e-8 in FMCA).  Vectors, inner
 * products and x updates are FP64 in every form; precision selects the matvec
 * that drives the iteration:
 *   FMCA_PRECISION_FP64       FP64 matvecs throughout;
 *   FMCA_PRECISION_FP32(_BF16) lower-precision matvecs with reliable updates:
 *                        once every active column's iterated residual has
 *                        dropped by refine_factor since its last replacement
 *                        (or reached tol), the true residual is recomputed
 *                        with the FP64 action and replaces it.
 * method:
 *   FMCA_SOLVER_CG   FMCA's CG; on replacement the search direction is
 *                    re-orthogonalised against the new residual;
 *   FMCA_SOLVER_GCR  GCR (Eisenstat, Elman & Schultz 1983; no FMCA
 *                    counterpart): the residual's image is orthogonalised
 *                    against a window of up to max_basis stored pairs
 *                    (p_j, (K + ridge I) p_j), and the residual minimised over
 *                    them.  Each stored pair is consistent as its matvec
 *                    computed it, so lower-precision matvecs do not delay
 *                    convergence as they do CG's; a replaced residual is
 *                    projected on the window.  Holds 2 max_basis vectors
 *                    of n x nrhs (FP64 for FP64 matvecs, FP32 otherwise)
 *                    during the solve.
 * Convergence is always decided on the FP64 residual. */
enum { FMCA_SOLVER_CG = 0, FMCA_SOLVER_GCR = 1 };
typedef struct {
  double ridge;
  double tol;
  int precision;
  double refine_factor; /* 0 < f < 1, default 0.1 */
  int max_iter;         /* cap on iterations (matvecs of `precision`) */
  int max_outer;        /* cap on FP64 residual replacements */
  int method;           /* FMCA_SOLVER_CG (0) or FMCA_SOLVER_GCR */
  int max_basis;        /* GCR window, at most 256 pairs; 0: as many as fit in
                           half the free device memory */
} fmca_cg_options;

typedef struct {
  int converged;        /* every column reached tol */
  int iterations;       /* CG iterations */
  int outer;            /* FP64 residual replacements */
  int matvecs_fp64;     /* FP64 actions */
  int matvecs_fp32;     /* FP32 actions */
  double rel_residual;  /* max over columns of ||b - (K + ridge I) x|| / ||b||, FP64 action,
                           at the last replacement, or of the returned x when a cap
                           stopped the solve */
  double seconds;       /* wall time of the solve */
  int indefinite;       /* 1 if a column met p.(K + ridge I)p <= 0 (the H2 matrix plus ridge
                           is not positive definite); the solve then stops */
} fmca_cg_result;

/* Solve of (K(P, P) + ridge I) x = b for nrhs independent columns (requires
 * m == n) by opts->method.  d_b: n x nrhs, d_x: n x nrhs initial guess on
 * entry (FMCA's pCG starts from x = b), solution on exit; device memory,
 * natural order.  Synchronises `stream`. */
fmca_status fmca_h2_cg_device(fmca_h2_actor *actor, const double *d_b, int64_t ld_b,
                              int64_t nrhs, double *d_x, int64_t ld_x,
                              const fmca_cg_options *opts, fmca_cg_result *result, void *stream);

/* Replaces the kernel of an actor (type and parameters) without rebuilding
 * the plan: tree, bases and block pattern do not depend on the kernel
 * (FMCA rebuilds the evaluator per action, py/FMCA.cpp:233-235). */
fmca_status fmca_h2_set_kernel(fmca_h2_actor *actor, const fmca_kernel_desc *kernel);

/* Replaces CovarianceKernel::setDistanceType(dist_type)
 * (FMCA/src/CovarianceKernel/CovarianceKernel.h:123-137) for the actor's
 * kernel: "EUCLIDEAN" (the default), |x - y|, or "GEODESIC",
 * acos(clamp(x . y, -1, 1)) (SphereClusterTree::geodesicDistance,
 * FMCA/src/Clustering/SphereClusterTree.h:65-69), the great-circle distance
 * of points on the unit sphere.  As in FMCA only the kernel entries change:
 * the trees and the admissibility stay Euclidean.  Geodesic entries are
 * formed in FP64 (x[0] y[0] followed by one fma per axis, then acos): actions
 * in FMCA_PRECISION_FP64 in both orders, transposed, FMCA_KERNEL_VALUE and
 * FMCA_KERNEL_DL variants, FP64 solves and FP64 assembled or projected
 * blocks; the FP32 forms, FMCA_KERNEL_PHI and the gradient form return
 * FMCA_ERR_NOT_SUPPORTED.  An unknown name is FMCA_ERR_INVALID_ARGUMENT
 * (FMCA: assert "desired distance not implemented"); names are
 * case-insensitive.  Assembled blocks are released on a change. */
fmca_status fmca_h2_set_distance_type(fmca_h2_actor *actor, const char *dist_type);

/* The actor's distance type: *geodesic = 1 for "GEODESIC", 0 for "EUCLIDEAN". */
fmca_status fmca_h2_distance_type(const fmca_h2_actor *actor, int *geodesic);

/* Kernel variants of an action (no FMCA counterpart; the differentiation
 * form).  The H2 operator is linear in the kernel function, so derivatives
 * of the compressed matrix are H2 actions with derivative kernels on the same
 * tree, bases and blocks:
 *   FMCA_KERNEL_VALUE  k(r)
 *   FMCA_KERNEL_DL     dk/dl (derivative with respect to the lengthscale)
 *   FMCA_KERNEL_PHI    k'(r) / r, the radial factor of the gradient with
 *                      respect to a point: d/dx_i sum_j k(|x_i - x_j|) w_j =
 *                      sum_j phi(r_ij) (x_i - x_j) w_j  (phi(0) = 0 where
 *                      k'(r)/r is singular) */
enum { FMCA_KERNEL_VALUE = 0, FMCA_KERNEL_DL = 1, FMCA_KERNEL_PHI = 2 };

/* out = K_variant(Peval, P) rhs (transpose = 0; rhs n x nrhs, out m x nrhs) or
 * K_variant(Peval, P)^T rhs (transpose = 1; rhs m x nrhs, out n x nrhs), device
 * memory, natural order, enqueued on `stream`. */
fmca_status fmca_h2_action_variant_device(fmca_h2_actor *actor, int variant, int transpose,
                                          const double *d_rhs, int64_t ld_rhs, int64_t nrhs,
                                          double *d_out, int64_t ld_out, int precision,
                                          void *stream);

/* Several kernel variants in one action: rhs columns [16 g, 16 g + 16) are
 * multiplied by the H2 matrix of variant variants[g] (FMCA_KERNEL_VALUE /
 * _DL / _PHI), g < ngroups <= 8, and 16 (ngroups - 1) < nrhs <= 16 ngroups.
 * On the matrix cores every kernel entry is evaluated once for all groups
 * (the fused backward of the differentiable matvec); where a phase runs on
 * VALU the groups are applied as separate actions.  No FMCA counterpart. */
fmca_status fmca_h2_action_variants_device(fmca_h2_actor *actor, const int *variants,
                                           int ngroups, int transpose, const double *d_rhs,
                                           int64_t ld_rhs, int64_t nrhs, double *d_out,
                                           int64_t ld_out, int precision, void *stream);

/* Gradient of L = <G, K~(P, P; l) alpha> for one point set (the backward of
 * the differentiable matvec y = K~ alpha with dL/dy = G; no FMCA
 * counterpart).  K~ is the H2 matrix of the actor (targets = sources, the
 * tensor-product interpolator), with the tree and boxes held fixed:
 *   d_dalpha  (n x nrhs, ld_dalpha) = K~^T G = K~ G
 *   d_dpoints (dim x n, row stride ld_points) = dL/dP: near blocks through
 *             phi(r) = k'(r)/r and the pair weights G_p.alpha_j + alpha_p.G_j,
 *             far blocks through the gradient of the interpolation basis
 *             dotted with the far-field sums of alpha and G (the derivative of
 *             the interpolated far field)
 *   d_dl      (one device double) = <G, dK~/dl alpha>
 * alpha, G: n x nrhs device FP64 (ld_alpha, ld_g), natural order, nrhs <= 16.
 * Enqueued on `stream` (one action on [alpha | G | alpha] with the last
 * group under dk/dl, then the gradient kernels); outputs are overwritten.
 * FMCA_ERR_NOT_SUPPORTED for two point sets, the total-degree interpolator,
 * more than 16 rhs, or a coupling that does not fit the matrix cores -- the
 * derivative-kernel actions above cover those. */
fmca_status fmca_h2_gradient_device(fmca_h2_actor *actor, const double *d_alpha,
                                   int64_t ld_alpha, const double *d_g, int64_t ld_g,
                                   int64_t nrhs, int precision, double *d_dalpha,
                                   int64_t ld_dalpha, double *d_dpoints, int64_t ld_points,
                                   double *d_dl, void *stream);

/* Permutation of a tree (cluster order position -> point index), as
 * pyH2Matrix::rindices / cindices (py/FMCA.cpp:224-229) and
 * ClusterTreeBase::indices (FMCA/src/Clustering/ClusterTreeBase.h:63-65). */
fmca_status fmca_h2_indices(const fmca_h2_actor *actor, int which, uint32_t *out);

/* H2MatrixBase::statistics (H2MatrixBase.h:70-105): rows, cols, low-rank
 * blocks, dense blocks, nonzeros per row, storage in GB. */
fmca_status fmca_h2_statistics(const fmca_h2_actor *actor, double out[6]);

/* Sizes: [0] dim, [1] m, [2] n, [3] q, [4] row clusters, [5] column clusters,
 * [6] far blocks, [7] near blocks, [8] near kernel entries, [9] far kernel
 * entries, [10] depth of the row tree, [11] depth of the column tree,
 * [12] 1 if both trees are the same object. */
fmca_status fmca_h2_info(const fmca_h2_actor *actor, double out[13]);

/* Clusters of a tree in FMCA level order (block ids): begin, size, level,
 * sons (2 per cluster, -1 for leaves), bounding boxes (3 x dim per cluster:
 * min, max, size).  Arrays sized by fmca_h2_info. */
fmca_status fmca_h2_clusters(const fmca_h2_actor *actor, int which, int64_t *begin,
                             int64_t *size, int32_t *level, int32_t *sons, double *bb);

/* Work of each workgroup of the forward action's coupling (which = 0) or
 * leaf output (which = 1), 4 ints per workgroup: target (cluster id / index
 * into the row tree's leaves), begin and end in its far / near list, flags
 * (1: the target is split over several workgroups that add into a zeroed
 * output, 2: this workgroup applies the leaf's far field); all -1 for an
 * idle workgroup.  Targets above three times the mean work are split; the
 * workgroups of a launch go to the 8 XCDs round-robin and each XCD gets a
 * contiguous piece of every tree level of about equal work.  No FMCA
 * counterpart (FMCA schedules with OpenMP, H2kernelActor.h:30-34).
 * out == NULL: *len receives the number of workgroups. */
fmca_status fmca_h2_launch_order(const fmca_h2_actor *actor, int which, int32_t *out,
                                 int64_t *len);

/* FMCA's leaf blocks: row / column block ids and low-rank flag; count = far + near. */
fmca_status fmca_h2_blocks(const fmca_h2_actor *actor, int32_t *row, int32_t *col,
                           int8_t *low_rank);

/* Per-phase timing (no FMCA counterpart; FMCA times whole calls with Tictoc,
 * FMCA/src/util/Tictoc.h:22-42).  When enabled, every action records HIP
 * events on its launch stream around its six phases:
 *   0 permute, 1 upward leaf moments, 2 upward transfers, 3 coupling,
 *   4 downward transfers, 5 leaf output (near field + leaf basis).
 * fmca_h2_phase_times waits for the recorded events and returns the summed
 * milliseconds and the number of actions per phase; reset != 0 clears them. */
fmca_status fmca_h2_set_profiling(fmca_h2_actor *actor, int enable);
fmca_status fmca_h2_phase_times(fmca_h2_actor *actor, double ms[6], int64_t count[6], int reset);

/* fmca_h2_phase_times restricted to the actions of one precision
 * (FMCA_PRECISION_FP64 / _FP32; -1: both, as fmca_h2_phase_times); reset
 * clears that precision's sums.  The events of read actions are kept for
 * later profiled actions, so that profiling records events but creates none
 * once warm (no FMCA counterpart). */
fmca_status fmca_h2_phase_times_precision(fmca_h2_actor *actor, int precision, double ms[6],
                                          int64_t count[6], int reset);

/* Diagnostic (no FMCA counterpart): evaluates the device exp (fn 0 FP64,
 * 2 FP32) or sqrt (fn 1 FP64, 3 FP32) that the kernel-entry loops use, on n
 * device values, so that their accuracy can be tested. */
fmca_status fmca_diag_math(int fn, const double *d_x, double *d_y, int64_t n, void *stream);

/* Device memory held by the library (every actor of the process): bytes
 * allocated now and the high-water mark since load or the last reset_peak.
 * No FMCA counterpart (FMCA allocates on the host heap); it states the device
 * footprint of a plan and its workspaces. */
fmca_status fmca_device_memory(int64_t *live_bytes, int64_t *peak_bytes, int reset_peak);

/* Caps the library's device memory (0 removes the cap): an allocation that
 * would take the live bytes above the cap fails with FMCA_ERR_OUT_OF_MEMORY,
 * as a full HBM would, and leaves the actor usable.  No FMCA counterpart; it
 * lets callers and tests exercise the out-of-memory path, and bounds the
 * library's share of HBM. */
fmca_status fmca_set_device_memory_limit(int64_t limit_bytes);

/* Static string for a status code. */
const char *fmca_status_string(fmca_status status);

/* Last error message of the calling thread ("" if none). */
const char *fmca_last_error(void);

/* ------------------------------------------------------------------------
 * Samplets: the second caller of the evaluator interface (interpolate_kernel
 * and compute_dense_block) besides the H2 action.
 * ------------------------------------------------------------------------ */

typedef struct fmca_samplet_tree fmca_samplet_tree;     /* opaque */
typedef struct fmca_samplet_matrix fmca_samplet_matrix; /* opaque */

/* Replaces pySampletTree(P, dtilde) (py/FMCA.cpp:51-70): H2SampletTree
 * <ClusterTree> (FMCA/src/Samplets/H2SampletTree.h:75-97) on
 * NystromMoments<TotalDegreeInterpolator>(P, 2 (dtilde - 1)) and
 * MinNystromSampletMoments<MonomialInterpolator>(P, dtilde - 1): the
 * ClusterTree with minimum leaf size q = binom(dim + 2 (dtilde - 1), dim),
 * samplets with dtilde vanishing moments from a Householder QR per cluster
 * (H2SampletTree.h:101-133), sampletMapper's coefficient order
 * (SampletMapper.h:19-31) and the multiscale cluster bases
 * (H2SampletTreeBase.h:68-86).  dtilde < 1 is taken as 1 (as FMCA's binding
 * does); dtilde <= 5.  Host only.  P: dim x n column-major. */
fmca_status fmca_samplet_tree_create(int dim, int64_t n, const double *P, int dtilde,
                                     fmca_samplet_tree **tree);
fmca_status fmca_samplet_tree_destroy(fmca_samplet_tree *tree);

/* info[7] = {n, dim, dtilde, clusters, q, mdtilde, depth}.  No direct FMCA
 * counterpart (FMCA exposes these as tree and moment members). */
fmca_status fmca_samplet_tree_info(const fmca_samplet_tree *tree, int64_t *info);

/* pySampletTree::indices (py/FMCA.cpp:72-74): the cluster-order permutation,
 * perm[i] = the point at cluster position i. */
fmca_status fmca_samplet_tree_indices(const fmca_samplet_tree *tree, uint32_t *perm);

/* pySampletTree::levels (py/FMCA.cpp:75-78, sampletLevelMapper,
 * SampletMapper.h:34-50): the level of each of the n coefficients. */
fmca_status fmca_samplet_tree_levels(const fmca_samplet_tree *tree, int32_t *levels);

/* Per cluster (level order; null outputs skipped): point range begin/size,
 * sons (2 per cluster, -1 for none), the order of its basis Q (qdim), its
 * scaling functions (nscal) and its first coefficient (start).  Node fields
 * of SampletTreeBase (SampletTreeBase.h:24-30, 192-197). */
fmca_status fmca_samplet_tree_clusters(const fmca_samplet_tree *tree, int64_t *begin,
                                       int64_t *size, int32_t *sons, int32_t *qdim,
                                       int32_t *nscal, int64_t *start);

/* SampletTreeBase::Q() and H2SampletTreeBase::V() of one cluster: Q is
 * qdim x qdim, V is q x qdim, both column-major (null outputs skipped). */
fmca_status fmca_samplet_tree_basis(const fmca_samplet_tree *tree, int cluster, double *Q,
                                    double *V);

/* sampletTransform / inverseSampletTransform (SampletTreeBase.h:68-85,
 * py/FMCA.cpp:396-420): in and out are n x k column-major (leading
 * dimensions ld_in, ld_out >= n); the forward transform maps data in
 * cluster order to samplet coefficients.  Host only. */
fmca_status fmca_samplet_transform(const fmca_samplet_tree *tree, int inverse, const double *in,
                                   int64_t ld_in, int64_t k, double *out, int64_t ld_out);

/* Replaces pySampletKernelCompressor(hst, ker, P, eta, thres)
 * (py/FMCA.cpp:256-300): SampletMatrixCompressor<H2SampletTree> with
 * CompareCluster -- init (samplet_matrix_compressor.h:37-71), compress
 * (:73-186, with recursivelyComputeBlock :330-388) over NystromEvaluator
 * (FMCA/src/MatrixEvaluator/NystromEvaluator.h:39-67) -- and triplets()
 * (:222-256): the upper triangle of the kernel matrix in samplet
 * coordinates, entries with |v| > threshold and the whole diagonal.  The
 * kernel blocks of the compression (dense leaf blocks and interpolated far
 * blocks, each with its cluster-basis products) are evaluated on the current
 * HIP device; their device memory is transient and counted against
 * fmca_set_device_memory_limit.  FMCA_ERR_NO_DEVICE without a device. */
fmca_status fmca_samplet_compress(const fmca_samplet_tree *tree, const fmca_kernel_desc *kernel,
                                  double eta, double threshold, fmca_samplet_matrix **out);
fmca_status fmca_samplet_matrix_destroy(fmca_samplet_matrix *mat);

/* Number of triplets, and the triplets (row, column, value), rows <= columns,
 * in samplet coordinates (null outputs skipped). */
fmca_status fmca_samplet_matrix_nnz(const fmca_samplet_matrix *mat, int64_t *nnz);
fmca_status fmca_samplet_matrix_triplets(const fmca_samplet_matrix *mat, int64_t *rows,
                                         int64_t *cols, double *vals);

/* aposteriori_triplets_fast(thres) (samplet_matrix_compressor.h:258-290):
 * *count kept triplets; index (nullable, nnz entries of room) receives their
 * positions in the triplet list. */
fmca_status fmca_samplet_matrix_aposteriori(const fmca_samplet_matrix *mat, double thres,
                                            int64_t *count, int64_t *index);

/* Work and time of the compression: {pattern entries, dense jobs, far jobs,
 * job block elements, program nodes, plan s, upload s, GPU s, download s,
 * combine s, triplets s, triplets}.  No FMCA counterpart. */
enum { FMCA_SAMPLET_STATS = 12 };
fmca_status fmca_samplet_matrix_stats(const fmca_samplet_matrix *mat, double *stats);

#ifdef __cplusplus
}
#endif

#endif /* FMCA_AMD_H_ */
