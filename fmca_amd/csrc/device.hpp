// Device-side views of the plan and the launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "plan.hpp"

namespace fmca_amd {

// One cluster tree resident in HBM, level order, points in cluster order.
// Coordinates are stored dimension-major (pts[i * npts + j]) so that a wave
// reading consecutive points of one dimension is coalesced.
template <typename T>
struct DevTreeView {
  int dim = 0, ncl = 0, nleaves = 0, depth = 0;
  int64_t npts = 0;
  const uint32_t *perm = nullptr;   // cluster position -> natural index
  const uint32_t *pinv = nullptr;   // natural index -> cluster position
  const T *pts = nullptr;           // dim x npts, cluster order
  const T *unit = nullptr;          // dim x npts: (x - bbmin) / bbsize of the point's leaf
  const T *bbmin = nullptr;         // ncl x dim
  const T *bbsize = nullptr;        // ncl x dim
  const int32_t *begin = nullptr;   // ncl
  const int32_t *size = nullptr;    // ncl
  const int32_t *son0 = nullptr;    // ncl
  const int32_t *son1 = nullptr;    // ncl
  const int32_t *parent = nullptr;  // ncl
  const int32_t *leaves = nullptr;  // nleaves
  const T *transfer = nullptr;      // tensor: ncl x dim x P1 x P1, [a][b] = l_a(x(b));
                                    // total degree: ncl x q x q, [a][b] = E_c(a, b)
};

// Interaction lists of one (target tree, source tree) pair.
// Tasks hold 4 ints (target, list begin, list end, flags kTaskSplit /
// kTaskFar); a split target's tasks add atomically into a zeroed output.
struct DevPatternView {
  const int32_t *far_ptr = nullptr;     // ncl_target + 1
  const int32_t *far_src = nullptr;
  int nfar_order = 0;                   // coupling grid size
  const int32_t *far_order = nullptr;   // workgroup -> coupling task (-1: none)
  const int32_t *far_task = nullptr;    // target cluster, far list range, flags
  int nfar_split = 0;
  const int32_t *far_split = nullptr;   // target clusters with several tasks
  int nnear_leaves = 0;
  const int32_t *near_ptr = nullptr;    // nleaves_target + 1
  const int32_t *near_src = nullptr;
  const int32_t *near_run = nullptr;    // contiguous sources from each entry on
  int nnear_order = 0;                  // leaf output grid size
  const int32_t *near_order = nullptr;  // workgroup -> leaf output task (-1: none)
  const int32_t *near_task = nullptr;   // target leaf index, near list range, flags
  int nnear_split = 0;
  const int32_t *near_split = nullptr;  // target leaf indices with several tasks
  // smallest target / source cluster id of any far block (ids are level
  // ordered, so the tree passes above their levels carry only zeros); 0: no
  // information, the passes run over every level
  int far_top_target = 0, far_top_source = 0;
};

// Interpolation data of the kernels.  Tensor scheme: the per-axis Chebyshev
// nodes xi and weights w.  Total-degree scheme (td = 1): the packed
// multi-index of every basis function (midx, a_0 | a_1 << 8 | a_2 << 16), the
// unit interpolation nodes unodes[d * q + k] and invV (q x q, column-major),
// all in device memory.
template <typename T>
struct InterpParams {
  int P1 = 0, q = 0;
  int td = 0;
  T xi[kMaxP1];
  T w[kMaxP1];
  const int32_t *midx = nullptr;
  const T *unodes = nullptr;
  const T *invV = nullptr;
};

// Kernel parameters with the radial scale folded in on the host, so that no
// division is evaluated per kernel entry:
//   EXPONENTIAL s = 1/l, MATERN32 s = sqrt(3)/l, MATERN52 s = sqrt(5)/l,
//   MATERN72 s = sqrt(7)/l, MATERN92 s = 3/l, GAUSSIAN s = -1/(2 l^2) (on r^2).
//
// kt >= KT_COUNT selects a derivative of kernel kt % KT_COUNT (variant
// kt / KT_COUNT): 1 = dk/dl, 2 = phi(r) = k'(r) / r (the radial factor of the
// gradient with respect to a point, phi(0) = 0 where k' / r is singular).
//
// Several variants in one action (MFMA forms): tv[g] is the variant of rhs
// columns [16 g, 16 g + 16) and vmask the OR of 1 << tv[g]; the kernel entry
// of every (target, source) pair is then evaluated once for all variants.
constexpr int kMaxVariantGroups = 8;

template <typename T>
struct KernelParams {
  int kt = 0;
  T l = 1, c = 1, nu = 1;
  T s = 1;      // radial scale (see above)
  T sl2 = 1;    // s * log2(e): FP32 exponentials are hardware exp2
  T inv_l = 1;  // 1 / l
  T c2 = 1;     // c * c
  int vmask = 1;
  int tv[kMaxVariantGroups] = {0, 0, 0, 0, 0, 0, 0, 0};
  int split_bf16 = 0;  // FP32 form: kernel-value contractions on BF16 matrix cores
  // distance of the kernel entries: 0 Euclidean, 1 geodesic on the unit
  // sphere, acos(clamp(x . y, -1, 1)) (CovarianceKernel.h:123-137 with
  // SphereClusterTree.h:65-69); geodesic entries run in FP64 on the VALU forms
  int geodesic = 0;
};

enum KernelVariant : int { KV_VALUE = 0, KV_DL = 1, KV_PHI = 2 };

// The radial scale s of KernelParams for kernel type kt and lengthscale l.
inline double kernel_radial_scale(int kt, double l) {
  switch (kt) {
    case KT_MATERN32:
      return std::sqrt(3.0) / l;
    case KT_MATERN52:
      return std::sqrt(5.0) / l;
    case KT_MATERN72:
      return std::sqrt(7.0) / l;
    case KT_MATERN92:
      return 3.0 / l;
    case KT_GAUSSIAN:
      return -0.5 / (l * l);
    default:
      return 1.0 / l;
  }
}

// Level structure of a tree on the host (drives the per-level launches).
struct HostLevels {
  int depth = 0;
  const int32_t *level_ptr = nullptr;  // depth + 2 entries
};

// Workspace of one action: permuted rhs, source moments, target moments.
template <typename T>
struct ActionBuffers {
  T *xs = nullptr;    // npts_src x nrhs (point-major, rhs fastest)
  T *mom = nullptr;   // ncl_src x q x nrhs
  T *tlhs = nullptr;  // ncl_tgt x q x nrhs
};

// Phases of one action, in launch order (timed with events when profiling).
enum Phase : int {
  PH_PERMUTE = 0,
  PH_UPWARD_LEAF = 1,
  PH_UPWARD_TRANSFER = 2,
  PH_COUPLING = 3,
  PH_DOWNWARD_TRANSFER = 4,
  PH_LEAF_OUTPUT = 5,
  PH_COUNT = 6
};

// ev[k] is recorded on the launch stream before phase k, ev[PH_COUNT] after
// the last phase.
struct PhaseEvents {
  hipEvent_t ev[PH_COUNT + 1];
};

// The blocks of an assembled H2 matrix (fmca_h2_assemble), kept in HBM as
// v_mfma 16x16x4 left-operand tiles.  In a block of R rows (a multiple of 16)
// and K columns (a multiple of 4 G), the 16 x 4 tile (rt, ks) gives lane
// (j % 4) * 16 + i % 16 element (i, j), the value the instruction takes from
// that lane; the G = 16 / sizeof(T) consecutive k-steps ks = g G + u of a
// lane are adjacent, at ((rt * K / (4 G) + g) * 64 + lane) * G + u, so one
// wave reads G tiles with one 16-byte load per lane.  Padding is zero.
constexpr int kStoredMaxRows = 256;  // rows of a stored block (q and leaf sizes)
template <typename T>
constexpr int kStoredGroup = 16 / (int)sizeof(T);
template <typename T>
__host__ __device__ inline int stored_kpad(int k) {
  constexpr int m = 4 * kStoredGroup<T>;
  return (k + m - 1) / m * m;
}
template <typename T>
struct StoredBlocks {
  const T *far = nullptr;             // far block b (far list order): S_ts, Rq x Kq
  const T *near = nullptr;            // per leaf: V_t^T (Rt x Kq), then its near blocks
  const int64_t *near_off = nullptr;  // near list entry e -> offset of K(P_t, P_s), Rt x kpad|s|
  const int64_t *v_off = nullptr;     // leaf index k -> offset of V_t^T
  // launch orders, longest list first: workgroup -> target cluster / leaf index
  const int32_t *far_order = nullptr, *near_order = nullptr;
  int Rq = 0, Kq = 0;                 // pad16(q), stored_kpad(q)
};

// Enqueues the whole action out = K(rows, cols) * rhs on `stream`: matrix-free,
// or with the coupling and leaf-output phases streaming stored blocks.
// rhs / out are FP64 column-major device arrays; natural_order selects the
// natural order (true) or FMCA's cluster order (false).
template <typename T>
hipError_t launch_action(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const HostLevels &lrows, const HostLevels &lcols,
                         const DevPatternView &pat, const InterpParams<T> &ip,
                         const KernelParams<T> &kp, const double *rhs, int64_t ld_rhs, int nrhs,
                         double *out, int64_t ld_out, bool natural_order, ActionBuffers<T> buf,
                         hipStream_t stream, const PhaseEvents *events = nullptr,
                         const StoredBlocks<T> *stored = nullptr,
                         const struct GradJob *grad = nullptr);

// Gradient of <G, K~ alpha> for one shared tree (kernels_grad.hip): the
// action on [alpha | G | alpha] (columns [0, R), [R, 2R), [2R, 2R + 16), the
// last group with dk/dl) stops after the downward pass, and these kernels
// take its moments and far sums.  Outputs are FP64, natural order, added
// atomically into arrays the caller has zeroed.
struct GradJob {
  int r = 0;            // right-hand sides (columns of alpha and G), r <= R
  int R = 8;            // 8 or 16
  int ncol = 0;         // 2 R + 16
  double *dalpha = nullptr;   // n x r, column stride ld_dalpha
  int64_t ld_dalpha = 0;
  double *dpoints = nullptr;  // dim x n, row stride ld_points
  int64_t ld_points = 0;
  double *dl_part = nullptr;  // nnear_parts + grad_far_dot_blocks() partial sums
  int nnear_parts = 0;        // = the leaf-output grid size
  double *dl_out = nullptr;   // one FP64 value
};
int grad_far_dot_blocks();
template <typename T>
hipError_t launch_grad_far_dot(const GradJob &job, int64_t nmom, const T *mom, const T *tlhs,
                               hipStream_t stream);
template <typename T>
hipError_t launch_grad_leaves(const DevTreeView<T> &tree, const DevPatternView &pat,
                              const InterpParams<T> &ip, const KernelParams<T> &kp,
                              const GradJob &job, const T *xs, const T *tlhs,
                              hipStream_t stream);
hipError_t launch_grad_zero(double *p, int64_t n, hipStream_t stream);

// Writes every block of the assembled matrix (assembled.hip).  far_tgt: far
// block -> target cluster; near_leaf: near list entry -> leaf index.
template <typename T>
hipError_t launch_assemble(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                           const DevPatternView &pat, const InterpParams<T> &ip,
                           const KernelParams<T> &kp, int64_t nfar, const int32_t *far_tgt,
                           int64_t nnear, const int32_t *near_leaf, T *far, T *near,
                           const int64_t *near_off, const int64_t *v_off, int Rq, int Kq,
                           hipStream_t stream);

// The coupling (output = false: tlhs) or the leaf output (output = true: out)
// from stored blocks.
template <typename T>
hipError_t launch_stored(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const DevPatternView &pat, const StoredBlocks<T> &st, int q, int nrhs,
                         const T *xs, const T *mom, T *tlhs, double *out, int64_t ldo,
                         bool natural, hipStream_t stream);

// Matrix-core (MFMA) forms of the coupling (output = false) and the leaf
// output (output = true), used from kMfmaMinRhs right-hand sides on when
// their LDS fits (td = 1: total-degree scheme).
constexpr int kMfmaMinRhs = 8;

// kt, vmask: the kernel parameters of the launch (several variants take
// 64-column passes, also in FP64).
template <typename T>
bool mfma_supported(bool output, int D, int P1, int q, int nrhs, int kt, int vmask, int td);

// Leaf moments of the FP32 forms on the matrix cores (kernels_mfma.hip):
// q <= 128 and >= kMfmaMinRhs rhs where the LDS staging fits.
bool upward_leaf_mfma_supported(int D, int P1, int q, int nrhs);
hipError_t launch_upward_leaf_mfma(const DevTreeView<float> &tr, const InterpParams<float> &ip,
                                   const float *xs, int nrhs, float *mom, hipStream_t stream);

// BF16 three-term forms of the FP32 leaf output / coupling (kernels_bf16.hip):
// kernel values only, >= kMfmaMinRhs rhs, where the LDS staging fits.
bool bf16_supported(bool output, int D, int P1, int q, int nrhs, int kt, int vmask);
// The right operand the launch reads (xs for the leaf output, mom for the
// coupling) is split into BF16 halves in place first, so it holds hi | lo << 16
// words afterwards.
hipError_t launch_bf16(bool output, const DevTreeView<float> &rows,
                       const DevTreeView<float> &cols, const DevPatternView &pat,
                       const InterpParams<float> &ip, const KernelParams<float> &kp, int nrhs,
                       float *xs, float *mom, float *tlhs, double *out, int64_t ldo,
                       bool natural, hipStream_t stream);

template <typename T>
hipError_t launch_mfma(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                       const DevPatternView &pat, const InterpParams<T> &ip,
                       const KernelParams<T> &kp, int nrhs, const T *xs, const T *mom, T *tlhs,
                       double *out, int64_t ldo, bool natural, hipStream_t stream);

// Column-batched FP64 vector kernels (blas1.hip); columns of n x nrhs arrays.
size_t blas1_dot_scratch(int nrhs);
hipError_t blas1_dot(int64_t n, int nrhs, const double *a, int64_t lda, const double *b,
                     int64_t ldb, double *out, double *scratch, hipStream_t stream);
hipError_t blas1_axpy(int64_t n, int nrhs, double s, const double *coef, const double *x,
                      int64_t ldx, double *y, int64_t ldy, hipStream_t stream);
hipError_t blas1_xpby(int64_t n, int nrhs, const double *coef, const double *z, int64_t ldz,
                      double *p, int64_t ldp, hipStream_t stream);
hipError_t blas1_residual(int64_t n, int nrhs, double ridge, const double *b, int64_t ldb,
                          const double *y, int64_t ldy, const double *x, int64_t ldx, double *r,
                          int64_t ldr, hipStream_t stream);
hipError_t blas1_add_ridge(int64_t n, int nrhs, double ridge, const double *x, int64_t ldx,
                           double *y, int64_t ldy, hipStream_t stream);
hipError_t blas1_fill(int64_t n, int nrhs, double v, double *x, int64_t ldx, hipStream_t stream);
// out[i, j] = in[perm[i], j] (gather = true) or out[perm[i], j] = in[i, j]
hipError_t blas1_permute(int64_t n, int nrhs, const uint32_t *perm, bool gather, const double *in,
                         int64_t ldi, double *out, int64_t ldo, hipStream_t stream);

// Krylov basis of the GCR solve (blas1.hip): a circular window of `count`
// pairs (p_j, q_j = (K + ridge I) p_j), window position j in slot
// (first + j) % cap, slots of n x nrhs values of type S at ptab / qtab (device
// arrays of slot pointers); qq[slot * nrhs + col] = q_slot . q_slot.
// Columns with active[col] == 0 are left as they are.  Deterministic sums.
constexpr int kBasisMax = 256;
size_t basis_scratch(int nrhs, int count);
// coef[col * count + j] = (v . q_j) / (q_j . q_j)
template <typename S>
hipError_t basis_dots(int64_t n, int nrhs, const double *v, const S *const *qtab, int first,
                      int count, int cap, const double *qq, const int *active, double *coef,
                      double *scratch, hipStream_t stream);
// q = y - sum_j coef_j q_j, p = r - sum_j coef_j p_j into (qk, pk);
// qqrq[2 col] = q.q, qqrq[2 col + 1] = r.q
template <typename S>
hipError_t basis_new_pair(int64_t n, int nrhs, const double *y, const double *r,
                          const double *coef, const S *const *qtab, const S *const *ptab,
                          int first, int count, int cap, S *qk, S *pk, const int *active,
                          double *qqrq, double *scratch, hipStream_t stream);
// x += a p_k, r -= a q_k with a = r.q / q.q; qq_slot[col] = q.q; rr[col] = r.r
template <typename S>
hipError_t basis_update(int64_t n, int nrhs, const S *pk, const S *qk, const double *qqrq,
                        double *qq_slot, double *x, double *r, const int *active, double *rr,
                        double *scratch, hipStream_t stream);
// x += sum_j coef_j p_j, r -= sum_j coef_j q_j; rr[col] = r.r
template <typename S>
hipError_t basis_project(int64_t n, int nrhs, const double *coef, const S *const *qtab,
                         const S *const *ptab, int first, int count, int cap, double *x, double *r,
                         const int *active, double *rr, double *scratch, hipStream_t stream);

// Evaluates the device exp (fn 0 / 2) or sqrt (fn 1 / 3) used by the kernel
// loops, in FP64 (fn 0, 1) or FP32 (fn 2, 3), on n device values.
hipError_t launch_math_probe(int fn, const double *x, double *y, int64_t n, hipStream_t stream);

}  // namespace fmca_amd
