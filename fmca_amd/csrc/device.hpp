// Device-side views of the plan and the launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

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
  const T *pts = nullptr;           // dim x npts, cluster order
  const T *bbmin = nullptr;         // ncl x dim
  const T *bbsize = nullptr;        // ncl x dim
  const int32_t *begin = nullptr;   // ncl
  const int32_t *size = nullptr;    // ncl
  const int32_t *son0 = nullptr;    // ncl
  const int32_t *son1 = nullptr;    // ncl
  const int32_t *parent = nullptr;  // ncl
  const int32_t *leaves = nullptr;  // nleaves
  const T *transfer = nullptr;      // ncl x dim x P1 x P1, [a][b] = l_a(x(b))
};

// Interaction lists of one (target tree, source tree) pair.
struct DevPatternView {
  int nfar_targets = 0;
  const int32_t *far_ptr = nullptr;      // ncl_target + 1
  const int32_t *far_src = nullptr;
  const int32_t *far_targets = nullptr;  // targets with a non-empty far list
  int nnear_leaves = 0;
  const int32_t *near_ptr = nullptr;     // nleaves_target + 1
  const int32_t *near_src = nullptr;
};

template <typename T>
struct InterpParams {
  int P1 = 0, q = 0;
  T xi[kMaxP1];
  T w[kMaxP1];
};

template <typename T>
struct KernelParams {
  int kt = 0;
  T l = 1, c = 1, nu = 1;
};

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

// Enqueues the whole matrix-free action out = K(rows, cols) * rhs on `stream`.
// rhs / out are FP64 column-major device arrays; perm_in / perm_out select the
// natural order (true) or FMCA's cluster order (false).
template <typename T>
hipError_t launch_action(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const HostLevels &lrows, const HostLevels &lcols,
                         const DevPatternView &pat, const InterpParams<T> &ip,
                         const KernelParams<T> &kp, const double *rhs, int64_t ld_rhs, int nrhs,
                         double *out, int64_t ld_out, bool natural_order, ActionBuffers<T> buf,
                         hipStream_t stream);

}  // namespace fmca_amd
