// Samplet trees and the samplet compression of kernel matrices, with the
// kernel blocks that the compression consumes evaluated on the GPU.
//
// FMCA builds this from H2SampletTree<ClusterTree>
// (FMCA/src/Samplets/H2SampletTree.h:53-137), MinNystromSampletMoments
// <MonomialInterpolator> (FMCA/src/Moments/MinNystromSampletMoments.h),
// NystromMoments<TotalDegreeInterpolator> for the multiscale cluster bases
// and SampletMatrixCompressor with NystromEvaluator
// (FMCA/src/Samplets/samplet_matrix_compressor.h:19-430,
//  FMCA/src/MatrixEvaluator/NystromEvaluator.h:25-70), as its Python binding
// assembles them (py/FMCA.cpp:51-118, 256-330).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "plan.hpp"

namespace fmca_amd {

// One samplet tree.  Cluster ids are the tree's level-order ids.
struct SampletTree {
  int dim = 0;
  int dtilde = 1;        // vanishing moments: polynomials of degree < dtilde
  int m = 1;             // mdtilde = binom(dim + dtilde - 1, dim) scaling functions
  int64_t n = 0;
  Tree tree;             // ClusterTree with minimum leaf size ip.q
  Interpolation ip;      // TotalDegreeInterpolator of degree 2 (dtilde - 1)
  std::vector<int32_t> mono;    // the m monomial exponents a_0 | a_1 << 8 | a_2 << 16
  std::vector<double> mcoef;    // m x m column-major, (j, i) = prod_k binom(a_j[k], a_i[k])
  // per cluster: Q is qdim x qdim (column-major at q_off), its first nscal
  // columns the scaling functions and the other nsamp the samplets in the
  // basis of the cluster's points (leaf) or of its sons' scaling functions
  std::vector<int32_t> qdim, nscal, nsamp;
  std::vector<int64_t> start;   // first samplet coefficient (sampletMapper)
  std::vector<int64_t> q_off, v_off;
  std::vector<double> Q;
  // multiscale cluster basis V_t (q x qdim, column-major at v_off):
  // leaf  V_t = V^leaf_t Q_t,  inner  V_t = [E_s0 V_s0[:, :nscal], E_s1 V_s1[:, :nscal]] Q_t
  std::vector<double> V;
  std::vector<double> E;        // compute_transfers_dense
  std::vector<double> pts;      // points in cluster order, pts[i * n + j] (coordinate i)
};

// H2SampletTree(mom, samp_mom, 0, P) with mom = NystromMoments<TotalDegree>(P,
// 2 (dtilde - 1)) and samp_mom = MinNystromSampletMoments<Monomial>(P,
// dtilde - 1) (py/FMCA.cpp:53-58).  P: dim x n column-major.  Empty string on
// success, else the reason.
std::string build_samplet_tree(SampletTree &S, int dim, int64_t n, const double *P, int dtilde);

// sampletTransform / inverseSampletTransform (SampletTreeBase.h:68-85,
// 160-208): in and out are n x k column-major with leading dimensions ld_in,
// ld_out; the forward transform takes data in cluster order and returns
// samplet coefficients, the inverse the converse.
void samplet_transform(const SampletTree &S, const double *in, int64_t ld_in, int64_t k,
                       double *out, int64_t ld_out);
void inverse_samplet_transform(const SampletTree &S, const double *in, int64_t ld_in, int64_t k,
                               double *out, int64_t ld_out);

// sampletLevelMapper (SampletMapper.h:34-50): the level of every coefficient.
std::vector<int32_t> samplet_levels(const SampletTree &S);

// A kernel block the compression consumes, evaluated on the GPU:
//   kind 0 (compute_dense_block, both leaves):  Q_r^T K(P_r, P_c) Q_c
//   kind 1 (interpolate_kernel, admissible):    W_r^T K(xi_r, xi_c) W_c,  W = invV^T V
// out: qdim_r x qdim_c column-major at element offset off.
struct SampletJob {
  int32_t kind, r, c, pad;
  int64_t off;
};

// The compression as a program over cluster pairs.  Node k computes the
// full block of the pair (row[k], col[k]) in the cluster bases Q:
//   op 0  a GPU job (arg[0] = job index)
//   op 1  columns from the sons of col:  [B(r, c0)[:, :nscal], B(r, c1)[:, :nscal]] Q_c
//   op 2  rows from the sons of row:     Q_r^T [B(r0, c)[:nscal, :]; B(r1, c)[:nscal, :]]
//   op 3  both:  Q_r^T [B(r_i, c_j)[:nscal_i, :nscal_j]]_ij Q_c
// Children come before their parents in the node order of each height.
struct SampletProgram {
  std::vector<int32_t> row, col, op, height;
  std::vector<int32_t> arg;      // 4 per node
  std::vector<int64_t> off;      // block offset (jobs: in the job buffer, others: in the combined buffer)
  std::vector<SampletJob> jobs;
  int64_t job_elems = 0, comb_elems = 0;
  int max_height = 0;
  // SampletMatrixCompressor::pattern_: the pattern entries and their nodes
  std::vector<int32_t> pat_row, pat_col, pat_node;
};

// SampletMatrixCompressor::init with CompareCluster (samplet_matrix_compressor.h
// :37-71) and the node program of compress (:73-186) with its
// recursivelyComputeBlock (:330-388).
void plan_samplet_compression(SampletProgram &prog, const SampletTree &S, double eta);

// Fills the combined blocks from the job blocks (host threads, by height).
void combine_samplet_blocks(const SampletProgram &prog, const SampletTree &S,
                            const double *job_blocks, std::vector<double> &comb);

// triplets() (samplet_matrix_compressor.h:222-256): the upper triangle of the
// compressed matrix in samplet coordinates, entries above threshold and the
// whole diagonal.
// A host array whose elements are not initialised on allocation (the
// triplet lists and job blocks run to several GB; they are written in full
// by parallel loops, which also spreads the first touch of their pages).
template <typename T>
struct HostArray {
  std::unique_ptr<T[]> p;
  size_t n = 0;
  void resize_uninit(size_t k) {
    p.reset(k ? new T[k] : nullptr);
    n = k;
  }
  T *data() { return p.get(); }
  const T *data() const { return p.get(); }
  size_t size() const { return n; }
};

// Two passes over the pattern on host threads: entry counts, then the fill
// at their prefix sums, so the order is FMCA's serial one.
void samplet_triplets(const SampletProgram &prog, const SampletTree &S, const double *job_blocks,
                      const std::vector<double> &comb, double threshold, HostArray<int64_t> &rows,
                      HostArray<int64_t> &cols, HostArray<double> &vals);

// aposteriori_triplets_fast (samplet_matrix_compressor.h:258-290): keep the
// magnitude decades, largest first, until the dropped part is below thres of
// the Frobenius norm; diagonal entries are always kept.  Returns the indices
// of the kept triplets in the order FMCA emits them.
std::vector<int64_t> samplet_aposteriori_fast(const int64_t *rows, const int64_t *cols,
                                              const double *vals, size_t count, double thres);

// The GPU part (samplets.hip): evaluate every job of the program into
// d_out (device, job_elems doubles).  Device inputs are prepared by the
// caller (see SampletDeviceData).
struct SampletDeviceData {
  const double *pts = nullptr;     // dim x n, cluster order, dim-major
  const double *Q = nullptr, *W = nullptr;
  const int64_t *q_off = nullptr, *v_off = nullptr, *begin = nullptr, *size = nullptr;
  const int32_t *qdim = nullptr;
  const double *bbmin = nullptr, *bbsize = nullptr, *unodes = nullptr;  // unodes: dim x q
  const SampletJob *jobs = nullptr;
  int64_t base = 0;                // job.off - base is the offset in the output buffer
  int64_t n = 0;
  int dim = 0, q = 0;
};
struct SampletKernel {
  int kt = 0;
  double l = 1, c = 1, nu = 1;
};
// hipError_t as int
int launch_samplet_jobs(const SampletDeviceData &d, const SampletKernel &k, int64_t njobs,
                        double *d_out, void *stream);

// The thread's C-ABI error message (fmca_last_error).
void set_last_error(const std::string &msg);

}  // namespace fmca_amd
