// Planner for the MI355X H2-matrix kernel action.
//
// Builds FMCA's cluster tree, block pattern and interpolation data, flattened
// into structure-of-arrays that are uploaded to HBM once.  Cluster ids are
// FMCA's level-order block ids.  plan.cpp runs on the host alone, so the
// planner can be exercised without a GPU; build_tree_device and
// expand_blocks_device (tree_device.hip, pattern_device.hip) move its O(n)
// and O(#blocks) parts onto the device with the same result.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace fmca_amd {

// CovarianceKernel types (FMCA/src/CovarianceKernel/CovarianceKernel.h:56-116).
enum KernelType : int {
  KT_EXPONENTIAL = 0,
  KT_MATERN32 = 1,
  KT_MATERN52 = 2,
  KT_MATERN72 = 3,
  KT_MATERN92 = 4,
  KT_GAUSSIAN = 5,
  KT_INVMULTIQUADRIC = 6,
  KT_MULTIQUADRIC = 7,
  KT_BIHARMONIC2D = 8,
  KT_BIHARMONIC3D = 9,
  KT_TRIHARMONIC3D = 10,
  KT_TPS2D = 11,
  KT_COUNT = 12,
};

// Case-insensitive name -> KernelType, -1 if FMCA does not define it.
int parse_kernel_type(const char *name);
const char *kernel_type_name(int kt);

constexpr int kMaxDim = 3;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X
constexpr int kTaskSplit = 1, kTaskFar = 2;  // task flags
constexpr int kMaxP1 = 10;  // polynomial degree <= 9

// Interpolation schemes of NystromMoments<Interpolator>.
enum InterpScheme : int {
  IS_TENSOR = 0,        // TensorProductInterpolator
  IS_TOTAL_DEGREE = 1,  // TotalDegreeInterpolator
};

// Interpolation on [0,1]^d.
//
// Tensor scheme (FMCA/src/Interpolator/TensorProductInterpolator.h:39-81):
// Chebyshev nodes xi and barycentric weights w per axis, invV = I.  Moments
// are held in tensor order lin = a_0 + P1*a_1 + P1^2*a_2; fmca_order maps
// FMCA's graded-lexicographic position k (MultiIndexSet.h:87-111) to lin.
//
// Total-degree scheme (FMCA/src/Interpolator/TotalDegreeInterpolator.h:27-48):
// the q = binom(deg+d, d) products of L2-normalised shifted Legendre
// polynomials (evalPolynomials.h:19-66) with |alpha|_1 <= deg, in FMCA's
// graded-lexicographic order (fmca_order is the identity); interpolation
// nodes are DeMarchi points (DeMarchiPoints.h:22-50) and invV = V^-1 with
// V(i, a) = P_a(node i).
struct Interpolation {
  int dim = 0, deg = 0, P1 = 0, q = 0;
  int scheme = IS_TENSOR;
  std::vector<double> xi, w;
  std::vector<int32_t> fmca_order;
  std::vector<int32_t> midx;   // total degree: a_0 | a_1 << 8 | a_2 << 16 per basis function
  std::vector<double> nodes;   // total degree: dim x q, node k in column k
  std::vector<double> invV;    // total degree: q x q, column-major
  void init(int dim, int deg, int scheme = IS_TENSOR);
  // l_a(u), a = 0..deg, with FMCA's rule that a factor within eps of its node
  // is 1 (TensorProductInterpolator.h:71-78).
  void lagrange_1d(double u, double *out) const;
  // sqrt(2a+1) P_a(2u-1), a = 0..deg (LegendrePolynomials::eval).
  void legendre_1d(double u, double *out) const;
  // all q basis values at a point of [0,1]^d (total degree)
  void basis(const double *u, double *out) const;
};

// A cluster tree in FMCA's level order.  Clusters of one level are contiguous.
struct Tree {
  int dim = 0;
  int64_t npts = 0;
  int ncl = 0;
  int depth = 0;                     // deepest level
  std::vector<uint32_t> perm;        // cluster order position -> point index
  std::vector<int64_t> begin, size;  // per cluster
  std::vector<int32_t> level, son0, son1, parent;
  std::vector<int32_t> level_ptr;    // clusters of level L: [level_ptr[L], level_ptr[L+1])
  std::vector<double> bbmin, bbmax, bbsize;  // ncl x dim
  std::vector<int32_t> leaves;       // leaf cluster ids, ascending
  bool is_leaf(int c) const { return son0[c] < 0; }
};

// ClusterTree with CardinalityBisection and shrink-to-fit boxes
// (FMCA/src/Clustering/ClusterTreeInitializer_ClusterTree.h:24-132,
//  FMCA/src/Clustering/ClusterTreeSplitter.h:60-81).
void build_tree(Tree &T, int dim, int64_t n, const double *P, int64_t min_csize);

// A range of the cluster-order permutation and a coordinate axis.
struct SplitSeg {
  int64_t begin, size;
  int axis;  // -1: none (a leaf without a father)
};

// The data-parallel part of a tree build: the host planner's splitter and
// the device build each implement it; build_tree_with does the rest.
struct TreeSplitter {
  virtual ~TreeSplitter() = default;
  // For every segment k: reorder its range so that its size / 2 points of
  // smallest coordinate on the axis come first; lmax[k] = the largest of
  // them, rmin[k] = the smallest of the others.  False if, for some k,
  // lmax[k] == rmin[k] (equal coordinates at a split point) and the build
  // must be redone exactly.
  virtual bool split(const std::vector<SplitSeg> &segs, double *lmax, double *rmin) = 0;
  // For every leaf k (axis: its father's split axis): put it in its father's
  // sort order if that is still to be done (false on equal coordinates in the
  // leaf), and store the min / max over its points in mn / mx[k * dim + i].
  virtual bool finish_leaves(const std::vector<SplitSeg> &leaves, double *mn, double *mx) = 0;
  // Hand over the final permutation.
  virtual void permutation(std::vector<uint32_t> &perm) = 0;
};

// f(lo, hi) over contiguous chunks of [0, n) on the planner's worker threads
// (FMCA_AMD_THREADS, else OMP_NUM_THREADS, else the hardware concurrency).
void parallel_ranges(int64_t n, const std::function<void(int64_t, int64_t)> &f);

// f(i) for i in [0, n) on the planner's worker threads, dynamically
// scheduled in chunks of `grain`.
void parallel_dynamic(int64_t n, int64_t grain, const std::function<void(int64_t)> &f);

// Min / max of each coordinate over n points.
void point_box(int dim, int64_t n, const double *P, double *mn, double *mx);

// Device memory of the library's own accounting (actor.cpp): device_malloc
// counts the bytes against the cap of fmca_set_device_memory_limit and
// returns a hipError_t (hipErrorOutOfMemory beyond the cap); device_free
// releases what device_malloc gave.
int device_malloc(void **p, size_t bytes);
void device_free(void *p, size_t bytes);

// build_tree with the O(n) work on the current HIP device (tree_device.hip).
// Returns 1 if built there, 0 if built on the host instead (equal coordinates
// at a split point or inside a leaf, a -0.0 coordinate, or leaves too large for
// the device sort; the result is the same either way), minus the hipError_t
// code on a HIP error (message in *err; T is then unspecified).
int build_tree_device(Tree &T, int dim, int64_t n, const double *P, int64_t min_csize,
                      std::string *err);

// The level loop of build_tree with a given splitter; root_min / root_max:
// the points' coordinate box.  False if the splitter reported equal keys.
bool build_tree_with(Tree &T, int dim, int64_t n, const double *P, const double *root_min,
                     const double *root_max, int64_t min_csize, TreeSplitter &sp);

// FMCA's leaf blocks (H2MatrixBase::computePattern, H2MatrixBase.h:160-191),
// regrouped by target for a gather-style GPU schedule.
struct Pattern {
  // every leaf block of FMCA's block tree (row id, column id, low-rank flag)
  std::vector<int32_t> block_row, block_col;
  std::vector<int8_t> block_lr;
  // far field: CSR over target clusters -> source clusters
  std::vector<int32_t> far_ptr, far_src;
  // near field: CSR over target leaves (index into Tree::leaves) -> source clusters
  std::vector<int32_t> near_ptr, near_src;
  // Work items of the two entry phases, 4 ints each: target (cluster id for
  // the coupling, leaf index for the leaf output), begin and end in its far /
  // near list, flags.  A target with more than three times the mean work is cut
  // into contiguous pieces of its list (flag kTaskSplit: its pieces add into
  // an output that is zeroed first); the first piece of a leaf also applies
  // the far field (kTaskFar).
  std::vector<int32_t> far_task, near_task;
  // near_run[e]: the number of near-list entries from e on, within e's task,
  // whose point ranges follow each other without a gap (1 if the next one
  // does not): the leaf output stages such a run as one source
  std::vector<int32_t> near_run;
  std::vector<int32_t> far_split, near_split;  // targets cut into several tasks
  // launch orders over the tasks (one workgroup per entry, -1 = nothing to
  // do, see xcd_launch_order)
  std::vector<int32_t> far_order, near_order;
  double near_entries = 0, far_entries = 0;
  int64_t n_far = 0, n_near = 0;
};

void compute_pattern(Pattern &pat, const Tree &rows, const Tree &cols, double eta, int q);

// The two halves of compute_pattern.  expand_blocks: the leaf blocks
// (block_row, block_col, block_lr), level by level, each level in the order
// of its parents' refinements.  finish_pattern: everything else, from them.
void expand_blocks(Pattern &pat, const Tree &rows, const Tree &cols, double eta, int q);
void finish_pattern(Pattern &pat, const Tree &rows, const Tree &cols, int q);

// expand_blocks on the current HIP device, with the same blocks in the same
// order.  Returns 1 if expanded there; 0 if some cluster pair's admissibility
// lies within rounding of eta (distance and radius compared on the device to
// a margin; such a plan is expanded on the host instead); minus the hipError_t
// code on a HIP error (message in *err).
int expand_blocks_device(Pattern &pat, const Tree &rows, const Tree &cols, double eta, int q,
                         std::string *err);

// Half the box diagonal of cluster c (Admissibility.h:19-29).
double cluster_radius(const Tree &T, int c);

// Launch order of one workgroup per item on MI355X, whose 8 XCDs (one L2
// each) receive the workgroups of a launch round-robin (workgroup b on XCD
// b % 8).  Each range [lo, hi) of items, taken in the order given, is cut into
// nxcd contiguous pieces of about equal work, piece x for XCD x: every XCD
// then works on tree neighbours, which read the same sources through its L2,
// and the XCDs get equal work.  order[nxcd * k + x] is the k-th item of XCD x,
// -1 where XCD x has fewer items.
std::vector<int32_t> xcd_launch_order(const std::vector<double> &work,
                                      const std::vector<std::pair<int, int>> &ranges, int nxcd);

// Transfer factors: E_c = T_c,0 (x) T_c,1 (x) T_c,2 for every non-root cluster c
// (FMCA/src/H2Matrix/compute_cluster_bases_impl.h:36-51), stored as
// T[((c*dim + i)*P1 + a)*P1 + b] = l_a(x_i(b)) with x_i(b) the son's node b in
// the father's reference frame.
void compute_transfers(std::vector<double> &T, const Tree &tree, const Interpolation &ip);

// Total-degree transfers (compute_cluster_bases_impl.h:36-51): for every
// non-root cluster c the dense q x q matrix
//   E_c[a * q + b] = sum_j P^father_a(x(j)) invV[b][j],
// x(j) = node j of c in the father's reference frame; zero for empty sons.
void compute_transfers_dense(std::vector<double> &E, const Tree &tree, const Interpolation &ip);

struct Statistics {
  double rows, cols, low_rank_blocks, dense_blocks, nz_per_row, storage_gb;
};
// H2MatrixBase::statistics (H2MatrixBase.h:70-105)
Statistics statistics(const Pattern &pat, const Tree &rows, const Tree &cols, int q);

}  // namespace fmca_amd
