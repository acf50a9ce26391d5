// FP64 CPU oracle for the H2-matrix kernel action.
//
// TEST INFRASTRUCTURE ONLY.  This file restates FMCA's matrix-free H2 action
// (muchip/fmca) in plain C++ without Eigen so that the GPU product can be
// checked against it.  Only tests/, __graft_entry__.smoke() and the CPU
// baseline leg of bench.py load it; the product library never links it.
//
// Each piece cites the reference file:line it follows (paths relative to the
// FMCA repository root):
//   tree build      FMCA/src/Clustering/ClusterTreeInitializer_ClusterTree.h:24-132
//   splitter        FMCA/src/Clustering/ClusterTreeSplitter.h:60-81
//   level order     FMCA/src/util/IDDFSForwardIterator.h:37-65
//   multi-indices   FMCA/src/util/MultiIndexSet.h:87-111,147-187
//   interpolation   FMCA/src/Interpolator/TensorProductInterpolator.h:39-81,
//                   FMCA/src/Interpolator/TotalDegreeInterpolator.h:27-48,
//                   FMCA/src/Interpolator/DeMarchiPoints.h:22-50,
//                   FMCA/src/Interpolator/evalPolynomials.h:19-66,
//                   FMCA/src/util/HaltonSet.h:23-131
//   leaf moments    FMCA/src/Moments/NystromMoments.h:24-34
//   transfers       FMCA/src/H2Matrix/compute_cluster_bases_impl.h:28-57
//   nestedness KAT  FMCA/src/H2Matrix/compute_cluster_bases_impl.h:59-80
//   admissibility   FMCA/src/Clustering/Admissibility.h:19-29,95-115
//   block pattern   FMCA/src/H2Matrix/H2MatrixBase.h:160-191
//   action          FMCA/src/H2Matrix/H2kernelActor.h:38-80 and
//                   FMCA/src/H2Matrix/H2MatrixBase.h:113-152
//   passes          FMCA/src/H2Matrix/forward_transform_impl.h:21-35,
//                   FMCA/src/H2Matrix/backward_transform_impl.h:21-33
//   evaluator       FMCA/src/MatrixEvaluator/unsymmetricNystromEvaluator.h:42-73
//   kernels         FMCA/src/CovarianceKernel/CovarianceKernel.h:50-143
//   CG              FMCA/src/util/PCG.h:15-44
//
// Deliberate differences from FMCA, none of which changes a result:
//   * tensor-scheme coupling blocks skip FMCA's products with invV = I
//     (unsymmetricNystromEvaluator.h:56); multiplying by an identity matrix is
//     exact in IEEE arithmetic, so S is unchanged;
//   * the DeMarchi points come from this file's own column-pivoted Householder
//     QR and invV from Gauss-Jordan elimination instead of Eigen's
//     ColPivHouseholderQR and inverse(): the pivot order is the same unless
//     two candidate norms tie to rounding, and invV agrees to rounding;
//   * kernel entries are evaluated through a switch instead of two
//     std::function calls with heap temporaries (CovarianceKernel.h:127-143).
// Both make this CPU action faster than FMCA's, i.e. a conservative baseline.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

namespace orc {

using Index = uint32_t;  // FMCA_INDEX, Macros.h:53
constexpr double kEps = std::numeric_limits<double>::epsilon();  // Macros.h:56-57
constexpr double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// Kernels (CovarianceKernel.h:50-121).  Parameters are held by value.
enum KType {
  K_EXPONENTIAL = 0,
  K_MATERN32 = 1,
  K_MATERN52 = 2,
  K_MATERN72 = 3,
  K_MATERN92 = 4,
  K_GAUSSIAN = 5,
  K_INVMULTIQUADRIC = 6,
  K_MULTIQUADRIC = 7,
  K_BIHARMONIC2D = 8,
  K_BIHARMONIC3D = 9,
  K_TRIHARMONIC3D = 10,
  K_TPS2D = 11,
};

int parse_kernel(const char *name) {
  std::string s(name);
  for (auto &c : s) c = (char)toupper(c);
  static const char *names[] = {"EXPONENTIAL",     "MATERN32",     "MATERN52",
                                "MATERN72",        "MATERN92",     "GAUSSIAN",
                                "INVMULTIQUADRIC", "MULTIQUADRIC", "BIHARMONIC2D",
                                "BIHARMONIC3D",    "TRIHARMONIC3D", "TPS2D"};
  for (int i = 0; i < 12; ++i)
    if (s == names[i]) return i;
  return -1;
}

// Bit of the kernel id that selects the geodesic distance.
constexpr int kGeodesicFlag = 0x100;

struct Kernel {
  int type;
  double l, c, nu;
  bool geodesic = false;
  Kernel() = default;
  // id: kernel type, or'ed with kGeodesicFlag for setDistanceType("GEODESIC")
  Kernel(int id, double l_, double c_, double nu_)
      : type(id & ~kGeodesicFlag), l(l_), c(c_), nu(nu_), geodesic((id & kGeodesicFlag) != 0) {}
  double operator()(double r) const {
    switch (type) {
      case K_BIHARMONIC2D:
        return r < kEps ? 0 : std::log(r / l) * (r / l) * (r / l);
      case K_BIHARMONIC3D:
        return r / l;
      case K_TRIHARMONIC3D:
        return (r / l) * (r / l) * (r / l);
      case K_EXPONENTIAL:
        return std::exp(-r / l);
      case K_MATERN32:
        return (1. + std::sqrt(3.) / l * r) * std::exp(-std::sqrt(3.) * r / l);
      case K_MATERN52:
        return (1. + std::sqrt(5.) / l * r + 5 * r * r / (3 * l * l)) *
               std::exp(-std::sqrt(5.) * r / l);
      case K_MATERN72:
        return (1. + std::sqrt(7.) / l * r + 14. / 5 * std::pow(r / l, 2.) +
                7. * std::sqrt(7.) / 15. * std::pow(r / l, 3.)) *
               std::exp(-std::sqrt(7.) * r / l);
      case K_MATERN92:
        return (1. + 3. / l * r + 27. / 7. * std::pow(r / l, 2.) +
                18. / 7. * std::pow(r / l, 3.) + 27. / 35. * std::pow(r / l, 4.)) *
               std::exp(-3 * r / l);
      case K_GAUSSIAN:
        return std::exp(-0.5 * r * r / (l * l));
      case K_INVMULTIQUADRIC:
        return 1. / std::sqrt((r / l) * (r / l) + c * c);
      case K_MULTIQUADRIC:
        return std::sqrt((r / l) * (r / l) + c * c);
      case K_TPS2D:
        return r > 0 ? r * r * std::log(r) : 0;
    }
    return std::numeric_limits<double>::quiet_NaN();
  }
  // Euclidean distance (CovarianceKernel.h:126-129): sqrt of the sum of
  // squares.  Geodesic distance (CovarianceKernel.h:130-133 with
  // SphereClusterTree.h:65-69): acos of the dot product clamped to [-1, 1];
  // the dot product is x[0] y[0] followed by one fma per further axis (Eigen
  // leaves the order of its reduction open; this one is fixed so that a
  // second implementation can round alike where acos amplifies rounding).
  double operator()(const double *x, const double *y, int d) const {
    if (geodesic) {
      double dot = x[0] * y[0];
      for (int i = 1; i < d; ++i) dot = std::fma(x[i], y[i], dot);
      return (*this)(std::acos(std::min(1., std::max(-1., dot))));
    }
    double s = 0;
    for (int i = 0; i < d; ++i) s += (x[i] - y[i]) * (x[i] - y[i]);
    return (*this)(std::sqrt(s));
  }
};

// ---------------------------------------------------------------------------
// Pieces of the total-degree interpolator (TotalDegreeInterpolator.h:27-48).

// HaltonSet<S> (util/HaltonSet.h:23-131): the first M primes as bases, b-adic
// digits incremented with carry, radical inverse by Horner in long double.
// The constructor advances `skip` times (init_HaltonVector); `v` is the
// current point.
struct Halton {
  static constexpr int kDigits = 40;  // HaltonSet = HaltonSetBase<S, 40>
  std::vector<uint32_t> primes;
  std::vector<std::vector<uint32_t>> digits;
  std::vector<int> top;  // highest digit position reached per base
  std::vector<double> v;

  Halton(int M, int skip) {
    primes.resize(M);
    int found = 0;
    if (M > 0) primes[found++] = 2;
    for (uint32_t cand = 3; found < M; cand += 2) {
      bool prime = true;
      const uint32_t lim = (uint32_t)std::sqrt((double)cand);
      for (int i = 0; i < found && primes[i] <= lim; ++i)
        if (cand % primes[i] == 0) {
          prime = false;
          break;
        }
      if (prime) primes[found++] = cand;
    }
    digits.assign(M, std::vector<uint32_t>(kDigits, 0));
    top.assign(M, 0);
    v.assign(M, 0.);
    for (int s = 0; s < skip; ++s) next();
  }

  void next() {
    for (size_t i = 0; i < primes.size(); ++i) {
      int k = 0;
      for (; k < kDigits; ++k) {
        if (digits[i][k] + 1 < primes[i]) {
          ++digits[i][k];
          break;
        }
        digits[i][k] = 0;
      }
      top[i] = std::max(top[i], k);
      const long double binv = (long double)1 / (long double)primes[i];
      long double r = digits[i][top[i]];
      for (int j = top[i] - 1; j >= 0; --j) r = binv * r + digits[i][j];
      r *= binv;
      v[i] = (double)r;
    }
  }
};

// LegendrePolynomials::eval (evalPolynomials.h:19-40): shifted Legendre
// polynomials on [0,1] by the three-term recurrence, each scaled by
// sqrt(2k+1) after the recurrence has used it.
static void legendre_values(double x, int deg, double *out) {
  double p0 = 0, p1 = 1;
  out[0] = p1;
  for (int k = 1; k <= deg; ++k) {
    const double pk = double(2 * k - 1) / double(k) * (2 * x - 1) * p1 -
                      double(k - 1) / double(k) * p0;
    p0 = p1;
    p1 = pk;
    out[k] = pk * std::sqrt(double(2 * k + 1));
  }
}

// Column order of Householder QR with column pivoting, as Eigen's
// ColPivHouseholderQR computes it (DeMarchiPoints.h:38-40 uses
// colsPermutation): at step k the remaining column of largest (downdated)
// norm, the first one on ties; norms downdated after each reflection and
// recomputed when the downdate loses more than half the digits (the xGEQP3
// safeguard).  A is rows x cols, column-major, and is overwritten.
static std::vector<int> pivoted_qr_order(std::vector<double> &A, int rows, int cols) {
  auto col = [&](int j) { return A.data() + (size_t)rows * j; };
  std::vector<double> nrm_direct(cols), nrm(cols);
  std::vector<int> order(cols);
  std::iota(order.begin(), order.end(), 0);
  for (int j = 0; j < cols; ++j) {
    double s = 0;
    for (int i = 0; i < rows; ++i) s += col(j)[i] * col(j)[i];
    nrm_direct[j] = nrm[j] = std::sqrt(s);
  }
  const double downdate_tol = std::sqrt(kEps);
  const int steps = std::min(rows, cols);
  for (int k = 0; k < steps; ++k) {
    int best = k;
    for (int j = k + 1; j < cols; ++j)
      if (nrm[j] > nrm[best]) best = j;
    if (best != k) {
      std::swap_ranges(col(k), col(k) + rows, col(best));
      std::swap(nrm[k], nrm[best]);
      std::swap(nrm_direct[k], nrm_direct[best]);
      std::swap(order[k], order[best]);
    }
    // reflector I - tau v v^T with v = [1; ess] mapping A(k:, k) to beta e_1
    double *a = col(k);
    double tail = 0;
    for (int i = k + 1; i < rows; ++i) tail += a[i] * a[i];
    double tau = 0, beta = a[k];
    if (tail > std::numeric_limits<double>::min()) {
      beta = std::sqrt(a[k] * a[k] + tail);
      if (a[k] >= 0) beta = -beta;
      const double den = a[k] - beta;
      for (int i = k + 1; i < rows; ++i) a[i] /= den;
      tau = (beta - a[k]) / beta;
    } else {
      for (int i = k + 1; i < rows; ++i) a[i] = 0;
    }
    a[k] = beta;
    for (int j = k + 1; j < cols; ++j) {
      double *b = col(j);
      if (tau != 0) {
        double wdot = b[k];
        for (int i = k + 1; i < rows; ++i) wdot += a[i] * b[i];
        b[k] -= tau * wdot;
        for (int i = k + 1; i < rows; ++i) b[i] -= tau * a[i] * wdot;
      }
      if (nrm[j] != 0) {
        double t = std::abs(b[k]) / nrm[j];
        t = std::max(0., (1 + t) * (1 - t));
        const double ratio = nrm[j] / nrm_direct[j];
        if (t * (ratio * ratio) <= downdate_tol) {
          double s = 0;
          for (int i = k + 1; i < rows; ++i) s += b[i] * b[i];
          nrm_direct[j] = nrm[j] = std::sqrt(s);
        } else {
          nrm[j] *= std::sqrt(t);
        }
      }
    }
  }
  return order;
}

// Inverse of an n x n column-major matrix by Gauss-Jordan elimination with
// partial pivoting (Eigen's inverse() of a dynamic matrix is an LU solve).
static std::vector<double> invert(std::vector<double> M, int n) {
  std::vector<double> R((size_t)n * n, 0.);
  for (int i = 0; i < n; ++i) R[i + (size_t)n * i] = 1;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int i = c + 1; i < n; ++i)
      if (std::abs(M[i + (size_t)n * c]) > std::abs(M[piv + (size_t)n * c])) piv = i;
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        std::swap(M[c + (size_t)n * j], M[piv + (size_t)n * j]);
        std::swap(R[c + (size_t)n * j], R[piv + (size_t)n * j]);
      }
    const double d = M[c + (size_t)n * c];
    for (int j = 0; j < n; ++j) {
      M[c + (size_t)n * j] /= d;
      R[c + (size_t)n * j] /= d;
    }
    for (int i = 0; i < n; ++i) {
      if (i == c) continue;
      const double f = M[i + (size_t)n * c];
      if (f == 0) continue;
      for (int j = 0; j < n; ++j) {
        M[i + (size_t)n * j] -= f * M[c + (size_t)n * j];
        R[i + (size_t)n * j] -= f * R[c + (size_t)n * j];
      }
    }
  }
  return R;
}

// ---------------------------------------------------------------------------
// Interpolation on [0,1]^d with the graded-lexicographic multi-index order of
// MultiIndexSet.h:87-100:
//   kTensor       tensor-product Chebyshev nodes, barycentric Lagrange values
//                 (TensorProductInterpolator.h:39-81), invV = I;
//   kTotalDegree  total-degree Legendre products at DeMarchi points
//                 (TotalDegreeInterpolator.h:27-48, DeMarchiPoints.h:22-50,
//                 evalPolynomials.h:42-66), dense invV = V^-1.
enum Scheme : int { kTensor = 0, kTotalDegree = 1 };

struct Interp {
  int dim = 0, deg = 0, q = 0;
  int scheme = kTensor;
  std::vector<double> xi, w;           // deg+1 nodes / barycentric weights
  std::vector<std::vector<int>> mi;    // q multi-indices, FMCA order
  std::vector<double> Xi;              // dim x q, column-major
  std::vector<double> V, invV;         // q x q, column-major (total degree only)

  void init(int d, int p, int sch = kTensor) {
    dim = d;
    deg = p;
    scheme = sch;
    if (sch == kTotalDegree) {
      init_total_degree();
      return;
    }
    const double weight = 2. * kPi / (2. * double(p) + 2.);
    xi.resize(p + 1);
    w.resize(p + 1);
    for (int k = 0; k <= p; ++k) {
      xi[k] = std::cos(weight * (double(k) + 0.5));
      xi[k] = 0.5 * (xi[k] + 1);
      w[k] = std::sin(weight * (double(k) + 0.5));
      if (k % 2 == 1) w[k] *= -1.;
    }
    // all multi-indices with max entry <= deg, ordered by (l1 norm, lexicographic)
    mi.clear();
    std::vector<int> a(d, 0);
    while (true) {
      mi.push_back(a);
      int i = 0;
      while (i < d && a[i] == p) a[i++] = 0;
      if (i == d) break;
      ++a[i];
    }
    std::sort(mi.begin(), mi.end(), [](const std::vector<int> &x, const std::vector<int> &y) {
      int nx = 0, ny = 0;
      for (int v : x) nx += v;
      for (int v : y) ny += v;
      if (nx != ny) return nx < ny;
      return std::lexicographical_compare(x.begin(), x.end(), y.begin(), y.end());
    });
    q = (int)mi.size();
    Xi.assign((size_t)d * q, 0.);
    for (int k = 0; k < q; ++k)
      for (int i = 0; i < d; ++i) Xi[i + (size_t)d * k] = xi[mi[k][i]];
  }

  // MultiIndexSet<TotalDegree> (MultiIndexSet.h:39-50,147-187), DeMarchi
  // points from 10 q Halton points after skipping 100 (DeMarchiPoints.h:22-40,
  // TotalDegreeInterpolator.h:41-46) and V(i, :) = P(Xi(:, i))^T, invV = V^-1.
  void init_total_degree() {
    mi.clear();
    std::vector<int> a(dim, 0);
    while (true) {
      int sum = 0;
      for (int v : a) sum += v;
      if (sum <= deg) mi.push_back(a);
      int i = 0;
      while (i < dim && a[i] == deg) a[i++] = 0;
      if (i == dim) break;
      ++a[i];
    }
    std::sort(mi.begin(), mi.end(), [](const std::vector<int> &x, const std::vector<int> &y) {
      int nx = 0, ny = 0;
      for (int v : x) nx += v;
      for (int v : y) ny += v;
      if (nx != ny) return nx < ny;
      return std::lexicographical_compare(x.begin(), x.end(), y.begin(), y.end());
    });
    q = (int)mi.size();
    const int ncand = 10 * q;
    std::vector<double> cand((size_t)dim * ncand), VT((size_t)q * ncand);
    Halton hs(dim, 100);
    for (int j = 0; j < ncand; ++j) {
      for (int i = 0; i < dim; ++i) cand[i + (size_t)dim * j] = hs.v[i];
      hs.next();
    }
    for (int j = 0; j < ncand; ++j) eval(cand.data() + (size_t)dim * j, VT.data() + (size_t)q * j);
    const std::vector<int> order = pivoted_qr_order(VT, q, ncand);
    Xi.assign((size_t)dim * q, 0.);
    for (int k = 0; k < q; ++k)
      for (int i = 0; i < dim; ++i) Xi[i + (size_t)dim * k] = cand[i + (size_t)dim * order[k]];
    V.assign((size_t)q * q, 0.);
    std::vector<double> row(q);
    for (int k = 0; k < q; ++k) {
      eval(Xi.data() + (size_t)dim * k, row.data());
      for (int a2 = 0; a2 < q; ++a2) V[k + (size_t)q * a2] = row[a2];
    }
    invV = invert(V, q);
  }

  // Values of all q basis polynomials at a point of [0,1]^d: barycentric
  // Lagrange products (tensor) or Legendre products (total degree,
  // evalPolynomials.h:53-64).
  void eval(const double *pt, double *out) const {
    if (scheme == kTotalDegree) {
      std::vector<double> pv((size_t)dim * (deg + 1));
      for (int i = 0; i < dim; ++i) legendre_values(pt[i], deg, pv.data() + (size_t)(deg + 1) * i);
      for (int k = 0; k < q; ++k) {
        double v = 1.;
        for (int i = 0; i < dim; ++i) v *= pv[(size_t)(deg + 1) * i + mi[k][i]];
        out[k] = v;
      }
      return;
    }
    std::vector<double> weight(dim);
    for (int i = 0; i < dim; ++i) {
      double s = 0;
      for (int j = 0; j <= deg; ++j) s += w[j] / (pt[i] - xi[j]);
      weight[i] = s;
    }
    for (int k = 0; k < q; ++k) {
      double v = 1.;
      for (int i = 0; i < dim; ++i)
        if (std::abs(pt[i] - xi[mi[k][i]]) > kEps)
          v *= w[mi[k][i]] / (pt[i] - xi[mi[k][i]]) / weight[i];
      out[k] = v;
    }
  }
};

// ---------------------------------------------------------------------------
// Cluster tree (ClusterTreeInitializer_ClusterTree.h, ClusterTreeSplitter.h).
struct Node {
  std::vector<double> bb;  // dim x 3 column-major: [min | max | size]
  Index begin = 0, size = 0;
  int son[2] = {-1, -1};
  int dad = -1;
  int level = 0;
  int block_id = -1;
  // H2 data (compute_cluster_bases_impl.h:28-57)
  std::vector<double> V;     // leaves: q x size, column-major
  std::vector<double> E[2];  // inner: q x q per son, column-major
};

struct Tree {
  int dim = 0;
  const double *P = nullptr;  // dim x n, column-major (borrowed during build)
  std::vector<Index> idx;
  std::vector<Node> nodes;     // storage order = creation order
  std::vector<int> by_id;      // block_id -> storage position
  int nclusters() const { return (int)nodes.size(); }
  double &bb(int node, int i, int col) { return nodes[node].bb[i + dim * col]; }
  double bb(int node, int i, int col) const { return nodes[node].bb[i + dim * col]; }
  bool is_leaf(int node) const { return nodes[node].son[0] < 0; }
};

// init_BoundingBox_impl (:44-56)
static void root_box(Tree &T, Index n) {
  const int d = T.dim;
  Node &r = T.nodes[0];
  r.bb.assign(3 * d, 0.);
  for (int i = 0; i < d; ++i) {
    double mn = T.P[i], mx = T.P[i];
    for (Index j = 1; j < n; ++j) {
      mn = std::min(mn, T.P[i + (size_t)d * j]);
      mx = std::max(mx, T.P[i + (size_t)d * j]);
    }
    mn -= kEps * std::abs(mn);
    mx += kEps * std::abs(mx);
    r.bb[i] = mn;
    r.bb[i + d] = mx;
    r.bb[i + 2 * d] = mx - mn;
  }
}

// CardinalityBisection (ClusterTreeSplitter.h:60-81) applied to node `t`,
// whose sons `a`, `b` start as copies of t.
static void cardinality_split(Tree &T, int a, int b) {
  const int d = T.dim;
  Node &c1 = T.nodes[a];
  Node &c2 = T.nodes[b];
  int longest = 0;  // first maximal edge, as Eigen's maxCoeff
  for (int i = 1; i < d; ++i)
    if (c1.bb[i + 2 * d] > c1.bb[longest + 2 * d]) longest = i;
  const double *P = T.P;
  Index *id = T.idx.data();
  std::sort(id + c1.begin, id + c1.begin + c1.size, [P, d, longest](Index i, Index j) {
    return P[longest + (size_t)d * i] < P[longest + (size_t)d * j];
  });
  c1.size /= 2;
  c2.size -= c1.size;
  c2.begin += c1.size;
  c1.bb[longest + d] = P[longest + (size_t)d * id[c1.begin + c1.size - 1]];
  c1.bb[longest + 2 * d] = c1.bb[longest + d] - c1.bb[longest];
  c2.bb[longest] = P[longest + (size_t)d * id[c2.begin]];
  c2.bb[longest + 2 * d] = c2.bb[longest + d] - c2.bb[longest];
}

// init_ClusterTree_impl (:61-82)
static void refine(Tree &T, int t, Index min_csize) {
  const Index split_threshold = min_csize >= 1 ? (2 * min_csize - 1) : 1;
  if (T.nodes[t].size > split_threshold) {
    for (int s = 0; s < 2; ++s) {
      Node son;
      son.bb = T.nodes[t].bb;
      son.begin = T.nodes[t].begin;
      son.size = T.nodes[t].size;
      son.dad = t;
      son.level = T.nodes[t].level + 1;
      T.nodes.push_back(son);
      T.nodes[t].son[s] = (int)T.nodes.size() - 1;
    }
    cardinality_split(T, T.nodes[t].son[0], T.nodes[t].son[1]);
    for (int s = 0; s < 2; ++s) refine(T, T.nodes[t].son[s], min_csize);
  }
}

// shrinkToFit_impl (:88-132)
static void shrink(Tree &T, int t) {
  const int d = T.dim;
  std::vector<double> m(3 * d);
  Node &nd = T.nodes[t];
  if (nd.size) {
    const double *p0 = T.P + (size_t)d * T.idx[nd.begin];
    for (int i = 0; i < d; ++i) m[i] = m[i + d] = p0[i];
    if (nd.son[0] >= 0) {
      for (int s = 0; s < 2; ++s) {
        const int c = T.nodes[t].son[s];
        shrink(T, c);
        if (T.nodes[c].size)
          for (int i = 0; i < d; ++i) {
            m[i] = std::min(m[i], T.nodes[c].bb[i]);
            m[i + d] = std::max(m[i + d], T.nodes[c].bb[i + d]);
          }
      }
    } else {
      for (Index j = 0; j < nd.size; ++j) {
        const double *p = T.P + (size_t)d * T.idx[nd.begin + j];
        for (int i = 0; i < d; ++i) {
          m[i] = std::min(p[i], m[i]);
          m[i + d] = std::max(p[i], m[i + d]);
        }
      }
    }
  } else {
    for (int i = 0; i < d; ++i) {
      m[i] = std::numeric_limits<double>::infinity();
      m[i + d] = -std::numeric_limits<double>::infinity();
    }
  }
  for (int i = 0; i < d; ++i) m[i + 2 * d] = m[i + d] - m[i];
  for (int i = 0; i < d; ++i)
    if (m[i + d] - m[i] < 2e2 * kEps) {
      m[i + d] += 1e2 * kEps;
      m[i] -= 1e2 * kEps;
      m[i + 2 * d] = m[i + d] - m[i];
    }
  T.nodes[t].bb = m;
}

// Level-order numbering: IDDFSForwardIterator visits depth 0, 1, 2, ... and
// each depth left to right, which is a breadth-first order.
static void number_level_order(Tree &T) {
  T.by_id.clear();
  std::vector<int> frontier{0};
  int id = 0;
  while (!frontier.empty()) {
    std::vector<int> next;
    for (int t : frontier) {
      T.nodes[t].block_id = id++;
      T.by_id.push_back(t);
      for (int s = 0; s < 2; ++s)
        if (T.nodes[t].son[s] >= 0) next.push_back(T.nodes[t].son[s]);
    }
    frontier.swap(next);
  }
}

static void build_tree(Tree &T, int d, Index n, const double *P, Index min_csize) {
  T.dim = d;
  T.P = P;
  T.nodes.assign(1, Node());
  root_box(T, n);
  T.nodes[0].begin = 0;
  T.nodes[0].size = n;
  T.idx.resize(n);
  std::iota(T.idx.begin(), T.idx.end(), 0u);
  refine(T, 0, min_csize);
  shrink(T, 0);
  number_level_order(T);
}

// compute_cluster_bases_impl::compute (:28-57) with NystromMoments (:24-34)
static void compute_bases(Tree &T, int t, const Interp &I) {
  const int d = T.dim;
  const int q = I.q;
  Node &nd = T.nodes[t];
  if (nd.son[0] >= 0) {
    for (int s = 0; s < 2; ++s) compute_bases(T, T.nodes[t].son[s], I);
    std::vector<double> pt(d);
    for (int s = 0; s < 2; ++s) {
      const int c = T.nodes[t].son[s];
      std::vector<double> &E = T.nodes[t].E[s];
      E.assign((size_t)q * q, 0.);
      if (T.nodes[c].size)
        for (int j = 0; j < q; ++j) {
          for (int i = 0; i < d; ++i)
            pt[i] = I.Xi[i + (size_t)d * j] * T.bb(c, i, 2) / T.bb(t, i, 2) +
                    (T.bb(c, i, 0) - T.bb(t, i, 0)) / T.bb(t, i, 2);
          I.eval(pt.data(), E.data() + (size_t)q * j);
        }
      if (I.scheme == kTotalDegree) {
        // E = E * invV^T (compute_cluster_bases_impl.h:50)
        std::vector<double> P = E;
        for (int b2 = 0; b2 < q; ++b2)
          for (int a2 = 0; a2 < q; ++a2) {
            double acc = 0;
            for (int j = 0; j < q; ++j) acc += P[a2 + (size_t)q * j] * I.invV[b2 + (size_t)q * j];
            E[a2 + (size_t)q * b2] = acc;
          }
      }
    }
  } else {
    nd.V.assign((size_t)q * nd.size, 0.);
    std::vector<double> pt(d);
    for (Index j = 0; j < nd.size; ++j) {
      const double *p = T.P + (size_t)d * T.idx[nd.begin + j];
      for (int i = 0; i < d; ++i) pt[i] = (p[i] - T.bb(t, i, 0)) / T.bb(t, i, 2);
      I.eval(pt.data(), nd.V.data() + (size_t)q * j);
    }
  }
}

// ---------------------------------------------------------------------------
// Block pattern (H2MatrixBase.h:160-191, Admissibility.h:19-29,95-115)
struct Block {
  int row, col;       // storage positions in the row / column tree
  bool low_rank;
};

static double radius(const Tree &T, int t) {
  double s = 0;
  for (int i = 0; i < T.dim; ++i) s += T.bb(t, i, 2) * T.bb(t, i, 2);
  return 0.5 * std::sqrt(s);
}

static double cluster_distance(const Tree &R, int t, const Tree &C, int s) {
  double s2 = 0;
  for (int i = 0; i < R.dim; ++i) {
    const double c =
        0.5 * (R.bb(t, i, 0) - C.bb(s, i, 0) + R.bb(t, i, 1) - C.bb(s, i, 1));
    s2 += c * c;
  }
  const double dist = std::sqrt(s2) - radius(R, t) - radius(C, s);
  return dist > 0 ? dist : 0;
}

enum Adm { Refine = 0, LowRank = 1, Dense = 2 };

static Adm compare(const Tree &R, int t, const Tree &C, int s, double eta) {
  const double dist = cluster_distance(R, t, C, s);
  const double rr = radius(R, t), rc = radius(C, s);
  const double rad = rr > rc ? rr : rc;
  if (rad > eta * dist) {
    if (R.is_leaf(t) || C.is_leaf(s)) return Dense;
    return Refine;
  }
  return LowRank;
}

struct BTNode {
  int row, col;
  bool low_rank = false;
  std::vector<int> sons;
};

// Returns the leaf blocks in the level order of the block tree.
static std::vector<Block> compute_pattern(const Tree &R, const Tree &C, double eta, int q) {
  std::vector<BTNode> bt;
  bt.push_back(BTNode{0, 0});
  std::vector<int> stack{0};
  while (!stack.empty()) {
    const int b = stack.back();
    stack.pop_back();
    const int t = bt[b].row, s = bt[b].col;
    const Adm adm = compare(R, t, C, s, eta);
    if (adm == LowRank) {
      // FMCA multiplies unsigned 32-bit indices here; the wrap is kept.
      const Index lhs = Index(q) * Index(q);
      const Index rhs = R.nodes[t].size * C.nodes[s].size;
      bt[b].low_rank = lhs < rhs;
    } else if (adm == Refine) {
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          BTNode son{R.nodes[t].son[i], C.nodes[s].son[j]};
          bt.push_back(son);
          const int sid = (int)bt.size() - 1;
          bt[b].sons.push_back(sid);
          stack.push_back(sid);
        }
    }
  }
  std::vector<Block> out;
  std::vector<int> frontier{0};
  while (!frontier.empty()) {
    std::vector<int> next;
    for (int b : frontier) {
      if (bt[b].sons.empty()) out.push_back(Block{bt[b].row, bt[b].col, bt[b].low_rank});
      for (int s : bt[b].sons) next.push_back(s);
    }
    frontier.swap(next);
  }
  return out;
}

// ---------------------------------------------------------------------------
struct Actor {
  int dim, deg, q;
  Index n, m;
  std::vector<double> P, Peval;  // owned copies (FMCA keeps references)
  Kernel K;
  double eta;
  Interp I;
  Tree cols, rows;               // source tree (P) and target tree (Peval)
  std::vector<Block> blocks;
  std::vector<std::vector<int>> scheduler;  // column block_id -> block ids
  double setup_seconds = 0;
};

static void forward_rec(const Actor &A, int t, const double *srhs, Index nrhs,
                        std::vector<std::vector<double>> &trhs) {
  const Tree &T = A.cols;
  const Node &nd = T.nodes[t];
  const int q = A.q;
  std::vector<double> &out = trhs[nd.block_id];
  out.assign((size_t)q * nrhs, 0.);
  if (nd.son[0] >= 0) {
    for (int s = 0; s < 2; ++s) {
      const int c = nd.son[s];
      forward_rec(A, c, srhs, nrhs, trhs);
      const std::vector<double> &E = nd.E[s];
      const std::vector<double> &x = trhs[T.nodes[c].block_id];
      for (Index r = 0; r < nrhs; ++r)
        for (int k = 0; k < q; ++k) {
          const double xv = x[k + (size_t)q * r];
          for (int i = 0; i < q; ++i) out[i + (size_t)q * r] += E[i + (size_t)q * k] * xv;
        }
    }
  } else {
    for (Index r = 0; r < nrhs; ++r)
      for (Index j = 0; j < nd.size; ++j) {
        const double xv = srhs[nd.begin + j + (size_t)A.n * r];
        for (int i = 0; i < q; ++i) out[i + (size_t)q * r] += nd.V[i + (size_t)q * j] * xv;
      }
  }
}

static void backward_rec(const Actor &A, int t, double *lhs, Index nrhs,
                         std::vector<std::vector<double>> &tlhs) {
  const Tree &T = A.rows;
  const Node &nd = T.nodes[t];
  const int q = A.q;
  const std::vector<double> &in = tlhs[nd.block_id];
  if (nd.son[0] >= 0) {
    for (int s = 0; s < 2; ++s) {
      const int c = nd.son[s];
      std::vector<double> &out = tlhs[T.nodes[c].block_id];
      const std::vector<double> &E = nd.E[s];
      for (Index r = 0; r < nrhs; ++r)
        for (int k = 0; k < q; ++k) {
          double acc = 0;
          for (int i = 0; i < q; ++i) acc += E[i + (size_t)q * k] * in[i + (size_t)q * r];
          out[k + (size_t)q * r] += acc;
        }
      backward_rec(A, c, lhs, nrhs, tlhs);
    }
  } else {
    for (Index r = 0; r < nrhs; ++r)
      for (Index j = 0; j < nd.size; ++j) {
        double acc = 0;
        for (int i = 0; i < q; ++i) acc += nd.V[i + (size_t)q * j] * in[i + (size_t)q * r];
        lhs[nd.begin + j + (size_t)A.m * r] += acc;
      }
  }
}

// interpolate_kernel (unsymmetricNystromEvaluator.h:42-58); the products with
// invV are skipped for the tensor scheme, where invV = I
static void coupling_matrix(const Actor &A, int t, int s, std::vector<double> &S) {
  const int d = A.dim, q = A.q;
  std::vector<double> X((size_t)d * q), Y((size_t)d * q);
  for (int k = 0; k < q; ++k)
    for (int i = 0; i < d; ++i) {
      X[i + (size_t)d * k] = A.I.Xi[i + (size_t)d * k] * A.rows.bb(t, i, 2) + A.rows.bb(t, i, 0);
      Y[i + (size_t)d * k] = A.I.Xi[i + (size_t)d * k] * A.cols.bb(s, i, 2) + A.cols.bb(s, i, 0);
    }
  S.resize((size_t)q * q);
  for (int j = 0; j < q; ++j)
    for (int i = 0; i < q; ++i)
      S[i + (size_t)q * j] = A.K(X.data() + (size_t)d * i, Y.data() + (size_t)d * j, d);
  if (A.I.scheme == kTotalDegree) {
    // S = (invV * S) * invV^T (unsymmetricNystromEvaluator.h:56)
    const std::vector<double> &iv = A.I.invV;
    std::vector<double> L((size_t)q * q, 0.);
    for (int j = 0; j < q; ++j)
      for (int k = 0; k < q; ++k) {
        const double sv = S[k + (size_t)q * j];
        for (int i = 0; i < q; ++i) L[i + (size_t)q * j] += iv[i + (size_t)q * k] * sv;
      }
    for (int j = 0; j < q; ++j)
      for (int i = 0; i < q; ++i) {
        double acc = 0;
        for (int k = 0; k < q; ++k) acc += L[i + (size_t)q * k] * iv[j + (size_t)q * k];
        S[i + (size_t)q * j] = acc;
      }
  }
}

// One leaf block of the hot loop (H2kernelActor.h:59-70); srhs/lhs in cluster order.
static void apply_block(const Actor &A, const Block &b, const double *srhs, double *lhs, Index nrhs,
                        const std::vector<std::vector<double>> &trhs,
                        std::vector<std::vector<double>> &tlhs, std::vector<double> &S) {
  const int q = A.q, d = A.dim;
  const Node &rt = A.rows.nodes[b.row];
  const Node &ct = A.cols.nodes[b.col];
  if (b.low_rank) {
    coupling_matrix(A, b.row, b.col, S);
    const std::vector<double> &x = trhs[ct.block_id];
    std::vector<double> &y = tlhs[rt.block_id];
    for (Index r = 0; r < nrhs; ++r)
      for (int j = 0; j < q; ++j) {
        const double xv = x[j + (size_t)q * r];
        for (int i = 0; i < q; ++i) y[i + (size_t)q * r] += S[i + (size_t)q * j] * xv;
      }
  } else {
    // compute_dense_block (:63-73) fused with lhs.middleRows += S * srhs.middleRows
    const Index br = rt.size, bc = ct.size;
    S.resize((size_t)br * bc);
    for (Index j = 0; j < bc; ++j) {
      const double *y = A.P.data() + (size_t)d * A.cols.idx[ct.begin + j];
      for (Index i = 0; i < br; ++i) {
        const double *x = A.Peval.data() + (size_t)d * A.rows.idx[rt.begin + i];
        S[i + (size_t)br * j] = A.K(x, y, d);
      }
    }
    for (Index r = 0; r < nrhs; ++r)
      for (Index j = 0; j < bc; ++j) {
        const double xv = srhs[ct.begin + j + (size_t)A.n * r];
        double *out = lhs + rt.begin + (size_t)A.m * r;
        for (Index i = 0; i < br; ++i) out[i] += S[i + (size_t)br * j] * xv;
      }
  }
}

// Cluster-order action (H2MatrixBase::action, H2MatrixBase.h:113-152).
static void action_cluster_order(const Actor &A, const double *srhs, Index nrhs, double *lhs) {
  std::fill(lhs, lhs + (size_t)A.m * nrhs, 0.);
  std::vector<std::vector<double>> trhs(A.cols.nclusters());
  forward_rec(A, 0, srhs, nrhs, trhs);
  std::vector<std::vector<double>> tlhs(A.rows.nclusters());
  for (auto &v : tlhs) v.assign((size_t)A.q * nrhs, 0.);
  for (const auto &group : A.scheduler) {
#pragma omp parallel
    {
      std::vector<double> S;
#pragma omp for schedule(dynamic)
      for (long k = 0; k < (long)group.size(); ++k)
        apply_block(A, A.blocks[group[k]], srhs, lhs, nrhs, trhs, tlhs, S);
    }
  }
  backward_rec(A, 0, lhs, nrhs, tlhs);
}

}  // namespace orc

// ===========================================================================
// C interface for the tests (ctypes).
using namespace orc;

extern "C" {

int orc_parse_kernel(const char *name) { return parse_kernel(name); }
int orc_geodesic_flag() { return kGeodesicFlag; }

// Kernel values k(r) for an array of distances.
void orc_kernel_eval(int type, double l, double c, double nu, const double *r, long n,
                     double *out) {
  Kernel K{type, l, c, nu};
  for (long i = 0; i < n; ++i) out[i] = K(r[i]);
}

// Interpolation data of a scheme (0 tensor, 1 total degree): xi, w (deg+1
// each, tensor only), multi-indices (q x dim, FMCA order), Xi (dim x q), and
// for total degree V, invV (q x q, column-major).  Returns q.
int orc_interp(int dim, int deg, int scheme, double *xi, double *w, int *mi, double *Xi,
               double *V, double *invV) {
  Interp I;
  I.init(dim, deg, scheme);
  if (xi) std::copy(I.xi.begin(), I.xi.end(), xi);
  if (w) std::copy(I.w.begin(), I.w.end(), w);
  if (V) std::copy(I.V.begin(), I.V.end(), V);
  if (invV) std::copy(I.invV.begin(), I.invV.end(), invV);
  if (mi)
    for (int k = 0; k < I.q; ++k)
      for (int i = 0; i < dim; ++i) mi[i + dim * k] = I.mi[k][i];
  if (Xi) std::copy(I.Xi.begin(), I.Xi.end(), Xi);
  return I.q;
}

// evalPolynomials at npts points (dim x npts) -> q x npts.
void orc_eval_polynomials(int dim, int deg, int scheme, const double *pts, long npts,
                          double *out) {
  Interp I;
  I.init(dim, deg, scheme);
  for (long j = 0; j < npts; ++j) I.eval(pts + (size_t)dim * j, out + (size_t)I.q * j);
}

void *orc_create(int dim, long n, const double *P, long m, const double *Peval, int ktype,
                 double l, double c, double nu, int deg, double eta, int scheme) {
  if (dim < 1 || n < 1 || m < 1 || deg < 0 || ktype < 0) return nullptr;
  if (scheme != kTensor && scheme != kTotalDegree) return nullptr;
  auto t0 = std::chrono::steady_clock::now();
  Actor *A = new Actor();
  A->dim = dim;
  A->deg = deg;
  A->n = (Index)n;
  A->m = (Index)m;
  A->P.assign(P, P + (size_t)dim * n);
  A->Peval.assign(Peval, Peval + (size_t)dim * m);
  A->K = Kernel{ktype, l, c, nu};
  A->eta = eta;
  A->I.init(dim, deg, scheme);
  A->q = A->I.q;
  // H2ClusterTree::init (H2ClusterTree.h:65-74): min cluster size = max(0, q)
  const Index mincsize = (Index)A->q;
  build_tree(A->cols, dim, A->n, A->P.data(), mincsize);
  compute_bases(A->cols, 0, A->I);
  build_tree(A->rows, dim, A->m, A->Peval.data(), mincsize);
  compute_bases(A->rows, 0, A->I);
  A->blocks = compute_pattern(A->rows, A->cols, eta, A->q);
  A->scheduler.assign(A->cols.nclusters(), {});
  for (int b = 0; b < (int)A->blocks.size(); ++b)
    A->scheduler[A->cols.nodes[A->blocks[b].col].block_id].push_back(b);
  A->setup_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return A;
}

void orc_destroy(void *h) { delete (Actor *)h; }

double orc_setup_seconds(void *h) { return ((Actor *)h)->setup_seconds; }
int orc_q(void *h) { return ((Actor *)h)->q; }

static Tree &tree_of(void *h, int which) {
  Actor *A = (Actor *)h;
  return which == 0 ? A->rows : A->cols;
}

// which: 0 = row (target / Peval) tree, 1 = column (source / P) tree.
int orc_num_clusters(void *h, int which) { return tree_of(h, which).nclusters(); }

void orc_indices(void *h, int which, uint32_t *out) {
  Tree &T = tree_of(h, which);
  std::copy(T.idx.begin(), T.idx.end(), out);
}

// Per cluster in block_id order: begin, size, level, son block ids (-1), bbox (dim x 3).
void orc_clusters(void *h, int which, int64_t *begin, int64_t *size, int32_t *level,
                  int32_t *sons, double *bb) {
  Tree &T = tree_of(h, which);
  const int d = T.dim;
  for (int id = 0; id < T.nclusters(); ++id) {
    const Node &nd = T.nodes[T.by_id[id]];
    begin[id] = nd.begin;
    size[id] = nd.size;
    level[id] = nd.level;
    for (int s = 0; s < 2; ++s)
      sons[2 * id + s] = nd.son[s] >= 0 ? T.nodes[nd.son[s]].block_id : -1;
    for (int k = 0; k < 3 * d; ++k) bb[(size_t)3 * d * id + k] = nd.bb[k];
  }
}

long orc_num_blocks(void *h) { return (long)((Actor *)h)->blocks.size(); }

// Leaf blocks in level order of the block tree: row / column block ids, low-rank flag.
void orc_blocks(void *h, int32_t *row, int32_t *col, int8_t *low_rank) {
  Actor *A = (Actor *)h;
  for (size_t b = 0; b < A->blocks.size(); ++b) {
    row[b] = A->rows.nodes[A->blocks[b].row].block_id;
    col[b] = A->cols.nodes[A->blocks[b].col].block_id;
    low_rank[b] = A->blocks[b].low_rank;
  }
}

// Leaf basis V (q x size) of a column-tree leaf, FMCA moment order.
int orc_leaf_basis(void *h, int which, int block_id, double *out) {
  Tree &T = tree_of(h, which);
  const Node &nd = T.nodes[T.by_id[block_id]];
  if (nd.son[0] >= 0) return -1;
  std::copy(nd.V.begin(), nd.V.end(), out);
  return 0;
}

// Transfer matrix E_s (q x q) of an inner cluster, FMCA moment order.
int orc_transfer(void *h, int which, int block_id, int son, double *out) {
  Tree &T = tree_of(h, which);
  const Node &nd = T.nodes[T.by_id[block_id]];
  if (nd.son[0] < 0) return -1;
  std::copy(nd.E[son].begin(), nd.E[son].end(), out);
  return 0;
}

// check_transfer_matrices (compute_cluster_bases_impl.h:59-80): the largest
// relative error || V_t - [E_1 V_1, E_2 V_2] || / || V_t || over inner clusters.
double orc_check_transfer(void *h, int which) {
  Actor *A = (Actor *)h;
  Tree &T = which == 0 ? A->rows : A->cols;
  const int q = A->q, d = A->dim;
  // leaf-style moment matrices of every cluster, assembled bottom-up through E
  std::vector<std::vector<double>> Vfull(T.nclusters());
  double worst = 0;
  for (int id = T.nclusters() - 1; id >= 0; --id) {
    const int t = T.by_id[id];
    const Node &nd = T.nodes[t];
    if (nd.son[0] < 0) {
      Vfull[id] = nd.V;
      continue;
    }
    std::vector<double> direct((size_t)q * nd.size), pt(d);
    for (Index j = 0; j < nd.size; ++j) {
      const double *p = T.P + (size_t)d * T.idx[nd.begin + j];
      for (int i = 0; i < d; ++i) pt[i] = (p[i] - T.bb(t, i, 0)) / T.bb(t, i, 2);
      A->I.eval(pt.data(), direct.data() + (size_t)q * j);
    }
    std::vector<double> nested((size_t)q * nd.size, 0.);
    Index off = 0;
    for (int s = 0; s < 2; ++s) {
      const int cid = T.nodes[nd.son[s]].block_id;
      const std::vector<double> &Vs = Vfull[cid];
      const Index cs = T.nodes[nd.son[s]].size;
      for (Index j = 0; j < cs; ++j)
        for (int k = 0; k < q; ++k) {
          const double v = Vs[k + (size_t)q * j];
          for (int i = 0; i < q; ++i)
            nested[i + (size_t)q * (off + j)] += nd.E[s][i + (size_t)q * k] * v;
        }
      off += cs;
    }
    double num = 0, den = 0;
    for (size_t k = 0; k < direct.size(); ++k) {
      num += (direct[k] - nested[k]) * (direct[k] - nested[k]);
      den += direct[k] * direct[k];
    }
    worst = std::max(worst, std::sqrt(num / den));
    Vfull[id] = direct;
  }
  return worst;
}

// H2kernelActor::action (H2kernelActor.h:38-80): rhs n x nrhs and out m x nrhs,
// column-major, natural order.  nthreads <= 0 keeps the OpenMP default.
int orc_action(void *h, const double *rhs, long nrhs, double *out, int nthreads) {
  Actor *A = (Actor *)h;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const Index n = A->n, m = A->m;
  std::vector<double> srhs((size_t)n * nrhs), lhs((size_t)m * nrhs);
  for (long r = 0; r < nrhs; ++r)
    for (Index i = 0; i < n; ++i) srhs[i + (size_t)n * r] = rhs[A->cols.idx[i] + (size_t)n * r];
  action_cluster_order(*A, srhs.data(), (Index)nrhs, lhs.data());
  for (long r = 0; r < nrhs; ++r)
    for (Index i = 0; i < m; ++i) out[A->rows.idx[i] + (size_t)m * r] = lhs[i + (size_t)m * r];
  return 0;
}

// H2MatrixBase::action: input in column-tree order, output in row-tree order.
int orc_action_cluster_order(void *h, const double *srhs, long nrhs, double *lhs, int nthreads) {
  Actor *A = (Actor *)h;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  action_cluster_order(*A, srhs, (Index)nrhs, lhs);
  return 0;
}

// The rows of the action that belong to selected target leaves (row block
// ids), in cluster order, concatenated leaf after leaf: the same blocks and
// passes as the full action restricted to the ancestors of each leaf, so that
// large problems can be checked without a full CPU action.
int orc_action_leaves(void *h, const double *rhs, long nrhs, const int32_t *leaves, int nleaves,
                      double *out, int nthreads) {
  Actor *A = (Actor *)h;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const Index n = A->n;
  const int q = A->q, d = A->dim;
  std::vector<double> srhs((size_t)n * nrhs);
  for (long r = 0; r < nrhs; ++r)
    for (Index i = 0; i < n; ++i) srhs[i + (size_t)n * r] = rhs[A->cols.idx[i] + (size_t)n * r];
  std::vector<std::vector<double>> trhs(A->cols.nclusters());
  forward_rec(*A, 0, srhs.data(), (Index)nrhs, trhs);
  // blocks by row storage position
  std::vector<std::vector<int>> by_row(A->rows.nclusters());
  for (int b = 0; b < (int)A->blocks.size(); ++b) by_row[A->blocks[b].row].push_back(b);
  size_t off = 0;
  for (int li = 0; li < nleaves; ++li) {
    const int t = A->rows.by_id[leaves[li]];
    const Node &lt = A->rows.nodes[t];
    if (lt.son[0] >= 0) return -1;
    std::vector<int> chain;
    for (int a = t; a >= 0; a = A->rows.nodes[a].dad) chain.push_back(a);
    std::reverse(chain.begin(), chain.end());  // root ... leaf
    std::vector<double> acc((size_t)q * nrhs, 0.), next;
    std::vector<double> near((size_t)lt.size * nrhs, 0.);
    std::vector<double> S;
    for (size_t ci = 0; ci < chain.size(); ++ci) {
      const int a = chain[ci];
      const Node &an = A->rows.nodes[a];
      if (ci > 0) {  // push the parent's far field down: acc = E^T acc_parent
        const Node &par = A->rows.nodes[chain[ci - 1]];
        const int s = par.son[0] == a ? 0 : 1;
        next.assign((size_t)q * nrhs, 0.);
        for (long r = 0; r < nrhs; ++r)
          for (int k = 0; k < q; ++k) {
            double v = 0;
            for (int i = 0; i < q; ++i) v += par.E[s][i + (size_t)q * k] * acc[i + (size_t)q * r];
            next[k + (size_t)q * r] = v;
          }
        acc.swap(next);
      }
      for (int b : by_row[a]) {
        const Block &blk = A->blocks[b];
        const Node &ct = A->cols.nodes[blk.col];
        if (blk.low_rank) {
          coupling_matrix(*A, a, blk.col, S);
          const std::vector<double> &x = trhs[ct.block_id];
          for (long r = 0; r < nrhs; ++r)
            for (int j = 0; j < q; ++j)
              for (int i = 0; i < q; ++i)
                acc[i + (size_t)q * r] += S[i + (size_t)q * j] * x[j + (size_t)q * r];
        } else {
          // dense block rows restricted to this leaf
          const Index r0 = lt.begin - an.begin;
          for (Index i = 0; i < lt.size; ++i) {
            const double *x = A->Peval.data() + (size_t)d * A->rows.idx[lt.begin + i];
            for (Index j = 0; j < ct.size; ++j) {
              const double *y = A->P.data() + (size_t)d * A->cols.idx[ct.begin + j];
              const double kv = A->K(x, y, d);
              for (long r = 0; r < nrhs; ++r)
                near[i + (size_t)lt.size * r] += kv * srhs[ct.begin + j + (size_t)n * r];
            }
          }
          (void)r0;
        }
      }
    }
    for (long r = 0; r < nrhs; ++r)
      for (Index j = 0; j < lt.size; ++j) {
        double v = 0;
        for (int i = 0; i < q; ++i) v += lt.V[i + (size_t)q * j] * acc[i + (size_t)q * r];
        out[off + j + (size_t)lt.size * r] = near[j + (size_t)lt.size * r] + v;
      }
    off += (size_t)lt.size * nrhs;
  }
  return 0;
}

// Dense reference K(Peval, P) * rhs (CovarianceKernel::eval, CovarianceKernel.h:145-151).
void orc_dense_action(int dim, long n, const double *P, long m, const double *Peval, int ktype,
                      double l, double c, double nu, const double *rhs, long nrhs, double *out,
                      int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  Kernel K{ktype, l, c, nu};
#pragma omp parallel for schedule(static)
  for (long i = 0; i < m; ++i) {
    std::vector<double> acc(nrhs, 0.);
    const double *x = Peval + (size_t)dim * i;
    for (long j = 0; j < n; ++j) {
      const double kv = K(x, P + (size_t)dim * j, dim);
      for (long r = 0; r < nrhs; ++r) acc[r] += kv * rhs[j + (size_t)n * r];
    }
    for (long r = 0; r < nrhs; ++r) out[i + (size_t)m * r] = acc[r];
  }
}

// Dense kernel matrix K(PR, PC), m x n column-major.
void orc_dense_matrix(int dim, long n, const double *P, long m, const double *Peval, int ktype,
                      double l, double c, double nu, double *out) {
  Kernel K{ktype, l, c, nu};
#pragma omp parallel for schedule(static)
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < m; ++i)
      out[i + (size_t)m * j] = K(Peval + (size_t)dim * i, P + (size_t)dim * j, dim);
}

// pCG (PCG.h:15-44) with the identity preconditioner on (A + ridge I) x = b,
// single right-hand side, natural order.  Returns the iteration count.
int orc_pcg(void *h, const double *b, double *x, double ridge, double tol, int max_iter,
            double *final_err, int nthreads) {
  Actor *A = (Actor *)h;
  const Index n = A->n;
  if (A->m != n) return -1;
  auto apply = [&](const std::vector<double> &v, std::vector<double> &Av) {
    orc_action(h, v.data(), 1, Av.data(), nthreads);
    for (Index i = 0; i < n; ++i) Av[i] += ridge * v[i];
  };
  auto dot = [n](const std::vector<double> &a, const std::vector<double> &c) {
    double s = 0;
    for (Index i = 0; i < n; ++i) s += a[i] * c[i];
    return s;
  };
  std::vector<double> xv(b, b + n), bv(b, b + n), r(n), Ap(n), p;
  apply(xv, Ap);
  for (Index i = 0; i < n; ++i) r[i] = bv[i] - Ap[i];
  std::vector<double> z = r;
  p = z;
  const double bn = std::sqrt(dot(bv, bv));
  double err = std::sqrt(dot(r, r)) / bn;
  double rdotz_old = dot(r, z);
  int iter = 0;
  while (err > tol && iter < max_iter) {
    apply(p, Ap);
    const double alpha = dot(r, z) / dot(p, Ap);
    for (Index i = 0; i < n; ++i) {
      xv[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    err = std::sqrt(dot(r, r)) / bn;
    z = r;
    const double rdotz = dot(r, z);
    const double beta = rdotz / rdotz_old;
    rdotz_old = rdotz;
    for (Index i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++iter;
  }
  std::copy(xv.begin(), xv.end(), x);
  if (final_err) *final_err = err;
  return iter;
}

// Kernel-entry counts of the block pattern: near entries, far entries, far blocks,
// near blocks.
void orc_entry_counts(void *h, double *out4) {
  Actor *A = (Actor *)h;
  double en = 0, ef = 0, nf = 0, nn = 0;
  for (const Block &b : A->blocks) {
    if (b.low_rank) {
      ef += double(A->q) * A->q;
      nf += 1;
    } else {
      en += double(A->rows.nodes[b.row].size) * A->cols.nodes[b.col].size;
      nn += 1;
    }
  }
  out4[0] = en;
  out4[1] = ef;
  out4[2] = nf;
  out4[3] = nn;
}

// CPU baseline sample: runs FMCA's schedule (forward pass, then the column
// groups in order, then the backward pass) but stops the coupling/near loop
// after the column groups that hold `entry_budget` kernel entries.  Returns the
// entries processed; times[0..2] = forward, sampled groups, backward seconds.
double orc_action_sample(void *h, const double *rhs, long nrhs, double entry_budget,
                         int nthreads, double *times) {
  Actor *A = (Actor *)h;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const Index n = A->n, m = A->m;
  std::vector<double> srhs((size_t)n * nrhs), lhs((size_t)m * nrhs, 0.);
  for (long r = 0; r < nrhs; ++r)
    for (Index i = 0; i < n; ++i) srhs[i + (size_t)n * r] = rhs[A->cols.idx[i] + (size_t)n * r];
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto t0 = now();
  std::vector<std::vector<double>> trhs(A->cols.nclusters());
  forward_rec(*A, 0, srhs.data(), (Index)nrhs, trhs);
  std::vector<std::vector<double>> tlhs(A->rows.nclusters());
  for (auto &v : tlhs) v.assign((size_t)A->q * nrhs, 0.);
  auto t1 = now();
  double done = 0;
  // Visit column groups in a strided order so that the sample covers all
  // levels of the column tree rather than only its top.
  const size_t G = A->scheduler.size();
  std::vector<size_t> order(G);
  std::iota(order.begin(), order.end(), 0);
  const size_t stride = 7919;  // prime, coprime with any power of two
  for (size_t g = 0; g < G; ++g) order[g] = (g * stride) % G;
  for (size_t gi = 0; gi < G && done < entry_budget; ++gi) {
    const auto &group = A->scheduler[order[gi]];
#pragma omp parallel
    {
      std::vector<double> S;
#pragma omp for schedule(dynamic)
      for (long k = 0; k < (long)group.size(); ++k)
        apply_block(*A, A->blocks[group[k]], srhs.data(), lhs.data(), (Index)nrhs, trhs, tlhs, S);
    }
    for (int b : group) {
      const Block &blk = A->blocks[b];
      done += blk.low_rank ? double(A->q) * A->q
                           : double(A->rows.nodes[blk.row].size) * A->cols.nodes[blk.col].size;
    }
  }
  auto t2 = now();
  backward_rec(*A, 0, lhs.data(), (Index)nrhs, tlhs);
  auto t3 = now();
  if (times) {
    times[0] = std::chrono::duration<double>(t1 - t0).count();
    times[1] = std::chrono::duration<double>(t2 - t1).count();
    times[2] = std::chrono::duration<double>(t3 - t2).count();
  }
  return done;
}

}  // extern "C"
