// Host-side planner: cluster tree, block pattern, Chebyshev data, transfers.
// Level-synchronous and thread-parallel; produces FMCA's tree and blocks.
#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <utility>

namespace fmca_amd {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();

// Host worker threads: FMCA_AMD_THREADS, else OMP_NUM_THREADS, else the
// hardware concurrency, at most 64.  The planner runs its parallel loops on
// plain std::threads so that it brings no OpenMP runtime into the process.
int num_workers() {
  for (const char *v : {"FMCA_AMD_THREADS", "OMP_NUM_THREADS"})
    if (const char *e = std::getenv(v)) {
      const int k = std::atoi(e);
      if (k > 0) return std::min(k, 64);
    }
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h, 64u));
}

// f(i) for i in [0, n), dynamically scheduled in chunks of `grain`.
template <typename F>
void parallel_for(int64_t n, int64_t grain, F &&f) {
  const int nw = (int)std::min<int64_t>(num_workers(), (n + grain - 1) / std::max<int64_t>(grain, 1));
  if (nw <= 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (;;) {
      const int64_t lo = next.fetch_add(grain);
      if (lo >= n) break;
      const int64_t hi = std::min(n, lo + grain);
      for (int64_t i = lo; i < hi; ++i) f(i);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nw; ++t) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

// f(tid, lo, hi) on nt contiguous static chunks of [0, n).
template <typename F>
void parallel_chunks(int64_t n, int nt, F &&f) {
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back([&, t] { f(t, n * t / nt, n * (t + 1) / nt); });
  f(0, 0, n / nt);
  for (auto &th : pool) th.join();
}
constexpr double kPi = 3.14159265358979323846264338327950288;
const char *const kNames[KT_COUNT] = {
    "EXPONENTIAL",     "MATERN32",     "MATERN52",     "MATERN72",
    "MATERN92",        "GAUSSIAN",     "INVMULTIQUADRIC", "MULTIQUADRIC",
    "BIHARMONIC2D",    "BIHARMONIC3D", "TRIHARMONIC3D",   "TPS2D"};
}  // namespace

int parse_kernel_type(const char *name) {
  if (!name) return -1;
  std::string s(name);
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  for (int i = 0; i < KT_COUNT; ++i)
    if (s == kNames[i]) return i;
  return -1;
}

const char *kernel_type_name(int kt) {
  return (kt >= 0 && kt < KT_COUNT) ? kNames[kt] : "UNKNOWN";
}

// ---------------------------------------------------------------------------
namespace {

// Point k (k >= 0) of the Halton sequence in the first `dim` prime bases,
// radical inverses by Horner from the leading digit in long double, as
// HaltonSet::next forms them (util/HaltonSet.h:37-66).
void halton_point(int64_t k, int dim, double *out) {
  static const int kPrimes[kMaxDim] = {2, 3, 5};
  for (int i = 0; i < dim; ++i) {
    const int b = kPrimes[i];
    int dig[64], nd = 0;
    for (int64_t r = k; r > 0; r /= b) dig[nd++] = (int)(r % b);
    const long double binv = 1.0L / (long double)b;
    long double v = nd ? dig[nd - 1] : 0;
    for (int j = nd - 2; j >= 0; --j) v = binv * v + dig[j];
    out[i] = (double)(v * binv);
  }
}

// Greedy column selection of Householder QR with column pivoting on the
// rows x cols column-major matrix A (overwritten): step k takes the remaining
// column of largest downdated norm (lowest index on ties), reflects it onto
// e_k and downdates the other norms, recomputing a norm once the downdate has
// cancelled more than half of its digits (the rule of LAPACK's xGEQP3, which
// Eigen's ColPivHouseholderQR follows).  Returns the original column index of
// each of the first min(rows, cols) pivots.
std::vector<int> greedy_pivots(std::vector<double> &A, int rows, int cols) {
  const int steps = std::min(rows, cols);
  std::vector<int> idx(cols);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<double> ref(cols), cur(cols);  // last recomputed / downdated norms
  auto c = [&](int j) { return &A[(size_t)rows * j]; };
  auto tail_norm = [&](int j, int from) {
    double s = 0;
    for (int i = from; i < rows; ++i) s += c(j)[i] * c(j)[i];
    return std::sqrt(s);
  };
  for (int j = 0; j < cols; ++j) ref[j] = cur[j] = tail_norm(j, 0);
  const double tol = std::sqrt(kEps);
  for (int k = 0; k < steps; ++k) {
    const int piv = (int)(std::max_element(cur.begin() + k, cur.end()) - cur.begin());
    if (piv != k) {
      for (int i = 0; i < rows; ++i) std::swap(c(k)[i], c(piv)[i]);
      std::swap(ref[k], ref[piv]);
      std::swap(cur[k], cur[piv]);
      std::swap(idx[k], idx[piv]);
    }
    double *v = c(k);
    double s2 = 0;
    for (int i = k + 1; i < rows; ++i) s2 += v[i] * v[i];
    double tau = 0;
    const double alpha = v[k];
    double beta = alpha;
    if (s2 > std::numeric_limits<double>::min()) {
      beta = std::sqrt(alpha * alpha + s2);
      if (alpha >= 0) beta = -beta;
      for (int i = k + 1; i < rows; ++i) v[i] /= (alpha - beta);
      tau = (beta - alpha) / beta;
    } else {
      for (int i = k + 1; i < rows; ++i) v[i] = 0;
    }
    v[k] = beta;
    for (int j = k + 1; j < cols; ++j) {
      double *u = c(j);
      if (tau != 0) {
        double dot = u[k];
        for (int i = k + 1; i < rows; ++i) dot += v[i] * u[i];
        u[k] -= tau * dot;
        for (int i = k + 1; i < rows; ++i) u[i] -= tau * v[i] * dot;
      }
      if (cur[j] == 0) continue;
      double t = std::abs(u[k]) / cur[j];
      t = std::max(0., (1 + t) * (1 - t));
      const double rr = cur[j] / ref[j];
      if (t * (rr * rr) <= tol)
        ref[j] = cur[j] = tail_norm(j, k + 1);
      else
        cur[j] *= std::sqrt(t);
    }
  }
  idx.resize(steps);
  return idx;
}

// Inverse of the n x n column-major matrix M by LU with partial pivoting.
std::vector<double> lu_inverse(std::vector<double> M, int n) {
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  auto at = [&](int i, int j) -> double & { return M[i + (size_t)n * j]; };
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(at(i, k)) > std::abs(at(p, k))) p = i;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(at(k, j), at(p, j));
      std::swap(perm[k], perm[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      at(i, k) /= at(k, k);
      for (int j = k + 1; j < n; ++j) at(i, j) -= at(i, k) * at(k, j);
    }
  }
  std::vector<double> X((size_t)n * n, 0.);
  for (int col = 0; col < n; ++col) {
    double *x = &X[(size_t)n * col];
    for (int i = 0; i < n; ++i) x[i] = perm[i] == col ? 1. : 0.;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) x[i] -= at(i, j) * x[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[i] -= at(i, j) * x[j];
      x[i] /= at(i, i);
    }
  }
  return X;
}

}  // namespace

void Interpolation::legendre_1d(double u, double *out) const {
  double prev = 0, cur = 1;
  out[0] = 1;
  for (int a = 1; a <= deg; ++a) {
    const double next =
        double(2 * a - 1) / double(a) * (2 * u - 1) * cur - double(a - 1) / double(a) * prev;
    prev = cur;
    cur = next;
    out[a] = next * std::sqrt(double(2 * a + 1));
  }
}

void Interpolation::basis(const double *u, double *out) const {
  double f[kMaxDim][kMaxP1];
  for (int i = 0; i < dim; ++i) legendre_1d(u[i], f[i]);
  for (int k = 0; k < q; ++k) {
    double v = 1;
    for (int i = 0; i < dim; ++i) v *= f[i][(midx[k] >> (8 * i)) & 0xff];
    out[k] = v;
  }
}

void Interpolation::init(int d, int p, int sch) {
  dim = d;
  deg = p;
  P1 = p + 1;
  scheme = sch;
  midx.clear();
  nodes.clear();
  invV.clear();
  if (sch == IS_TOTAL_DEGREE) {
    // {alpha : |alpha|_1 <= p} in graded-lexicographic order
    std::vector<std::vector<int>> mi;
    std::vector<int> a(d, 0);
    for (;;) {
      int sum = 0;
      for (int v : a) sum += v;
      if (sum <= p) mi.push_back(a);
      int i = 0;
      while (i < d && a[i] == p) a[i++] = 0;
      if (i == d) break;
      ++a[i];
    }
    std::sort(mi.begin(), mi.end(), [](const std::vector<int> &x, const std::vector<int> &y) {
      const int sx = std::accumulate(x.begin(), x.end(), 0);
      const int sy = std::accumulate(y.begin(), y.end(), 0);
      return sx != sy ? sx < sy : x < y;
    });
    q = (int)mi.size();
    midx.resize(q);
    for (int k = 0; k < q; ++k) {
      int pk = 0;
      for (int i = 0; i < d; ++i) pk |= mi[k][i] << (8 * i);
      midx[k] = pk;
    }
    fmca_order.resize(q);
    std::iota(fmca_order.begin(), fmca_order.end(), 0);
    // DeMarchi points: Halton points 100 .. 100 + 10 q - 1, the first q
    // pivots of the column-pivoted QR of their basis values
    const int ncand = 10 * q;
    std::vector<double> cand((size_t)d * ncand), B((size_t)q * ncand);
    for (int j = 0; j < ncand; ++j) {
      halton_point(100 + j, d, &cand[(size_t)d * j]);
      basis(&cand[(size_t)d * j], &B[(size_t)q * j]);
    }
    const std::vector<int> piv = greedy_pivots(B, q, ncand);
    nodes.resize((size_t)d * q);
    for (int k = 0; k < q; ++k)
      for (int i = 0; i < d; ++i) nodes[(size_t)d * k + i] = cand[(size_t)d * piv[k] + i];
    std::vector<double> V((size_t)q * q);
    std::vector<double> row(q);
    for (int k = 0; k < q; ++k) {
      basis(&nodes[(size_t)d * k], row.data());
      for (int a2 = 0; a2 < q; ++a2) V[k + (size_t)q * a2] = row[a2];
    }
    invV = lu_inverse(V, q);
    return;
  }
  q = 1;
  for (int i = 0; i < d; ++i) q *= P1;
  const double weight = 2. * kPi / (2. * double(p) + 2.);
  xi.resize(P1);
  w.resize(P1);
  for (int k = 0; k < P1; ++k) {
    xi[k] = 0.5 * (std::cos(weight * (double(k) + 0.5)) + 1);
    w[k] = std::sin(weight * (double(k) + 0.5));
    if (k & 1) w[k] = -w[k];
  }
  // graded-lexicographic order of {0..p}^d: by l1 norm, then lexicographic
  // with the first entry most significant (MultiIndexSet.h:87-100).
  std::vector<std::vector<int>> idx;
  idx.reserve(q);
  for (int lin = 0; lin < q; ++lin) {
    std::vector<int> a(d);
    int r = lin;
    for (int i = 0; i < d; ++i) {
      a[i] = r % P1;
      r /= P1;
    }
    idx.push_back(a);
  }
  std::vector<int> order(q);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    int sx = 0, sy = 0;
    for (int v : idx[x]) sx += v;
    for (int v : idx[y]) sy += v;
    if (sx != sy) return sx < sy;
    return idx[x] < idx[y];
  });
  fmca_order.assign(order.begin(), order.end());
}

void Interpolation::lagrange_1d(double u, double *out) const {
  double s = 0;
  for (int j = 0; j < P1; ++j) s += w[j] / (u - xi[j]);
  for (int a = 0; a < P1; ++a) {
    const double d = u - xi[a];
    out[a] = std::abs(d) > kEps ? w[a] / d / s : 1.;
  }
}

// ---------------------------------------------------------------------------
namespace {
struct WorkNode {
  int64_t begin = 0, size = 0;
  int32_t parent = -1, level = 0, son0 = -1, son1 = -1;
  int32_t axis = -1;       // split axis of a refined node
  double bb[3 * kMaxDim];  // [min | max | size], each dim entries
};

// The host splitter.  exact: every split sorts its range as FMCA does
// (std::sort by key), so ties between equal coordinates resolve as in FMCA.
// Otherwise the splits only partition (std::nth_element) and each leaf is
// sorted once, by its father's split key, at the end.  With pairwise distinct
// keys at every split point and inside every leaf the two give the same
// permutation: FMCA's order inside a leaf is the order of its father's sort,
// and a partition at the median is the same set either way.  Equal keys there
// make the fast build report a tie; the caller then builds exactly.
class HostSplitter final : public TreeSplitter {
 public:
  HostSplitter(int dim, int64_t n, const double *P, bool exact)
      : dim_(dim), P_(P), exact_(exact), perm_(n) {
    std::iota(perm_.begin(), perm_.end(), 0u);
  }
  bool split(const std::vector<SplitSeg> &segs, double *lmax, double *rmin) override {
    std::atomic<bool> tie{false};
    // Sorting (key, index) pairs by key performs the same comparisons as
    // sorting indices through the coordinate, so std::sort yields the same
    // permutation, ties included.
    parallel_for((int64_t)segs.size(), 1, [&](int64_t k) {
      const SplitSeg &sg = segs[k];
      std::vector<std::pair<double, uint32_t>> kv(sg.size);
      uint32_t *id = perm_.data() + sg.begin;
      for (int64_t j = 0; j < sg.size; ++j) kv[j] = {P_[sg.axis + (size_t)dim_ * id[j]], id[j]};
      const int64_t half = sg.size / 2;
      if (exact_) {
        std::sort(kv.begin(), kv.end(), by_key);
        lmax[k] = kv[half - 1].first;
      } else {
        std::nth_element(kv.begin(), kv.begin() + half, kv.end(), by_key);
        lmax[k] = std::max_element(kv.begin(), kv.begin() + half, by_key)->first;
        if (!(lmax[k] < kv[half].first)) tie = true;
      }
      rmin[k] = kv[half].first;
      for (int64_t j = 0; j < sg.size; ++j) id[j] = kv[j].second;
    });
    return !tie;
  }
  bool finish_leaves(const std::vector<SplitSeg> &leaves, double *mn, double *mx) override {
    std::atomic<bool> tie{false};
    parallel_for((int64_t)leaves.size(), 16, [&](int64_t k) {
      const SplitSeg &lf = leaves[k];
      uint32_t *id = perm_.data() + lf.begin;
      if (!exact_ && lf.axis >= 0 && lf.size > 1) {
        // leaves in their father's sort order
        std::vector<std::pair<double, uint32_t>> kv(lf.size);
        for (int64_t j = 0; j < lf.size; ++j) kv[j] = {P_[lf.axis + (size_t)dim_ * id[j]], id[j]};
        std::sort(kv.begin(), kv.end(), by_key);
        for (int64_t j = 0; j < lf.size; ++j) {
          id[j] = kv[j].second;
          if (j && !(kv[j - 1].first < kv[j].first)) tie = true;
        }
      }
      double *a = mn + (size_t)k * dim_, *b = mx + (size_t)k * dim_;
      const double *p0 = P_ + (size_t)dim_ * id[0];
      for (int i = 0; i < dim_; ++i) a[i] = b[i] = p0[i];
      for (int64_t j = 0; j < lf.size; ++j) {
        const double *p = P_ + (size_t)dim_ * id[j];
        for (int i = 0; i < dim_; ++i) {
          a[i] = std::min(p[i], a[i]);
          b[i] = std::max(p[i], b[i]);
        }
      }
    });
    return !tie;
  }
  void permutation(std::vector<uint32_t> &perm) override { perm.swap(perm_); }

 private:
  static bool by_key(const std::pair<double, uint32_t> &a, const std::pair<double, uint32_t> &b) {
    return a.first < b.first;
  }
  int dim_;
  const double *P_;
  bool exact_;
  std::vector<uint32_t> perm_;
};
}  // namespace

void parallel_ranges(int64_t n, const std::function<void(int64_t, int64_t)> &f) {
  const int nw = (int)std::max<int64_t>(1, std::min<int64_t>(num_workers(), n / 65536 + 1));
  parallel_chunks(n, nw, [&](int, int64_t lo, int64_t hi) { f(lo, hi); });
}

void parallel_dynamic(int64_t n, int64_t grain, const std::function<void(int64_t)> &f) {
  parallel_for(n, grain, f);
}

void point_box(int dim, int64_t n, const double *P, double *mn, double *mx) {
  const int nw = (int)std::max<int64_t>(1, std::min<int64_t>(num_workers(), n / 65536 + 1));
  for (int i = 0; i < dim; ++i) {
    std::vector<double> tmn(nw, P[i]), tmx(nw, P[i]);
    parallel_chunks(n, nw, [&](int t, int64_t lo, int64_t hi) {
      double a = P[i], b = P[i];
      for (int64_t j = lo; j < hi; ++j) {
        const double v = P[i + (size_t)dim * j];
        a = v < a ? v : a;
        b = v > b ? v : b;
      }
      tmn[t] = a;
      tmx[t] = b;
    });
    mn[i] = mx[i] = P[i];
    for (int t = 0; t < nw; ++t) {
      mn[i] = tmn[t] < mn[i] ? tmn[t] : mn[i];
      mx[i] = tmx[t] > mx[i] ? tmx[t] : mx[i];
    }
  }
}

void build_tree(Tree &T, int dim, int64_t n, const double *P, int64_t min_csize) {
  double mn[kMaxDim], mx[kMaxDim];
  point_box(dim, n, P, mn, mx);
  {
    HostSplitter fast(dim, n, P, false);
    if (build_tree_with(T, dim, n, P, mn, mx, min_csize, fast)) return;
  }
  HostSplitter exact(dim, n, P, true);
  build_tree_with(T, dim, n, P, mn, mx, min_csize, exact);
}

bool build_tree_with(Tree &T, int dim, int64_t n, const double *P, const double *root_min,
                     const double *root_max, int64_t min_csize, TreeSplitter &sp) {
  T = Tree();
  T.dim = dim;
  T.npts = n;
  std::vector<WorkNode> nodes(1);
  WorkNode &root = nodes[0];
  root.begin = 0;
  root.size = n;
  // eps-enlarged root box (ClusterTreeInitializer_ClusterTree.h:44-56)
  for (int i = 0; i < dim; ++i) {
    double mn = root_min[i], mx = root_max[i];
    mn -= kEps * std::abs(mn);
    mx += kEps * std::abs(mx);
    root.bb[i] = mn;
    root.bb[i + dim] = mx;
    root.bb[i + 2 * dim] = mx - mn;
  }
  const int64_t thr = min_csize >= 1 ? (2 * min_csize - 1) : 1;
  std::vector<int32_t> level_ptr{0};
  int64_t lb = 0, le = 1;
  int level = 0;
  std::vector<SplitSeg> segs;
  std::vector<int64_t> seg_node;
  std::vector<double> lmax, rmin;
  while (lb < le) {
    // Split every node of this level above the leaf size: sort its index
    // range by the longest edge of its split box and halve it
    // (ClusterTreeSplitter.h:60-81).
    segs.clear();
    seg_node.clear();
    for (int64_t k = lb; k < le; ++k) {
      WorkNode &nd = nodes[k];
      if (nd.size <= thr) continue;
      int longest = 0;
      for (int i = 1; i < dim; ++i)
        if (nd.bb[i + 2 * dim] > nd.bb[longest + 2 * dim]) longest = i;
      nd.axis = longest;
      segs.push_back({nd.begin, nd.size, longest});
      seg_node.push_back(k);
    }
    lmax.assign(segs.size(), 0.);
    rmin.assign(segs.size(), 0.);
    if (!segs.empty() && !sp.split(segs, lmax.data(), rmin.data())) return false;
    for (size_t s = 0; s < segs.size(); ++s) {
      const int64_t k = seg_node[s];
      const int ax = segs[s].axis;
      WorkNode c1, c2;
      std::memcpy(c1.bb, nodes[k].bb, sizeof(c1.bb));
      std::memcpy(c2.bb, nodes[k].bb, sizeof(c2.bb));
      c1.begin = nodes[k].begin;
      c1.size = nodes[k].size / 2;
      c2.begin = nodes[k].begin + c1.size;
      c2.size = nodes[k].size - c1.size;
      c1.bb[ax + dim] = lmax[s];
      c1.bb[ax + 2 * dim] = c1.bb[ax + dim] - c1.bb[ax];
      c2.bb[ax] = rmin[s];
      c2.bb[ax + 2 * dim] = c2.bb[ax + dim] - c2.bb[ax];
      c1.level = c2.level = level + 1;
      c1.parent = c2.parent = (int32_t)k;
      nodes[k].son0 = (int32_t)nodes.size();
      nodes.push_back(c1);
      nodes[k].son1 = (int32_t)nodes.size();
      nodes.push_back(c2);
    }
    level_ptr.push_back((int32_t)le);
    lb = le;
    le = (int64_t)nodes.size();
    ++level;
  }
  const int ncl = (int)nodes.size();
  T.ncl = ncl;
  T.depth = (int)level_ptr.size() - 2;
  T.level_ptr = level_ptr;
  T.begin.resize(ncl);
  T.size.resize(ncl);
  T.level.resize(ncl);
  T.son0.resize(ncl);
  T.son1.resize(ncl);
  T.parent.resize(ncl);
  T.bbmin.assign((size_t)ncl * dim, 0.);
  T.bbmax.assign((size_t)ncl * dim, 0.);
  T.bbsize.assign((size_t)ncl * dim, 0.);
  std::vector<SplitSeg> leaves;
  for (int c = 0; c < ncl; ++c) {
    T.begin[c] = nodes[c].begin;
    T.size[c] = nodes[c].size;
    T.level[c] = nodes[c].level;
    T.son0[c] = nodes[c].son0;
    T.son1[c] = nodes[c].son1;
    T.parent[c] = nodes[c].parent;
    if (nodes[c].son0 < 0) {
      T.leaves.push_back(c);
      const int fa = nodes[c].parent;
      leaves.push_back({nodes[c].begin, nodes[c].size, fa >= 0 ? nodes[fa].axis : -1});
    }
  }
  // Leaf point boxes (and, for the fast build, the leaves' final order).
  std::vector<double> lmn(leaves.size() * dim), lmx(leaves.size() * dim);
  if (!sp.finish_leaves(leaves, lmn.data(), lmx.data())) return false;
  sp.permutation(T.perm);
  std::vector<int32_t> leaf_index(ncl, -1);
  for (size_t k = 0; k < T.leaves.size(); ++k) leaf_index[T.leaves[k]] = (int32_t)k;
  // Shrink to tight boxes, deepest level first (ClusterTreeInitializer_ClusterTree.h:88-132).
  // A leaf's box is the min / max over its points; a father's over its first
  // point and its sons' (widened) boxes.
  for (int L = T.depth; L >= 0; --L) {
    const int cl0 = T.level_ptr[L];
    parallel_for(T.level_ptr[L + 1] - cl0, 16, [&](int64_t ci) {
      const int c = cl0 + (int)ci;
      double mn[kMaxDim], mx[kMaxDim];
      if (T.son0[c] >= 0) {
        const double *p0 = P + (size_t)dim * T.perm[T.begin[c]];
        for (int i = 0; i < dim; ++i) mn[i] = mx[i] = p0[i];
        for (int s : {T.son0[c], T.son1[c]})
          for (int i = 0; i < dim; ++i) {
            mn[i] = std::min(mn[i], T.bbmin[(size_t)s * dim + i]);
            mx[i] = std::max(mx[i], T.bbmax[(size_t)s * dim + i]);
          }
      } else {
        const int k = leaf_index[c];
        for (int i = 0; i < dim; ++i) {
          mn[i] = lmn[(size_t)k * dim + i];
          mx[i] = lmx[(size_t)k * dim + i];
        }
      }
      for (int i = 0; i < dim; ++i) {
        double sz = mx[i] - mn[i];
        if (mx[i] - mn[i] < 2e2 * kEps) {
          mx[i] += 1e2 * kEps;
          mn[i] -= 1e2 * kEps;
          sz = mx[i] - mn[i];
        }
        T.bbmin[(size_t)c * dim + i] = mn[i];
        T.bbmax[(size_t)c * dim + i] = mx[i];
        T.bbsize[(size_t)c * dim + i] = sz;
      }
    });
  }
  return true;
}

// ---------------------------------------------------------------------------
namespace {
double radius(const Tree &T, int c) {
  double s = 0;
  for (int i = 0; i < T.dim; ++i) {
    const double v = T.bbsize[(size_t)c * T.dim + i];
    s += v * v;
  }
  return 0.5 * std::sqrt(s);
}

// What the admissibility test reads of a cluster, in one cache line: box min,
// box max, radius and whether it is a leaf.
struct alignas(64) ClusterGeo {
  double mn[kMaxDim], mx[kMaxDim], rad;
  int leaf;
};

std::vector<ClusterGeo> cluster_geo(const Tree &T) {
  std::vector<ClusterGeo> g(T.ncl);
  for (int c = 0; c < T.ncl; ++c) {
    for (int i = 0; i < T.dim; ++i) {
      g[c].mn[i] = T.bbmin[(size_t)c * T.dim + i];
      g[c].mx[i] = T.bbmax[(size_t)c * T.dim + i];
    }
    g[c].rad = radius(T, c);
    g[c].leaf = T.is_leaf(c) ? 1 : 0;
  }
  return g;
}

// CompareCluster::compare (Admissibility.h:95-115) with computeDistance (:19-29).
// 0 = refine, 1 = low rank, 2 = dense.
int admissibility(const ClusterGeo &r, const ClusterGeo &c, int d, double eta) {
  double s2 = 0;
  for (int i = 0; i < d; ++i) {
    const double v = 0.5 * (r.mn[i] - c.mn[i] + r.mx[i] - c.mx[i]);
    s2 += v * v;
  }
  double dist = std::sqrt(s2) - r.rad - c.rad;
  dist = dist > 0 ? dist : 0;
  const double rad = r.rad > c.rad ? r.rad : c.rad;
  if (rad > eta * dist) return (r.leaf || c.leaf) ? 2 : 0;
  return 1;
}
}  // namespace

namespace {
// CSR of the (key, value) pairs that items b in [0, nb) emit: emit(b, put)
// calls put(key, value) for each pair of item b.  The values of one key keep
// the order of their items, as a serial fill would: the items are cut into
// one contiguous chunk per worker, counted per chunk, and each chunk then
// fills from its own offsets.
template <typename Emit>
void build_csr(int64_t nb, int nkeys, Emit emit, std::vector<int32_t> &ptr,
               std::vector<int32_t> &out) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(num_workers(), nb / 65536 + 1));
  std::vector<std::vector<int32_t>> cnt(nt, std::vector<int32_t>(nkeys, 0));
  parallel_chunks(nb, nt, [&](int t, int64_t lo, int64_t hi) {
    int32_t *c = cnt[t].data();
    for (int64_t b = lo; b < hi; ++b) emit(b, [&](int k, int32_t) { ++c[k]; });
  });
  ptr.assign(nkeys + 1, 0);
  for (int k = 0; k < nkeys; ++k) {
    int32_t sum = 0;
    for (int t = 0; t < nt; ++t) sum += cnt[t][k];
    ptr[k + 1] = ptr[k] + sum;
  }
  parallel_for(nkeys, 1024, [&](int64_t k) {  // counts -> each chunk's first position
    int32_t o = ptr[k];
    for (int t = 0; t < nt; ++t) {
      const int32_t c = cnt[t][k];
      cnt[t][k] = o;
      o += c;
    }
  });
  out.assign(ptr[nkeys], 0);
  parallel_chunks(nb, nt, [&](int t, int64_t lo, int64_t hi) {
    int32_t *f = cnt[t].data();
    for (int64_t b = lo; b < hi; ++b) emit(b, [&](int k, int32_t v) { out[f[k]++] = v; });
  });
}
}  // namespace

double cluster_radius(const Tree &T, int c) { return radius(T, c); }

void compute_pattern(Pattern &pat, const Tree &R, const Tree &C, double eta, int q) {
  pat = Pattern();
  expand_blocks(pat, R, C, eta, q);
  finish_pattern(pat, R, C, q);
}

void expand_blocks(Pattern &pat, const Tree &R, const Tree &C, double eta, int q) {
  // Level-synchronous expansion of the block tree; the set of leaf blocks is
  // the one of FMCA's depth-first construction (H2MatrixBase.h:160-191).
  std::vector<std::pair<int32_t, int32_t>> frontier{{0, 0}};
  const int nth = num_workers();
  const std::vector<ClusterGeo> gr = cluster_geo(R);
  const std::vector<ClusterGeo> gc_own = &C == &R ? std::vector<ClusterGeo>() : cluster_geo(C);
  const std::vector<ClusterGeo> &gc = &C == &R ? gr : gc_own;
  while (!frontier.empty()) {
    std::vector<std::vector<std::pair<int32_t, int32_t>>> next(nth);
    std::vector<std::vector<int32_t>> brow(nth), bcol(nth);
    std::vector<std::vector<int8_t>> blr(nth);
    const int64_t F = (int64_t)frontier.size();
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nth, F / 256 + 1));
    parallel_chunks(F, nt, [&](int tid, int64_t lo, int64_t hi) {
      // no reallocation while expanding (every pair refines into at most 4)
      next[tid].reserve(4 * (hi - lo));
      brow[tid].reserve(hi - lo);
      bcol[tid].reserve(hi - lo);
      blr[tid].reserve(hi - lo);
      for (int64_t k = lo; k < hi; ++k) {
        const int t = frontier[k].first, s = frontier[k].second;
        const int adm = admissibility(gr[t], gc[s], R.dim, eta);
        if (adm == 0) {
          for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i)
              next[tid].push_back({i ? R.son1[t] : R.son0[t], j ? C.son1[s] : C.son0[s]});
        } else {
          int8_t lr = 0;
          if (adm == 1) {
            // FMCA compares 32-bit unsigned products; the wrap is kept.
            const uint32_t lhs = uint32_t(q) * uint32_t(q);
            const uint32_t rhs = uint32_t(R.size[t]) * uint32_t(C.size[s]);
            lr = lhs < rhs;
          }
          brow[tid].push_back(t);
          bcol[tid].push_back(s);
          blr[tid].push_back(lr);
        }
      }
    });
    frontier.clear();
    for (int t = 0; t < nth; ++t) {
      frontier.insert(frontier.end(), next[t].begin(), next[t].end());
      pat.block_row.insert(pat.block_row.end(), brow[t].begin(), brow[t].end());
      pat.block_col.insert(pat.block_col.end(), bcol[t].begin(), bcol[t].end());
      pat.block_lr.insert(pat.block_lr.end(), blr[t].begin(), blr[t].end());
    }
  }
}

void finish_pattern(Pattern &pat, const Tree &R, const Tree &C, int q) {
  const int64_t nb = (int64_t)pat.block_row.size();
  pat.n_far = pat.n_near = 0;
  pat.near_entries = pat.far_entries = 0;
  // far field CSR by target cluster
  build_csr(
      nb, R.ncl,
      [&](int64_t b, auto put) {
        if (pat.block_lr[b]) put(pat.block_row[b], pat.block_col[b]);
      },
      pat.far_ptr, pat.far_src);
  pat.n_far = pat.far_ptr[R.ncl];
  parallel_for(R.ncl, 64, [&](int64_t c) {
    std::sort(pat.far_src.begin() + pat.far_ptr[c], pat.far_src.begin() + pat.far_ptr[c + 1]);
  });
  pat.far_entries = double(pat.n_far) * q * q;
  // near field CSR by target leaf; a dense block whose row cluster is not a
  // leaf is split over the leaves below it (their point ranges partition it).
  const int nl = (int)R.leaves.size();
  std::vector<int32_t> leaf_index(R.ncl, -1);
  for (int k = 0; k < nl; ++k) leaf_index[R.leaves[k]] = k;
  std::vector<int32_t> by_begin(R.leaves);
  std::sort(by_begin.begin(), by_begin.end(),
            [&](int a, int b) { return R.begin[a] < R.begin[b]; });
  std::vector<int64_t> begins(nl);
  for (int k = 0; k < nl; ++k) begins[k] = R.begin[by_begin[k]];
  // the leaves below each cluster: [leaf_lo, leaf_hi) in by_begin
  std::vector<int32_t> leaf_lo(R.ncl), leaf_hi(R.ncl);
  parallel_for(R.ncl, 256, [&](int64_t t) {
    leaf_lo[t] = int(std::lower_bound(begins.begin(), begins.end(), R.begin[t]) - begins.begin());
    leaf_hi[t] = int(std::lower_bound(begins.begin(), begins.end(), R.begin[t] + R.size[t]) -
                     begins.begin());
  });
  auto leaves_under = [&](int t, int &lo, int &hi) {
    lo = leaf_lo[t];
    hi = leaf_hi[t];
  };
  {
    // dense blocks and their entries (integer sums: exact in any order)
    std::vector<int64_t> cnt(num_workers(), 0), ent(num_workers(), 0);
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(num_workers(), nb / 65536 + 1));
    parallel_chunks(nb, nt, [&](int t, int64_t lo, int64_t hi) {
      int64_t c = 0, en = 0;  // local: the per-thread slots share cache lines
      for (int64_t b = lo; b < hi; ++b)
        if (!pat.block_lr[b]) {
          ++c;
          en += R.size[pat.block_row[b]] * C.size[pat.block_col[b]];
        }
      cnt[t] = c;
      ent[t] = en;
    });
    int64_t e = 0;
    for (int t = 0; t < nt; ++t) {
      pat.n_near += cnt[t];
      e += ent[t];
    }
    pat.near_entries = double(e);
  }
  build_csr(
      nb, nl,
      [&](int64_t b, auto put) {
        if (pat.block_lr[b]) return;
        int lo, hi;
        leaves_under(pat.block_row[b], lo, hi);
        for (int k = lo; k < hi; ++k) put(leaf_index[by_begin[k]], pat.block_col[b]);
      },
      pat.near_ptr, pat.near_src);
  parallel_for(nl, 16, [&](int64_t k) {
    std::sort(pat.near_src.begin() + pat.near_ptr[k], pat.near_src.begin() + pat.near_ptr[k + 1],
              [&](int a, int b) { return C.begin[a] < C.begin[b]; });
  });
  // Tasks and launch orders.  Coupling work = far blocks (plus the zero fill
  // of an empty list); leaf output work = near-field entries (plus the far
  // field V^T l_t).  Targets above three times the mean work are cut into
  // pieces of about that size.
  {
    double fsum = 0;
    int fcnt = 0;
    for (int c = 0; c < R.ncl; ++c)
      if (pat.far_ptr[c + 1] > pat.far_ptr[c]) {
        fsum += pat.far_ptr[c + 1] - pat.far_ptr[c];
        ++fcnt;
      }
    const int cap = std::max(16, (int)std::ceil(3.0 * fsum / std::max(1, fcnt)));
    std::vector<double> wt;
    std::vector<std::pair<int, int>> levels;
    for (int L = R.depth; L >= 0; --L) {
      const int first = (int)(pat.far_task.size() / 4);
      for (int c = R.level_ptr[L]; c < R.level_ptr[L + 1]; ++c) {
        const int lo = pat.far_ptr[c], hi = pat.far_ptr[c + 1], f = hi - lo;
        const int m = std::max(1, (f + cap - 1) / cap);
        if (m > 1) pat.far_split.push_back(c);
        for (int k = 0; k < m; ++k) {
          const int a = lo + (int)((int64_t)f * k / m), b = lo + (int)((int64_t)f * (k + 1) / m);
          pat.far_task.insert(pat.far_task.end(), {c, a, b, m > 1 ? kTaskSplit : 0});
          wt.push_back(0.05 + (b - a));
        }
      }
      levels.push_back({first, (int)(pat.far_task.size() / 4)});
    }
    pat.far_order = xcd_launch_order(wt, levels, kXcds);
  }
  {
    std::vector<double> wl(nl, 0.);
    for (int k = 0; k < nl; ++k)
      for (int e = pat.near_ptr[k]; e < pat.near_ptr[k + 1]; ++e)
        wl[k] += double(R.size[R.leaves[k]]) * C.size[pat.near_src[e]];
    double mean = 0;
    for (double w : wl) mean += w;
    mean /= std::max(1, nl);
    const double cap = std::max(3.0 * mean, 1.0);
    std::vector<double> wt;
    for (int k = 0; k < nl; ++k) {
      const int rows = (int)R.size[R.leaves[k]];
      const double wfar = double(rows) * q;  // V^T l_t
      const size_t first = pat.near_task.size();
      int a = pat.near_ptr[k];
      double acc = wfar;
      for (int e = pat.near_ptr[k]; e < pat.near_ptr[k + 1]; ++e) {
        const double w = double(rows) * C.size[pat.near_src[e]];
        if (acc + w > cap && e > a) {
          pat.near_task.insert(pat.near_task.end(), {k, a, e, 0});
          wt.push_back(acc);
          a = e;
          acc = 0;
        }
        acc += w;
      }
      pat.near_task.insert(pat.near_task.end(), {k, a, pat.near_ptr[k + 1], 0});
      wt.push_back(acc);
      const size_t ntasks = (pat.near_task.size() - first) / 4;
      pat.near_task[first + 3] |= kTaskFar;
      if (ntasks > 1) {
        pat.near_split.push_back(k);
        for (size_t i = first; i < pat.near_task.size(); i += 4) pat.near_task[i + 3] |= kTaskSplit;
      }
    }
    pat.near_order = xcd_launch_order(wt, {{0, (int)wt.size()}}, kXcds);
    // runs of contiguous sources (the lists are sorted by point range)
    pat.near_run.assign(pat.near_src.size(), 1);
    for (size_t t = 0; t < pat.near_task.size(); t += 4) {
      const int a = pat.near_task[t + 1], b = pat.near_task[t + 2];
      for (int e = b - 2; e >= a; --e) {
        const int s = pat.near_src[e], sn = pat.near_src[e + 1];
        if (C.begin[sn] == C.begin[s] + C.size[s]) pat.near_run[e] = pat.near_run[e + 1] + 1;
      }
    }
  }
}

std::vector<int32_t> xcd_launch_order(const std::vector<double> &work,
                                      const std::vector<std::pair<int, int>> &ranges, int nxcd) {
  std::vector<std::vector<int32_t>> lists(nxcd);
  for (const auto &rg : ranges) {
    double total = 0;
    for (int i = rg.first; i < rg.second; ++i) total += work[i];
    double acc = 0;
    for (int i = rg.first; i < rg.second; ++i) {
      // the piece that holds the midpoint of item i's work
      int x = total > 0 ? (int)((acc + 0.5 * work[i]) / total * nxcd)
                        : (int)((int64_t)(i - rg.first) * nxcd / (rg.second - rg.first));
      x = std::min(std::max(x, 0), nxcd - 1);
      lists[x].push_back(i);
      acc += work[i];
    }
  }
  size_t len = 0;
  for (const auto &l : lists) len = std::max(len, l.size());
  std::vector<int32_t> order(len * nxcd, -1);
  for (int x = 0; x < nxcd; ++x)
    for (size_t k = 0; k < lists[x].size(); ++k) order[k * nxcd + x] = lists[x][k];
  return order;
}

// ---------------------------------------------------------------------------
void compute_transfers(std::vector<double> &T, const Tree &tree, const Interpolation &ip) {
  const int d = tree.dim, P1 = ip.P1;
  T.assign((size_t)tree.ncl * d * P1 * P1, 0.);
  parallel_for(tree.ncl - 1, 64, [&](int64_t ci) {
    const int c = 1 + (int)ci;
    const int p = tree.parent[c];
    double vals[kMaxP1];
    for (int i = 0; i < d; ++i) {
      const double sc = tree.bbsize[(size_t)c * d + i];
      const double sp = tree.bbsize[(size_t)p * d + i];
      const double off = (tree.bbmin[(size_t)c * d + i] - tree.bbmin[(size_t)p * d + i]) / sp;
      double *Ti = T.data() + ((size_t)c * d + i) * P1 * P1;
      for (int b = 0; b < P1; ++b) {
        if (tree.size[c] == 0) continue;  // FMCA leaves E = 0 for empty sons
        const double x = ip.xi[b] * sc / sp + off;
        ip.lagrange_1d(x, vals);
        for (int a = 0; a < P1; ++a) Ti[a * P1 + b] = vals[a];
      }
    }
  });
}

void compute_transfers_dense(std::vector<double> &E, const Tree &tree, const Interpolation &ip) {
  const int d = tree.dim, q = ip.q;
  E.assign((size_t)tree.ncl * q * q, 0.);
  parallel_for(tree.ncl - 1, 16, [&](int64_t ci) {
    const int c = 1 + (int)ci;
    if (tree.size[c] == 0) return;  // FMCA leaves E = 0 for empty sons
    const int p = tree.parent[c];
    std::vector<double> Pv((size_t)q * q);  // Pv[j * q + a] = P^father_a(x(j))
    double x[kMaxDim];
    for (int j = 0; j < q; ++j) {
      for (int i = 0; i < d; ++i) {
        const double sp = tree.bbsize[(size_t)p * d + i];
        x[i] = ip.nodes[(size_t)d * j + i] * tree.bbsize[(size_t)c * d + i] / sp +
               (tree.bbmin[(size_t)c * d + i] - tree.bbmin[(size_t)p * d + i]) / sp;
      }
      ip.basis(x, &Pv[(size_t)q * j]);
    }
    double *Ec = E.data() + (size_t)c * q * q;
    for (int a = 0; a < q; ++a)
      for (int b = 0; b < q; ++b) {
        double acc = 0;
        for (int j = 0; j < q; ++j) acc += Pv[(size_t)q * j + a] * ip.invV[b + (size_t)q * j];
        Ec[(size_t)a * q + b] = acc;
      }
  });
}

Statistics statistics(const Pattern &pat, const Tree &R, const Tree &C, int q) {
  Statistics s{};
  double mem = 0;
  for (size_t b = 0; b < pat.block_row.size(); ++b) {
    if (pat.block_lr[b]) {
      s.low_rank_blocks += 1;
      mem += double(q) * q;
    } else {
      s.dense_blocks += 1;
      mem += double(R.size[pat.block_row[b]]) * C.size[pat.block_col[b]];
    }
  }
  s.rows = double(R.npts);
  s.cols = double(C.npts);
  s.nz_per_row = std::round(mem / double(R.npts));
  s.storage_gb = mem * sizeof(double) / 1e9;
  return s;
}

}  // namespace fmca_amd
