// Samplet trees, samplet transforms and the host side of the samplet
// compression: the pattern, the program of block operations over cluster
// pairs, their combination and the triplet output.  The kernel blocks the
// program starts from are evaluated on the GPU (samplets.hip).
#include "samplets.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>

namespace fmca_amd {

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();

// {a in N^dim : |a|_1 <= deg} in FMCA's graded-lexicographic order
// (MultiIndexSet<TotalDegree>, MultiIndexSet.h:87-111), packed 8 bits per axis.
std::vector<int32_t> total_degree_set(int dim, int deg) {
  std::vector<std::vector<int>> mi;
  std::vector<int> a(dim, 0);
  for (;;) {
    int sum = 0;
    for (int v : a) sum += v;
    if (sum <= deg) mi.push_back(a);
    int i = 0;
    while (i < dim && a[i] == deg) a[i++] = 0;
    if (i == dim) break;
    ++a[i];
  }
  std::sort(mi.begin(), mi.end(), [](const std::vector<int> &x, const std::vector<int> &y) {
    const int sx = std::accumulate(x.begin(), x.end(), 0);
    const int sy = std::accumulate(y.begin(), y.end(), 0);
    return sx != sy ? sx < sy : x < y;
  });
  std::vector<int32_t> out(mi.size());
  for (size_t k = 0; k < mi.size(); ++k) {
    int pk = 0;
    for (int i = 0; i < dim; ++i) pk |= mi[k][i] << (8 * i);
    out[k] = pk;
  }
  return out;
}

inline int expo(int32_t pk, int i) { return (pk >> (8 * i)) & 0xff; }

double binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  double r = 1;
  for (int j = 1; j <= k; ++j) r = r * double(n - k + j) / double(j);
  return std::round(r);
}

// Moments about a new centre: mom_new = S mom_old with
// S(j, i) = C(a_j, a_i) shift^(a_j - a_i), the identity for a shift below
// machine epsilon (SampletHelper::monomialMomentShifter, SampletHelper.h:38-63).
void shift_matrix(const SampletTree &S, const double *shift, double *out) {
  const int m = S.m, d = S.dim;
  double nrm = 0;
  for (int i = 0; i < d; ++i) nrm += shift[i] * shift[i];
  if (std::sqrt(nrm) < kEps) {
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) out[j + (size_t)m * i] = i == j ? 1.0 : 0.0;
    return;
  }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      double v = S.mcoef[j + (size_t)m * i];
      if (v != 0.0)
        for (int k = 0; k < d; ++k) {
          const int e = expo(S.mono[j], k) - expo(S.mono[i], k);
          if (e) v *= std::pow(shift[k], e);
        }
      out[j + (size_t)m * i] = v;
    }
}

// Householder QR of the rows x cols matrix A (column-major, rows > cols) in
// the convention of Eigen's HouseholderQR (makeHouseholder: beta = -sign(c0)
// ||x||, tau = (beta - c0) / beta, essential part x_tail / (c0 - beta), no
// reflection where the tail's squared norm is at most DBL_MIN).  On return
// A holds R in its upper triangle and the essential parts below it.
void householder_qr(int rows, int cols, double *A, double *tau) {
  for (int k = 0; k < cols; ++k) {
    double *col = A + (size_t)rows * k;
    const double c0 = col[k];
    double tail = 0;
    for (int i = k + 1; i < rows; ++i) tail += col[i] * col[i];
    double beta, t;
    if (tail <= DBL_MIN) {
      t = 0;
      beta = c0;
      for (int i = k + 1; i < rows; ++i) col[i] = 0;
    } else {
      beta = std::sqrt(c0 * c0 + tail);
      if (c0 >= 0) beta = -beta;
      const double den = c0 - beta;
      for (int i = k + 1; i < rows; ++i) col[i] = col[i] / den;
      t = (beta - c0) / beta;
    }
    col[k] = beta;
    tau[k] = t;
    if (t == 0) continue;
    for (int j = k + 1; j < cols; ++j) {
      double *cj = A + (size_t)rows * j;
      double w = cj[k];
      for (int i = k + 1; i < rows; ++i) w += col[i] * cj[i];
      cj[k] -= t * w;
      for (int i = k + 1; i < rows; ++i) cj[i] -= t * col[i] * w;
    }
  }
}

// Q = H_0 H_1 ... H_{cols-1} (rows x rows), applied to the identity from the
// last reflector on (HouseholderSequence::evalTo).
void householder_q(int rows, int cols, const double *A, const double *tau, double *Q) {
  for (int j = 0; j < rows; ++j)
    for (int i = 0; i < rows; ++i) Q[i + (size_t)rows * j] = i == j ? 1.0 : 0.0;
  for (int k = cols - 1; k >= 0; --k) {
    const double t = tau[k];
    if (t == 0) continue;
    const double *v = A + (size_t)rows * k;  // essential part in v[k+1 ..]
    for (int j = k; j < rows; ++j) {
      double *qj = Q + (size_t)rows * j;
      double w = qj[k];
      for (int i = k + 1; i < rows; ++i) w += v[i] * qj[i];
      qj[k] -= t * w;
      for (int i = k + 1; i < rows; ++i) qj[i] -= t * v[i] * w;
    }
  }
}

// C (m x n) = op(A) (m x k) * B (k x n), column-major, op = transpose if ta.
void gemm(bool ta, int m, int n, int k, const double *A, int lda, const double *B, int ldb,
          double *C, int ldc) {
  for (int j = 0; j < n; ++j) {
    double *c = C + (size_t)ldc * j;
    const double *b = B + (size_t)ldb * j;
    if (ta) {
      for (int i = 0; i < m; ++i) {
        const double *a = A + (size_t)lda * i;
        double s = 0;
        for (int l = 0; l < k; ++l) s += a[l] * b[l];
        c[i] = s;
      }
    } else {
      // column of C as a sum of columns of A (unit stride), in the same
      // order over l as the dot-product form
      for (int i = 0; i < m; ++i) c[i] = 0;
      for (int l = 0; l < k; ++l) {
        const double bl = b[l];
        const double *a = A + (size_t)lda * l;
        for (int i = 0; i < m; ++i) c[i] += a[i] * bl;
      }
    }
  }
}

// CompareCluster::compare (Admissibility.h:95-115): 1 = low rank (admissible).
bool admissible(const Tree &T, int a, int b, double eta) {
  const int d = T.dim;
  double s2 = 0;
  for (int i = 0; i < d; ++i) {
    const double v = 0.5 * (T.bbmin[(size_t)a * d + i] - T.bbmin[(size_t)b * d + i] +
                            T.bbmax[(size_t)a * d + i] - T.bbmax[(size_t)b * d + i]);
    s2 += v * v;
  }
  const double ra = cluster_radius(T, a), rb = cluster_radius(T, b);
  double dist = std::sqrt(s2) - ra - rb;
  dist = dist > 0 ? dist : 0;
  const double rad = ra > rb ? ra : rb;
  return !(rad > eta * dist);
}

}  // namespace

std::string build_samplet_tree(SampletTree &S, int dim, int64_t n, const double *P, int dtilde) {
  if (dim < 1 || dim > kMaxDim) return "dimension must be 1, 2 or 3";
  if (n < 1 || !P) return "no points";
  if (dtilde < 1) dtilde = 1;  // pySampletTree: dtilde > 0 ? dtilde : 1
  if (2 * (dtilde - 1) > kMaxP1 - 1) return "dtilde too large (multipole degree 2 (dtilde - 1) <= 9)";
  for (int64_t j = 0; j < n * dim; ++j)
    if (!std::isfinite(P[j])) return "non-finite point coordinate";
  S = SampletTree();
  S.dim = dim;
  S.dtilde = dtilde;
  S.n = n;
  S.ip.init(dim, 2 * (dtilde - 1), IS_TOTAL_DEGREE);
  const int q = S.ip.q;
  build_tree(S.tree, dim, n, P, q);
  const Tree &T = S.tree;
  const int ncl = T.ncl;
  S.pts.resize((size_t)dim * n);
  parallel_dynamic(n, 4096, [&](int64_t j) {
    const uint32_t p = T.perm[j];
    for (int i = 0; i < dim; ++i) S.pts[(size_t)i * n + j] = P[(size_t)dim * p + i];
  });
  S.mono = total_degree_set(dim, dtilde - 1);
  const int m = S.m = (int)S.mono.size();
  S.mcoef.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      double v = 1;
      for (int k = 0; k < dim; ++k) v *= binom(expo(S.mono[j], k), expo(S.mono[i], k));
      S.mcoef[j + (size_t)m * i] = v;
    }
  compute_transfers_dense(S.E, T, S.ip);

  // sizes of Q per cluster: leaf |t|, inner the sum of the sons' scaling functions
  S.qdim.assign(ncl, 0);
  S.nscal.assign(ncl, 0);
  S.nsamp.assign(ncl, 0);
  for (int L = T.depth; L >= 0; --L)
    for (int c = T.level_ptr[L]; c < T.level_ptr[L + 1]; ++c) {
      const int cols = T.is_leaf(c) ? (int)T.size[c] : S.nscal[T.son0[c]] + S.nscal[T.son1[c]];
      S.qdim[c] = cols;
      S.nscal[c] = m < cols ? m : cols;
      S.nsamp[c] = cols - S.nscal[c];
    }
  S.q_off.assign(ncl + 1, 0);
  S.v_off.assign(ncl + 1, 0);
  for (int c = 0; c < ncl; ++c) {
    S.q_off[c + 1] = S.q_off[c] + (int64_t)S.qdim[c] * S.qdim[c];
    S.v_off[c + 1] = S.v_off[c] + (int64_t)q * S.qdim[c];
  }
  S.Q.assign(S.q_off[ncl], 0.0);
  S.V.assign(S.v_off[ncl], 0.0);

  // computeSamplets (H2SampletTree.h:101-133): moments of the cluster's
  // functions about its box midpoint (leaf: monomials of P - mid; inner: the
  // sons' scaling-function moments shifted to this midpoint), QR of their
  // transpose; Q's first m columns are the scaling functions, whose moments
  // R^T go to the father.
  std::vector<std::vector<double>> mom(ncl);
  for (int L = T.depth; L >= 0; --L) {
    const int cb = T.level_ptr[L], ce = T.level_ptr[L + 1];
    parallel_dynamic(ce - cb, 4, [&](int64_t k) {
      const int c = cb + (int)k;
      const int cols = S.qdim[c];
      std::vector<double> M((size_t)m * cols);
      double mid[kMaxDim];
      for (int i = 0; i < dim; ++i)
        mid[i] = 0.5 * (T.bbmin[(size_t)c * dim + i] + T.bbmax[(size_t)c * dim + i]);
      if (T.is_leaf(c)) {
        for (int j = 0; j < cols; ++j) {
          const int64_t p = T.begin[c] + j;
          for (int a = 0; a < m; ++a) {
            double v = 1;
            for (int i = 0; i < dim; ++i) {
              const int e = expo(S.mono[a], i);
              if (e) v *= std::pow(S.pts[(size_t)i * n + p] - mid[i], e);
            }
            M[a + (size_t)m * j] = v;
          }
        }
      } else {
        int off = 0;
        std::vector<double> Sh((size_t)m * m);
        for (int s : {T.son0[c], T.son1[c]}) {
          double shift[kMaxDim];
          for (int i = 0; i < dim; ++i)
            shift[i] = 0.5 * (T.bbmin[(size_t)s * dim + i] - T.bbmin[(size_t)c * dim + i] +
                              T.bbmax[(size_t)s * dim + i] - T.bbmax[(size_t)c * dim + i]);
          shift_matrix(S, shift, Sh.data());
          const int sc = (int)(mom[s].size() / m);
          gemm(false, m, sc, m, Sh.data(), m, mom[s].data(), m, M.data() + (size_t)m * off, m);
          off += sc;
          std::vector<double>().swap(mom[s]);
        }
      }
      double *Qc = S.Q.data() + S.q_off[c];
      if (m < cols) {
        std::vector<double> A((size_t)cols * m), tau(m);
        for (int a = 0; a < m; ++a)
          for (int j = 0; j < cols; ++j) A[j + (size_t)cols * a] = M[a + (size_t)m * j];
        householder_qr(cols, m, A.data(), tau.data());
        householder_q(cols, m, A.data(), tau.data(), Qc);
        // moments of the scaling functions: (upper triangle of R)^T, m x m
        std::vector<double> R((size_t)m * m, 0.0);
        for (int j = 0; j < m; ++j)
          for (int i = 0; i <= j; ++i) R[j + (size_t)m * i] = A[i + (size_t)cols * j];
        mom[c].swap(R);
      } else {
        for (int j = 0; j < cols; ++j)
          for (int i = 0; i < cols; ++i) Qc[i + (size_t)cols * j] = i == j ? 1.0 : 0.0;
        mom[c].swap(M);
      }
    });
  }

  // sampletMapper (SampletMapper.h:19-31): level order, the root's scaling
  // functions first
  S.start.assign(ncl, 0);
  int64_t sum = 0;
  for (int c = 0; c < ncl; ++c) {
    S.start[c] = sum;
    sum += S.nsamp[c] + (c == 0 ? S.nscal[c] : 0);
  }

  // computeMultiscaleClusterBasis (H2SampletTreeBase.h:68-86)
  for (int L = T.depth; L >= 0; --L) {
    const int cb = T.level_ptr[L], ce = T.level_ptr[L + 1];
    parallel_dynamic(ce - cb, 4, [&](int64_t k) {
      const int c = cb + (int)k;
      const int cols = S.qdim[c];
      std::vector<double> B((size_t)q * cols);
      if (T.is_leaf(c)) {
        // NystromMoments::moment_matrix: basis at (P - bbmin) / bbsize
        double u[kMaxDim];
        for (int j = 0; j < cols; ++j) {
          const int64_t p = T.begin[c] + j;
          for (int i = 0; i < dim; ++i)
            u[i] = (S.pts[(size_t)i * n + p] - T.bbmin[(size_t)c * dim + i]) /
                   T.bbsize[(size_t)c * dim + i];
          S.ip.basis(u, B.data() + (size_t)q * j);
        }
      } else {
        int off = 0;
        for (int s : {T.son0[c], T.son1[c]}) {
          const double *Es = S.E.data() + (size_t)s * q * q;  // E[a * q + b]
          const double *Vs = S.V.data() + S.v_off[s];
          for (int j = 0; j < S.nscal[s]; ++j)
            for (int a = 0; a < q; ++a) {
              double v = 0;
              for (int b = 0; b < q; ++b) v += Es[(size_t)a * q + b] * Vs[b + (size_t)q * j];
              B[a + (size_t)q * (off + j)] = v;
            }
          off += S.nscal[s];
        }
      }
      gemm(false, q, cols, cols, B.data(), q, S.Q.data() + S.q_off[c], cols,
           S.V.data() + S.v_off[c], q);
    });
  }
  return std::string();
}

void samplet_transform(const SampletTree &S, const double *in, int64_t ld_in, int64_t k,
                       double *out, int64_t ld_out) {
  const Tree &T = S.tree;
  const int ncl = T.ncl;
  std::vector<std::vector<double>> scal(ncl);  // nscal x k scaling coefficients per cluster
  for (int L = T.depth; L >= 0; --L) {
    const int cb = T.level_ptr[L], ce = T.level_ptr[L + 1];
    parallel_dynamic(ce - cb, 4, [&](int64_t kk) {
      const int c = cb + (int)kk;
      const int cols = S.qdim[c];
      std::vector<double> X((size_t)cols * k);  // the cluster's functions' coefficients
      if (T.is_leaf(c)) {
        for (int64_t r = 0; r < k; ++r)
          for (int j = 0; j < cols; ++j) X[j + (size_t)cols * r] = in[T.begin[c] + j + ld_in * r];
      } else {
        int off = 0;
        for (int s : {T.son0[c], T.son1[c]}) {
          const int ns = S.nscal[s];
          for (int64_t r = 0; r < k; ++r)
            for (int j = 0; j < ns; ++j) X[off + j + (size_t)cols * r] = scal[s][j + (size_t)ns * r];
          off += ns;
          std::vector<double>().swap(scal[s]);
        }
      }
      const double *Qc = S.Q.data() + S.q_off[c];
      const int nsc = S.nscal[c], nsa = S.nsamp[c];
      std::vector<double> Y((size_t)cols * k);
      gemm(true, cols, (int)k, cols, Qc, cols, X.data(), cols, Y.data(), cols);
      const int64_t shift = c == 0 ? nsc : 0;
      for (int64_t r = 0; r < k; ++r)
        for (int j = 0; j < nsa; ++j) out[S.start[c] + shift + j + ld_out * r] = Y[nsc + j + (size_t)cols * r];
      std::vector<double> sc((size_t)nsc * k);
      for (int64_t r = 0; r < k; ++r)
        for (int j = 0; j < nsc; ++j) sc[j + (size_t)nsc * r] = Y[j + (size_t)cols * r];
      if (c == 0)
        for (int64_t r = 0; r < k; ++r)
          for (int j = 0; j < nsc; ++j) out[S.start[c] + j + ld_out * r] = sc[j + (size_t)nsc * r];
      scal[c].swap(sc);
    });
  }
}

void inverse_samplet_transform(const SampletTree &S, const double *in, int64_t ld_in, int64_t k,
                               double *out, int64_t ld_out) {
  const Tree &T = S.tree;
  const int ncl = T.ncl;
  std::vector<std::vector<double>> scal(ncl);  // scaling coefficients handed down
  for (int L = 0; L <= T.depth; ++L) {
    const int cb = T.level_ptr[L], ce = T.level_ptr[L + 1];
    parallel_dynamic(ce - cb, 4, [&](int64_t kk) {
      const int c = cb + (int)kk;
      const int cols = S.qdim[c], nsc = S.nscal[c], nsa = S.nsamp[c];
      std::vector<double> Y((size_t)cols * k);
      const int64_t s0 = S.start[c] + (c == 0 ? nsc : 0);
      for (int64_t r = 0; r < k; ++r) {
        for (int j = 0; j < nsc; ++j)
          Y[j + (size_t)cols * r] = c == 0 ? in[S.start[c] + j + ld_in * r] : scal[c][j + (size_t)nsc * r];
        for (int j = 0; j < nsa; ++j) Y[nsc + j + (size_t)cols * r] = in[s0 + j + ld_in * r];
      }
      std::vector<double>().swap(scal[c]);
      std::vector<double> X((size_t)cols * k);
      gemm(false, cols, (int)k, cols, S.Q.data() + S.q_off[c], cols, Y.data(), cols, X.data(), cols);
      if (T.is_leaf(c)) {
        for (int64_t r = 0; r < k; ++r)
          for (int j = 0; j < cols; ++j) out[T.begin[c] + j + ld_out * r] = X[j + (size_t)cols * r];
      } else {
        int off = 0;
        for (int s : {T.son0[c], T.son1[c]}) {
          const int ns = S.nscal[s];
          std::vector<double> v((size_t)ns * k);
          for (int64_t r = 0; r < k; ++r)
            for (int j = 0; j < ns; ++j) v[j + (size_t)ns * r] = X[off + j + (size_t)cols * r];
          scal[s].swap(v);
          off += ns;
        }
      }
    });
  }
}

std::vector<int32_t> samplet_levels(const SampletTree &S) {
  const Tree &T = S.tree;
  const int d = T.dim;
  auto diam = [&](int c) {
    double s = 0;
    for (int i = 0; i < d; ++i) s += T.bbsize[(size_t)c * d + i] * T.bbsize[(size_t)c * d + i];
    return std::sqrt(s);
  };
  const double geo = diam(0);
  std::vector<int32_t> out;
  out.reserve(S.n);
  for (int c = 0; c < T.ncl; ++c) {
    const int lvl = (int)std::round(-std::log(diam(c) / geo) / std::log(2.0));
    if (c == 0)
      for (int i = 0; i < S.nscal[c]; ++i) out.push_back(lvl);
    for (int i = 0; i < S.nsamp[c]; ++i) out.push_back(lvl);
  }
  return out;
}

namespace {

struct PairHash {
  size_t operator()(uint64_t k) const { return std::hash<uint64_t>()(k * 0x9E3779B97F4A7C15ull); }
};

inline uint64_t pair_key(int r, int c) { return ((uint64_t)(uint32_t)r << 32) | (uint32_t)c; }

struct ProgramBuilder {
  SampletProgram &P;
  const SampletTree &S;
  const Tree &T;
  double eta;
  std::unordered_map<uint64_t, int32_t, PairHash> pattern, rec, pat_nodes;

  int add_node(int r, int c, int op, std::initializer_list<int32_t> children) {
    const int id = (int)P.row.size();
    P.row.push_back(r);
    P.col.push_back(c);
    P.op.push_back(op);
    int h = 0;
    int a[4] = {-1, -1, -1, -1}, k = 0;
    for (int32_t ch : children) {
      a[k++] = ch;
      if (op != 0) h = std::max(h, P.height[ch] + 1);
    }
    P.height.push_back(h);
    for (int i = 0; i < 4; ++i) P.arg.push_back(a[i]);
    const int64_t sz = (int64_t)S.qdim[r] * S.qdim[c];
    if (op == 0) {
      P.off.push_back(P.job_elems);
      P.job_elems += sz;
    } else {
      P.off.push_back(P.comb_elems);
      P.comb_elems += sz;
    }
    P.max_height = std::max(P.max_height, h);
    return id;
  }

  int add_job(int kind, int r, int c) {
    const int j = (int)P.jobs.size();
    SampletJob job;
    job.kind = kind;
    job.r = r;
    job.c = c;
    job.pad = 0;
    job.off = P.job_elems;
    P.jobs.push_back(job);
    return add_node(r, c, 0, {j});
  }

  // recursivelyComputeBlock (samplet_matrix_compressor.h:330-388)
  int recursion(int r, int c) {
    const uint64_t key = pair_key(r, c);
    auto it = rec.find(key);
    if (it != rec.end()) return it->second;
    int id;
    const bool lr = T.is_leaf(r), lc = T.is_leaf(c);
    if (admissible(T, r, c, eta)) {
      id = add_job(1, r, c);
    } else if (lr && lc) {
      id = add_job(0, r, c);
    } else if (lr) {
      const int a = recursion(r, T.son0[c]), b = recursion(r, T.son1[c]);
      id = add_node(r, c, 1, {a, b});
    } else if (lc) {
      const int a = recursion(T.son0[r], c), b = recursion(T.son1[r], c);
      id = add_node(r, c, 2, {a, b});
    } else {
      const int a = recursion(T.son0[r], T.son0[c]), b = recursion(T.son0[r], T.son1[c]);
      const int e = recursion(T.son1[r], T.son0[c]), f = recursion(T.son1[r], T.son1[c]);
      id = add_node(r, c, 3, {a, b, e, f});
    }
    rec.emplace(key, id);
    return id;
  }

  // a stored pattern block if (r, c) is a pattern entry, else the recursion
  int pattern_or_recursion(int r, int c) {
    if (pattern.count(pair_key(r, c))) return pattern_node(r, c);
    return recursion(r, c);
  }

  // compress (samplet_matrix_compressor.h:73-186) for one pattern entry
  int pattern_node(int r, int c) {
    const uint64_t key = pair_key(r, c);
    auto it = pat_nodes.find(key);
    if (it != pat_nodes.end()) return it->second;
    int id;
    const bool lr = T.is_leaf(r), lc = T.is_leaf(c);
    if (lr && lc) {
      id = recursion(r, c);
    } else if (lc) {  // (noleaf, leaf): recycle from below
      const int a = pattern_or_recursion(T.son0[r], c), b = pattern_or_recursion(T.son1[r], c);
      id = add_node(r, c, 2, {a, b});
    } else {  // (*, noleaf): recycle from the right
      const int a = pattern_or_recursion(r, T.son0[c]), b = pattern_or_recursion(r, T.son1[c]);
      id = add_node(r, c, 1, {a, b});
    }
    pat_nodes.emplace(key, id);
    return id;
  }
};

}  // namespace

void plan_samplet_compression(SampletProgram &prog, const SampletTree &S, double eta) {
  prog = SampletProgram();
  const Tree &T = S.tree;
  const int ncl = T.ncl;
  // init: for every column cluster a depth-first search over the row
  // clusters that are not admissible to it; entries with id(col) >= id(row)
  std::vector<std::vector<std::pair<int32_t, int32_t>>> per_col(ncl);
  parallel_dynamic(ncl, 16, [&](int64_t k) {
    const int pc = (int)k;
    std::vector<int32_t> stack{0};
    auto &out = per_col[pc];
    while (!stack.empty()) {
      const int pr = stack.back();
      stack.pop_back();
      if (!T.is_leaf(pr))
        for (int s : {T.son0[pr], T.son1[pr]})
          if (!admissible(T, s, pc, eta)) stack.push_back(s);
      if (pc >= pr) out.emplace_back(pr, pc);
    }
  });
  // levels, deepest first; within a level by id = row + ncl * col, descending
  // (the LevelBuffer map's order)
  const int nlev = 2 * T.depth + 1;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> lev(nlev);
  for (int pc = 0; pc < ncl; ++pc)
    for (auto &e : per_col[pc]) lev[T.level[e.first] + T.level[e.second]].push_back(e);
  ProgramBuilder B{prog, S, T, eta, {}, {}, {}};
  size_t total = 0;
  for (auto &v : lev) total += v.size();
  B.pattern.reserve(total * 2);
  B.pat_nodes.reserve(total * 2);
  B.rec.reserve(total * 4);
  for (auto &v : lev)
    for (auto &e : v) B.pattern.emplace(pair_key(e.first, e.second), 1);
  for (int ll = nlev - 1; ll >= 0; --ll) {
    auto &v = lev[ll];
    std::sort(v.begin(), v.end(), [&](const std::pair<int32_t, int32_t> &a,
                                      const std::pair<int32_t, int32_t> &b) {
      const int64_t ia = a.first + (int64_t)ncl * a.second, ib = b.first + (int64_t)ncl * b.second;
      return ia > ib;
    });
    for (auto &e : v) {
      prog.pat_row.push_back(e.first);
      prog.pat_col.push_back(e.second);
      prog.pat_node.push_back(B.pattern_node(e.first, e.second));
    }
  }
}

void combine_samplet_blocks(const SampletProgram &prog, const SampletTree &S,
                            const double *job_blocks, std::vector<double> &comb) {
  const Tree &T = S.tree;
  comb.assign(prog.comb_elems, 0.0);
  auto block = [&](int node) -> const double * {
    return prog.op[node] == 0 ? job_blocks + prog.off[node] : comb.data() + prog.off[node];
  };
  std::vector<std::vector<int32_t>> by_height(prog.max_height + 1);
  for (size_t k = 0; k < prog.op.size(); ++k)
    if (prog.op[k] != 0) by_height[prog.height[k]].push_back((int32_t)k);
  for (int h = 1; h <= prog.max_height; ++h) {
    const auto &nodes = by_height[h];
    parallel_dynamic((int64_t)nodes.size(), 16, [&](int64_t k) {
      const int id = nodes[k];
      const int r = prog.row[id], c = prog.col[id], op = prog.op[id];
      const int dr = S.qdim[r], dc = S.qdim[c];
      const int32_t *a = prog.arg.data() + 4 * (size_t)id;
      std::vector<double> X((size_t)dr * dc, 0.0);
      if (op == 1) {  // columns from the sons of c
        int off = 0;
        for (int j = 0; j < 2; ++j) {
          const int s = j ? T.son1[c] : T.son0[c];
          const double *Bj = block(a[j]);
          for (int col = 0; col < S.nscal[s]; ++col)
            std::memcpy(&X[(size_t)dr * (off + col)], Bj + (size_t)dr * col, sizeof(double) * dr);
          off += S.nscal[s];
        }
        gemm(false, dr, dc, dc, X.data(), dr, S.Q.data() + S.q_off[c], dc,
             const_cast<double *>(block(id)), dr);
      } else if (op == 2) {  // rows from the sons of r
        int off = 0;
        for (int i = 0; i < 2; ++i) {
          const int s = i ? T.son1[r] : T.son0[r];
          const double *Bi = block(a[i]);
          const int ds = S.qdim[s];
          for (int col = 0; col < dc; ++col)
            for (int row = 0; row < S.nscal[s]; ++row)
              X[off + row + (size_t)dr * col] = Bi[row + (size_t)ds * col];
          off += S.nscal[s];
        }
        gemm(true, dr, dc, dr, S.Q.data() + S.q_off[r], dr, X.data(), dr,
             const_cast<double *>(block(id)), dr);
      } else {  // both: Q_r^T [B(r_i, c_j) top-left] Q_c
        int roff = 0;
        for (int i = 0; i < 2; ++i) {
          const int ri = i ? T.son1[r] : T.son0[r];
          int coff = 0;
          for (int j = 0; j < 2; ++j) {
            const int cj = j ? T.son1[c] : T.son0[c];
            const double *Bij = block(a[2 * i + j]);
            const int ds = S.qdim[ri];
            for (int col = 0; col < S.nscal[cj]; ++col)
              for (int row = 0; row < S.nscal[ri]; ++row)
                X[roff + row + (size_t)dr * (coff + col)] = Bij[row + (size_t)ds * col];
            coff += S.nscal[cj];
          }
          roff += S.nscal[ri];
        }
        std::vector<double> Y((size_t)dr * dc);
        gemm(false, dr, dc, dc, X.data(), dr, S.Q.data() + S.q_off[c], dc, Y.data(), dr);
        gemm(true, dr, dc, dr, S.Q.data() + S.q_off[r], dr, Y.data(), dr,
             const_cast<double *>(block(id)), dr);
      }
    });
  }
}

void samplet_triplets(const SampletProgram &prog, const SampletTree &S, const double *job_blocks,
                      const std::vector<double> &comb, double threshold, HostArray<int64_t> &rows,
                      HostArray<int64_t> &cols, HostArray<double> &vals) {
  const int64_t np = (int64_t)prog.pat_row.size();
  // storeBlock's window of pattern entry e: the samplet rows / columns, all
  // of the root's rows when the row cluster is the root
  struct Window {
    const double *B;
    int dr, nr, nc, r0, c0;
    int64_t srow, scol;
  };
  auto window = [&](int64_t e) {
    const int r = prog.pat_row[e], c = prog.pat_col[e], node = prog.pat_node[e];
    Window w;
    w.B = prog.op[node] == 0 ? job_blocks + prog.off[node] : comb.data() + prog.off[node];
    w.dr = S.qdim[r];
    if (r != 0 && c != 0) {
      w.nr = S.nsamp[r], w.nc = S.nsamp[c];
      w.r0 = w.dr - w.nr;
      w.c0 = S.qdim[c] - w.nc;
    } else if (c != 0) {
      w.nr = w.dr, w.nc = S.nsamp[c];
      w.r0 = 0;
      w.c0 = S.qdim[c] - w.nc;
    } else {
      w.nr = w.dr, w.nc = S.qdim[c];
      w.r0 = w.c0 = 0;
    }
    w.srow = S.start[r];
    w.scol = S.start[c];
    return w;
  };
  // the entries kept: upper triangle above the threshold, and the diagonal
  auto keep = [&](const Window &w, int j, int k, double v) {
    return (w.srow + j <= w.scol + k && std::abs(v) > threshold) || (w.srow == w.scol && j == k);
  };
  std::vector<int64_t> off(np + 1, 0);
  parallel_dynamic(np, 64, [&](int64_t e) {
    const Window w = window(e);
    int64_t cnt = 0;
    for (int k = 0; k < w.nc; ++k)
      for (int j = 0; j < w.nr; ++j)
        cnt += keep(w, j, k, w.B[(w.r0 + j) + (size_t)w.dr * (w.c0 + k)]) ? 1 : 0;
    off[e + 1] = cnt;
  });
  for (int64_t e = 0; e < np; ++e) off[e + 1] += off[e];
  const size_t total = (size_t)off[np];
  rows.resize_uninit(total);
  cols.resize_uninit(total);
  vals.resize_uninit(total);
  parallel_dynamic(np, 64, [&](int64_t e) {
    const Window w = window(e);
    int64_t t = off[e];
    for (int k = 0; k < w.nc; ++k)
      for (int j = 0; j < w.nr; ++j) {
        const double v = w.B[(w.r0 + j) + (size_t)w.dr * (w.c0 + k)];
        if (keep(w, j, k, v)) {
          rows.data()[t] = w.srow + j;
          cols.data()[t] = w.scol + k;
          vals.data()[t] = v;
          ++t;
        }
      }
  });
}

std::vector<int64_t> samplet_aposteriori_fast(const int64_t *rows, const int64_t *cols,
                                              const double *vals, size_t count, double thres) {
  std::vector<std::vector<int64_t>> buckets(17);
  double norms2[17] = {0};
  const double invlog10 = 1. / std::log(10.);
  for (size_t i = 0; i < count; ++i) {
    const double entry = std::abs(vals[i]);
    const double val = -std::floor(invlog10 * std::log(entry));
    // decade of the entry, the last bucket for 1e-16 and below (zero included)
    const int b = val < 0 ? 0 : (val >= 16 ? 16 : (int)val);
    buckets[b].push_back((int64_t)i);
    norms2[b] += entry * entry;
  }
  double fnorm2 = 0;
  for (int i = 16; i >= 0; --i) fnorm2 += norms2[i];
  double cut = 0;
  int cut_off = 17;
  for (int i = 16; i >= 0; --i) {
    cut += norms2[i];
    if (std::sqrt(cut / fnorm2) >= thres) break;
    --cut_off;
  }
  std::vector<int64_t> keep;
  for (int i = 0; i < cut_off; ++i) keep.insert(keep.end(), buckets[i].begin(), buckets[i].end());
  for (int i = cut_off; i < 17; ++i)
    for (int64_t t : buckets[i])
      if (rows[t] == cols[t]) keep.push_back(t);
  return keep;
}

}  // namespace fmca_amd
