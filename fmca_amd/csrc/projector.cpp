// H2MatrixProjector (FMCA/src/H2Matrix/H2MatrixProjector.h:21-75): the
// projection of a given matrix M onto the H2 format of a plan.  Near blocks
// are M's blocks; a far block (t, s) becomes the coupling matrix
//   S_ts = (V_t V_t^T)^-1 (V_t M_ts V_s^T) (V_s V_s^T)^-1,
// the Frobenius-orthogonal projection of M_ts onto span(V_t^T) x span(V_s^T)
// (FMCA: rinvR rinvR^T tmat^T cinvR cinvR^T with the dual cluster bases of
// compute_dual_cluster_bases_impl.h:25-57, whose invRT^T invRT is the
// inverse of the basis Gram matrix V V^T; computed here the same way, from
// QR factors).  Host code, FP64, tensor-product
// interpolation (invV = I): the projection is set-up work on matrices that fit
// in memory, as in FMCA; the result is streamed by the device actions as the
// assembled blocks are.
//
// V_c is the q x |c| basis matrix of cluster c in the planner's tensor order:
// at a leaf V_c[a][p] = prod_i l_{a_i}(u_i(p)) with u the point's unit
// coordinates in the leaf box; at an inner cluster the nested
// V_c = [E_0 V_son0, E_1 V_son1] with E = T_0 (x) T_1 (x) T_2 the transfer
// factors (compute_cluster_bases_impl.h:36-54).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <vector>

#include "plan.hpp"
#include "projector.hpp"

namespace fmca_amd {

namespace {

// Tensor index a -> per-axis indices.
void unpack_index(int a, int dim, int P1, int *ai) {
  for (int i = 0; i < dim; ++i) {
    ai[i] = a % P1;
    a /= P1;
  }
}

// The basis matrices V_c (q x |c|, column p = the c-local point p) of every
// cluster of a tree, nested from the leaves.
std::vector<std::vector<double>> basis_matrices(const Tree &T, const Interpolation &ip,
                                                const std::vector<double> &trans,
                                                const std::vector<double> &pts) {
  const int d = T.dim, P1 = ip.P1, q = ip.q;
  const size_t n = (size_t)T.npts;
  std::vector<std::vector<double>> V(T.ncl);
  // leaves
  parallel_ranges((int64_t)T.leaves.size(), [&](int64_t lo, int64_t hi) {
    std::vector<double> f((size_t)d * P1);
    int ai[kMaxDim];
    for (int64_t k = lo; k < hi; ++k) {
      const int c = T.leaves[k];
      const int64_t b0 = T.begin[c], sz = T.size[c];
      std::vector<double> &Vc = V[c];
      Vc.assign((size_t)q * sz, 0.0);
      for (int64_t p = 0; p < sz; ++p) {
        for (int i = 0; i < d; ++i) {
          const double u = (pts[i * n + b0 + p] - T.bbmin[(size_t)c * d + i]) /
                           T.bbsize[(size_t)c * d + i];
          ip.lagrange_1d(u, f.data() + (size_t)i * P1);
        }
        for (int a = 0; a < q; ++a) {
          unpack_index(a, d, P1, ai);
          double v = 1.0;
          for (int i = 0; i < d; ++i) v *= f[(size_t)i * P1 + ai[i]];
          Vc[(size_t)p * q + a] = v;
        }
      }
    }
  });
  // inner clusters, deepest level first: V_c = [E_0 V_son0 | E_1 V_son1]
  for (int L = T.depth - 1; L >= 0; --L) {
    const int cb = T.level_ptr[L], ce = T.level_ptr[L + 1];
    parallel_ranges(ce - cb, [&](int64_t lo, int64_t hi) {
      int ai[kMaxDim], bi[kMaxDim];
      std::vector<double> E((size_t)q * q);
      for (int64_t k = lo; k < hi; ++k) {
        const int c = cb + (int)k;
        if (T.is_leaf(c)) continue;
        std::vector<double> &Vc = V[c];
        Vc.assign((size_t)q * T.size[c], 0.0);
        for (int son : {T.son0[c], T.son1[c]}) {
          if (son < 0 || T.size[son] == 0) continue;
          // E[a][b] = prod_i T_son,i[a_i][b_i]
          for (int a = 0; a < q; ++a) {
            unpack_index(a, d, P1, ai);
            for (int b = 0; b < q; ++b) {
              unpack_index(b, d, P1, bi);
              double v = 1.0;
              for (int i = 0; i < d; ++i)
                v *= trans[(((size_t)son * d + i) * P1 + ai[i]) * P1 + bi[i]];
              E[(size_t)a * q + b] = v;
            }
          }
          const std::vector<double> &Vs = V[son];
          const int64_t off = T.begin[son] - T.begin[c];
          for (int64_t p = 0; p < T.size[son]; ++p) {
            const double *vs = Vs.data() + (size_t)p * q;
            double *vc = Vc.data() + (size_t)(off + p) * q;
            for (int a = 0; a < q; ++a) {
              double s = 0;
              const double *e = E.data() + (size_t)a * q;
              for (int b = 0; b < q; ++b) s += e[b] * vs[b];
              vc[a] = s;
            }
          }
        }
      }
    });
  }
  return V;
}

// R (q x q, upper triangular, row-major) of the Householder QR of the tall
// matrix A (m x q, column-major, m >= q; overwritten).  false if R has a zero
// diagonal entry (A of lower rank).
bool householder_r(std::vector<double> &A, int64_t m, int q, std::vector<double> &R) {
  R.assign((size_t)q * q, 0.0);
  for (int k = 0; k < q; ++k) {
    double *ak = A.data() + (size_t)k * m;
    double nrm = 0;
    for (int64_t i = k; i < m; ++i) nrm += ak[i] * ak[i];
    nrm = std::sqrt(nrm);
    if (nrm == 0) return false;
    const double alpha = ak[k] > 0 ? -nrm : nrm;
    // v = a_k - alpha e_k (stored in place), H = I - 2 v v^T / v^T v
    ak[k] -= alpha;
    double vv = 0;
    for (int64_t i = k; i < m; ++i) vv += ak[i] * ak[i];
    for (int j = k + 1; j < q; ++j) {
      double *aj = A.data() + (size_t)j * m;
      double dot = 0;
      for (int64_t i = k; i < m; ++i) dot += ak[i] * aj[i];
      const double f = 2 * dot / vv;
      for (int64_t i = k; i < m; ++i) aj[i] -= f * ak[i];
    }
    R[(size_t)k * q + k] = alpha;
    for (int j = k + 1; j < q; ++j) R[(size_t)k * q + j] = A[(size_t)j * m + k];
  }
  return true;
}

// Li = (R^T)^-1 for upper triangular R (row-major): lower triangular.
void inverse_rt(const std::vector<double> &R, int q, std::vector<double> &Li) {
  Li.assign((size_t)q * q, 0.0);
  // R^T is lower triangular with RT[i][j] = R[j][i]
  for (int j = 0; j < q; ++j) {
    Li[(size_t)j * q + j] = 1.0 / R[(size_t)j * q + j];
    for (int i = j + 1; i < q; ++i) {
      double t = 0;
      for (int k = j; k < i; ++k) t -= R[(size_t)k * q + i] * Li[(size_t)k * q + j];
      Li[(size_t)i * q + j] = t / R[(size_t)i * q + i];
    }
  }
}

// The dual cluster bases of compute_dual_cluster_bases_impl.h:25-57: per
// cluster a factor RT with RT RT^T = V V^T, from a QR factorisation of V^T at a
// leaf and of [E_0 RT_0, E_1 RT_1]^T at an inner cluster (V V^T = sum_i E_i
// V_i V_i^T E_i^T), so that V V^T is never formed; the inverse Gram matrix is
// Li^T Li with Li = RT^-1.  Returns (V V^T)^-1 of every cluster and FMCA's
// check: the largest ||V^T - V^T (V V^T)^-1 V V^T||_F / ||V^T||_F, with
// (V V^T)^-1 V = Li^T (Li V) as FMCA forms it.
bool gram_inverses(const Tree &T, const Interpolation &ip, const std::vector<double> &trans,
                   const std::vector<std::vector<double>> &V,
                   std::vector<std::vector<double>> &Ginv, double *check, std::string *err) {
  const int d = T.dim, P1 = ip.P1, q = ip.q;
  std::vector<std::vector<double>> RT(T.ncl);  // q x q, row-major, lower triangular
  Ginv.assign(T.ncl, {});
  std::atomic<bool> ok{true};
  std::vector<double> worst(T.ncl, 0.0);
  for (int L = T.depth; L >= 0; --L) {
    const int cb = T.level_ptr[L], ce = T.level_ptr[L + 1];
    parallel_ranges(ce - cb, [&](int64_t lo, int64_t hi) {
      int ai[kMaxDim], bi[kMaxDim];
      std::vector<double> A, R, E((size_t)q * q), Li;
      for (int64_t k = lo; k < hi; ++k) {
        const int c = cb + (int)k;
        const int64_t sz = T.size[c];
        if (sz == 0) continue;
        int64_t mrows;
        if (T.is_leaf(c)) {
          // A = V_c^T (|c| x q, column-major): A[a * sz + p] = V_c[a][p]
          mrows = sz;
          A.assign((size_t)sz * q, 0.0);
          for (int64_t p = 0; p < sz; ++p)
            for (int a = 0; a < q; ++a) A[(size_t)a * sz + p] = V[c][(size_t)p * q + a];
        } else {
          // A = [E_0 RT_0, E_1 RT_1]^T (2q x q): row block i holds (E_i RT_i)^T
          mrows = 0;
          for (int son : {T.son0[c], T.son1[c]})
            if (son >= 0 && T.size[son] > 0) mrows += q;
          A.assign((size_t)mrows * q, 0.0);
          int64_t r0 = 0;
          for (int son : {T.son0[c], T.son1[c]}) {
            if (son < 0 || T.size[son] == 0) continue;
            for (int a = 0; a < q; ++a) {
              unpack_index(a, d, P1, ai);
              for (int b = 0; b < q; ++b) {
                unpack_index(b, d, P1, bi);
                double v = 1.0;
                for (int i = 0; i < d; ++i)
                  v *= trans[(((size_t)son * d + i) * P1 + ai[i]) * P1 + bi[i]];
                E[(size_t)a * q + b] = v;
              }
            }
            const std::vector<double> &Rs = RT[son];
            // (E RT)[a][j] = sum_b E[a][b] RT[b][j]; A[r0 + j][a] = (E RT)[a][j]
            for (int a = 0; a < q; ++a)
              for (int j = 0; j < q; ++j) {
                double v = 0;
                for (int b = j; b < q; ++b) v += E[(size_t)a * q + b] * Rs[(size_t)b * q + j];
                A[(size_t)a * mrows + r0 + j] = v;
              }
            r0 += q;
          }
        }
        if (mrows < q || !householder_r(A, mrows, q, R)) {
          ok = false;
          continue;
        }
        std::vector<double> &rt = RT[c];
        rt.assign((size_t)q * q, 0.0);
        for (int i = 0; i < q; ++i)
          for (int j = 0; j <= i; ++j) rt[(size_t)i * q + j] = R[(size_t)j * q + i];
        inverse_rt(R, q, Li);
        // Ginv = Li^T Li
        std::vector<double> G((size_t)q * q, 0.0);
        for (int i = 0; i < q; ++i)
          for (int j = 0; j <= i; ++j) {
            double v = 0;
            for (int kk = i; kk < q; ++kk) v += Li[(size_t)kk * q + i] * Li[(size_t)kk * q + j];
            G[(size_t)i * q + j] = G[(size_t)j * q + i] = v;
          }
        // FMCA's check: Vdag = Li^T (Li V), err = ||V^T - V^T (Vdag V^T)|| / ||V^T||
        const std::vector<double> &Vc = V[c];
        std::vector<double> W((size_t)q * sz, 0.0), Vd((size_t)q * sz, 0.0),
            X((size_t)q * q, 0.0);
        for (int64_t p = 0; p < sz; ++p) {
          const double *v = Vc.data() + (size_t)p * q;
          double *w = W.data() + (size_t)p * q;
          for (int i = 0; i < q; ++i) {
            double t = 0;
            for (int kk = 0; kk <= i; ++kk) t += Li[(size_t)i * q + kk] * v[kk];
            w[i] = t;
          }
          double *vd = Vd.data() + (size_t)p * q;
          for (int i = 0; i < q; ++i) {
            double t = 0;
            for (int kk = i; kk < q; ++kk) t += Li[(size_t)kk * q + i] * w[kk];
            vd[i] = t;
          }
          for (int i = 0; i < q; ++i)
            for (int j = 0; j < q; ++j) X[(size_t)i * q + j] += vd[i] * v[j];
        }
        double num = 0, den = 0;
        for (int64_t p = 0; p < sz; ++p) {
          const double *v = Vc.data() + (size_t)p * q;
          for (int j = 0; j < q; ++j) {
            double t = 0;
            for (int i = 0; i < q; ++i) t += v[i] * X[(size_t)i * q + j];
            num += (v[j] - t) * (v[j] - t);
            den += v[j] * v[j];
          }
        }
        worst[c] = den > 0 ? std::sqrt(num / den) : 0.0;
        Ginv[c] = std::move(G);
      }
    });
    if (!ok) break;
  }
  if (!ok) {
    if (err) *err = "a cluster basis has lower rank than q (too few points)";
    return false;
  }
  if (check) *check = *std::max_element(worst.begin(), worst.end());
  return true;
}

}  // namespace

bool project_matrix(const Tree &rows, const Tree &cols, bool shared, const Interpolation &ip,
                    const std::vector<double> &trans_rows, const std::vector<double> &trans_cols,
                    const std::vector<double> &pts_rows, const std::vector<double> &pts_cols,
                    const Pattern &pat, const double *M, int64_t ld, ProjectedMatrix &out,
                    std::string *err) {
  if (ip.scheme != IS_TENSOR) {
    if (err) *err = "the projector takes the tensor-product interpolator";
    return false;
  }
  const int q = ip.q;
  const std::vector<std::vector<double>> Vr = basis_matrices(rows, ip, trans_rows, pts_rows);
  const std::vector<std::vector<double>> Vc_storage =
      shared ? std::vector<std::vector<double>>() : basis_matrices(cols, ip, trans_cols, pts_cols);
  const std::vector<std::vector<double>> &Vc = shared ? Vr : Vc_storage;
  std::vector<std::vector<double>> Gr, Gc_storage;
  if (!gram_inverses(rows, ip, trans_rows, Vr, Gr, &out.dual_check_rows, err)) return false;
  if (!shared && !gram_inverses(cols, ip, trans_cols, Vc, Gc_storage, &out.dual_check_cols, err))
    return false;
  if (shared) out.dual_check_cols = out.dual_check_rows;
  const std::vector<std::vector<double>> &Gc = shared ? Gr : Gc_storage;
  // far blocks in far-list order (target-major CSR)
  const int nrow_cl = rows.ncl;
  std::vector<int32_t> far_tgt(pat.far_src.size());
  for (int t = 0; t < nrow_cl; ++t)
    for (int b = pat.far_ptr[t]; b < pat.far_ptr[t + 1]; ++b) far_tgt[b] = t;
  const int64_t nfar = (int64_t)pat.far_src.size();
  out.far.assign((size_t)nfar * q * q, 0.0);
  parallel_ranges(nfar, [&](int64_t lo, int64_t hi) {
    std::vector<double> W, Tm((size_t)q * q), X((size_t)q * q);
    for (int64_t b = lo; b < hi; ++b) {
      const int t = far_tgt[b], s = pat.far_src[b];
      const int64_t bt = rows.begin[t], nt = rows.size[t];
      const int64_t bs = cols.begin[s], ns = cols.size[s];
      // W = M_ts V_s^T (nt x q), M in natural order
      W.assign((size_t)nt * q, 0.0);
      const std::vector<double> &Vs = Vc[s];
      for (int64_t j = 0; j < ns; ++j) {
        const double *mcol = M + (size_t)ld * cols.perm[bs + j];
        const double *vs = Vs.data() + (size_t)j * q;
        for (int64_t i = 0; i < nt; ++i) {
          const double m = mcol[rows.perm[bt + i]];
          if (m == 0) continue;
          double *w = W.data() + (size_t)i * q;
          for (int a = 0; a < q; ++a) w[a] += m * vs[a];
        }
      }
      // T = V_t W (q x q), T[a][c]
      std::fill(Tm.begin(), Tm.end(), 0.0);
      const std::vector<double> &Vt = Vr[t];
      for (int64_t i = 0; i < nt; ++i) {
        const double *vt = Vt.data() + (size_t)i * q;
        const double *w = W.data() + (size_t)i * q;
        for (int a = 0; a < q; ++a) {
          const double va = vt[a];
          double *tr = Tm.data() + (size_t)a * q;
          for (int c = 0; c < q; ++c) tr[c] += va * w[c];
        }
      }
      // S = Ginv_t T Ginv_s
      const std::vector<double> &A = Gr[t], &B = Gc[s];
      for (int a = 0; a < q; ++a)
        for (int c = 0; c < q; ++c) {
          double v = 0;
          for (int k = 0; k < q; ++k) v += A[(size_t)a * q + k] * Tm[(size_t)k * q + c];
          X[(size_t)a * q + c] = v;
        }
      double *S = out.far.data() + (size_t)b * q * q;
      for (int a = 0; a < q; ++a)
        for (int c = 0; c < q; ++c) {
          double v = 0;
          for (int k = 0; k < q; ++k) v += X[(size_t)a * q + k] * B[(size_t)k * q + c];
          S[(size_t)a * q + c] = v;
        }
    }
  });
  // near blocks in near-list order: M_ts with rows / columns in cluster order,
  // column-major (|t| x |s|)
  const int nl = (int)rows.leaves.size();
  out.near_off.assign(pat.near_src.size() + 1, 0);
  for (int k = 0; k < nl; ++k)
    for (int e = pat.near_ptr[k]; e < pat.near_ptr[k + 1]; ++e)
      out.near_off[e + 1] = out.near_off[e] + rows.size[rows.leaves[k]] * cols.size[pat.near_src[e]];
  out.near.assign((size_t)out.near_off.back(), 0.0);
  parallel_ranges(nl, [&](int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
      const int t = rows.leaves[k];
      const int64_t bt = rows.begin[t], nt = rows.size[t];
      for (int e = pat.near_ptr[k]; e < pat.near_ptr[k + 1]; ++e) {
        const int s = pat.near_src[e];
        const int64_t bs = cols.begin[s], ns = cols.size[s];
        double *dst = out.near.data() + out.near_off[e];
        for (int64_t j = 0; j < ns; ++j) {
          const double *mcol = M + (size_t)ld * cols.perm[bs + j];
          for (int64_t i = 0; i < nt; ++i) dst[(size_t)j * nt + i] = mcol[rows.perm[bt + i]];
        }
      }
    }
  });
  return true;
}

}  // namespace fmca_amd
