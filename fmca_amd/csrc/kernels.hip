// HIP kernels of the matrix-free H2 action for gfx950 (MI355X).
//
// Pipeline of one action y = K(rows, cols) x (FMCA H2kernelActor::action,
// FMCA/src/H2Matrix/H2kernelActor.h:38-80), all on one stream:
//   1. permute_in        rhs -> cluster order, point-major (rhs fastest)
//   2. upward_leaf       leaf moments  m_t = V_t x_t, V generated from the points
//   3. upward_transfer   per level, m_t = sum_sons (T0 (x) T1 (x) T2) m_son
//   4. coupling          per target cluster, l_t = sum_far K(xi_t, xi_s) m_s
//   5. downward_transfer per level, l_son += (T0 (x) T1 (x) T2)^T l_t
//   6. leaf_output       per target leaf, y = V_t^T l_t + sum_near K(P_t, P_s) x_s,
//                        written straight to the caller's layout
// Kernel entries are generated in registers and never stored.  Every output
// element is owned by exactly one workgroup (gather by target), so no atomics
// are needed and results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device.hpp"
#include "device_math.hpp"

namespace fmca_amd {

namespace {

using namespace dev;

// x[0, count) = 0 (a fill kernel, not a memset, so that captured graphs
// replay it like any other launch)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_range(int64_t count, T *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = T(0);
}

// ---------------------------------------------------------------------------
// 1. rhs (column-major, natural order) -> xs (cluster order, point-major).
// A workgroup moves kPermRows natural rows in column chunks of kPermCols
// through LDS: the reads run down the columns (consecutive rows, coalesced)
// and the writes along each row at its cluster position pinv[i]
// (consecutive columns, contiguous).
constexpr int kPermRows = 64, kPermCols = 64;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute_in(const double *__restrict__ rhs,
                                                       int64_t ld, int nrhs, int64_t n,
                                                       const uint32_t *__restrict__ pinv,
                                                       T *__restrict__ xs) {
  __shared__ T S[kPermRows * (kPermCols + 1)];
  const int64_t i0 = (int64_t)blockIdx.x * kPermRows;
  const int rows = (int)min<int64_t>(kPermRows, n - i0);
  for (int c0 = 0; c0 < nrhs; c0 += kPermCols) {
    const int cols = min(kPermCols, nrhs - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < kPermRows * cols; e += blockDim.x) {
      const int ii = e % kPermRows, r = e / kPermRows;
      if (ii < rows) S[ii * (kPermCols + 1) + r] = (T)rhs[i0 + ii + (int64_t)(c0 + r) * ld];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kPermRows * cols; e += blockDim.x) {
      const int ii = e / cols, r = e % cols;
      if (ii < rows) {
        const int64_t pos = pinv ? (int64_t)pinv[i0 + ii] : i0 + ii;
        xs[pos * nrhs + c0 + r] = S[ii * (kPermCols + 1) + r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Zero fill of the outputs that split targets accumulate into (a kernel, not
// a memset, so that the action can be captured in a graph): the moments
// tlhs[t] of the listed clusters, or the output rows of the listed leaves.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_clusters(const int32_t *__restrict__ list,
                                                          int64_t per_cluster,
                                                          T *__restrict__ tlhs) {
  const int64_t base = (int64_t)list[blockIdx.x] * per_cluster;
  for (int64_t e = threadIdx.x; e < per_cluster; e += blockDim.x) tlhs[base + e] = T(0);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_leaf_rows(DevTreeView<T> rows,
                                                           const int32_t *__restrict__ list,
                                                           int nrhs, double *__restrict__ out,
                                                           int64_t ldo, int natural) {
  const int c = rows.leaves[list[blockIdx.x]];
  const int b0 = rows.begin[c], sz = rows.size[c];
  for (int e = threadIdx.x; e < sz * nrhs; e += blockDim.x) {
    const int i = e % sz, r = e / sz;
    const int64_t row = natural ? (int64_t)rows.perm[b0 + i] : (int64_t)(b0 + i);
    out[row + (int64_t)r * ldo] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// 2. leaf moments: mom[t][alpha][r] = sum_p V_t[alpha][p] xs[p][r]
// One workgroup per leaf; the 1-D factors (Lagrange or Legendre) of 64 points
// at a time are staged in LDS and the basis value, a product of D factors, is
// formed on the fly.  G groups of kBlock / G threads split the NR columns, so
// that q <= kBlock / G basis functions keep every lane busy.
template <typename T, int D, int NR, int G>
__global__ __launch_bounds__(kBlock) void k_upward_leaf(DevTreeView<T> tr, InterpParams<T> ip,
                                                        const T *__restrict__ xs, int nrhs,
                                                        T *__restrict__ mom) {
  constexpr int A = kBlock / G;  // basis functions per pass
  constexpr int NRG = NR / G;    // columns per thread
  static_assert(NR % G == 0 && A % 64 == 0, "groups of whole waves, whole columns");
  extern __shared__ unsigned char smem_raw[];
  T *F = reinterpret_cast<T *>(smem_raw);  // [D][P1][64]
  T *X = F + D * ip.P1 * 64;               // [64][NR]
  const int c = tr.leaves[blockIdx.x];
  const int b0 = tr.begin[c], sz = tr.size[c];
  const int P1 = ip.P1, q = ip.q;
  const int al = threadIdx.x % A, g0 = (threadIdx.x / A) * NRG;
  for (int rc = 0; rc < nrhs; rc += NR) {
    const int nr = min(NR, nrhs - rc);
    for (int abase = 0; abase < q; abase += A) {
      const int alpha = abase + al;
      int a[D];
      {
        const int pk = basis_index<T, D>(alpha < q ? alpha : 0, ip);
#pragma unroll
        for (int i = 0; i < D; ++i) a[i] = (pk >> (8 * i)) & 0xff;
      }
      T acc[NRG];
#pragma unroll
      for (int r = 0; r < NRG; ++r) acc[r] = 0;
      for (int p0 = 0; p0 < sz; p0 += 64) {
        const int np = min(64, sz - p0);
        __syncthreads();
        for (int e = threadIdx.x; e < D * 64; e += kBlock) {
          const int i = e / 64, jj = e % 64;
          if (jj < np)
            basis_1d<T>(tr.unit[(int64_t)i * tr.npts + b0 + p0 + jj], ip, F + (i * P1) * 64 + jj,
                        64);
        }
        for (int e = threadIdx.x; e < 64 * NR; e += kBlock) {
          const int jj = e / NR, r = e % NR;
          X[e] = (jj < np && r < nr) ? xs[(int64_t)(b0 + p0 + jj) * nrhs + rc + r] : T(0);
        }
        __syncthreads();
        if (alpha < q) {
          for (int jj = 0; jj < np; ++jj) {
            T v = F[(0 * P1 + a[0]) * 64 + jj];
#pragma unroll
            for (int i = 1; i < D; ++i) v *= F[(i * P1 + a[i]) * 64 + jj];
#pragma unroll
            for (int r = 0; r < NRG; ++r) acc[r] += v * X[jj * NR + g0 + r];
          }
        }
      }
      if (alpha < q) {
#pragma unroll
        for (int r = 0; r < NRG; ++r)
          if (g0 + r < nr) mom[((int64_t)c * q + alpha) * nrhs + rc + g0 + r] = acc[r];
      }
    }
  }
}

// Mode-i product with a P1 x P1 matrix on a [q][nr] tensor held in LDS:
// out[lin with a_i] = sum_b M[a_i][b] in[lin with b]   (transpose: M[b][a_i]).
template <typename T, int NRC>
__device__ __forceinline__ void mode_product(const T *__restrict__ in, T *__restrict__ out,
                                             const T *__restrict__ M,
                                             const int *__restrict__ aidx, int dimi, int P1,
                                             int q, int stride, bool transpose) {
  for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
    const int lin = e / NRC, r = e % NRC;
    const int ai = (aidx[lin] >> (8 * dimi)) & 0xff;
    const int base = lin - ai * stride;
    T s = 0;
    for (int b = 0; b < P1; ++b) {
      const T m = transpose ? M[b * P1 + ai] : M[ai * P1 + b];
      s += m * in[(base + b * stride) * NRC + r];
    }
    out[e] = s;
  }
}

// aidx[lin] = a0 | a1 << 8 | a2 << 16 for the moment index lin = a0 + P1 a1 + P1^2 a2
template <int D>
__device__ __forceinline__ void fill_moment_index(int *aidx, int q, int P1) {
  for (int e = threadIdx.x; e < q; e += blockDim.x) {
    int rem = e, pk = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      pk |= (rem % P1) << (8 * d);
      rem /= P1;
    }
    aidx[e] = pk;
  }
}

// ---------------------------------------------------------------------------
// 3. upward transfer for the clusters [cbeg, cend) of one level:
//    mom[t] = sum_{s in sons(t)} (T_s0 (x) ... ) mom[s]
// NRC rhs columns per pass (compile time), the moments staged [q][NRC] in LDS
// and the Kronecker factors applied as D mode products.
template <typename T, int D, int NRC>
__global__ __launch_bounds__(kBlock) void k_upward_transfer(DevTreeView<T> tr, InterpParams<T> ip,
                                                            int cbeg, int nrhs,
                                                            T *__restrict__ mom) {
  extern __shared__ unsigned char smem_raw[];
  const int P1 = ip.P1, q = ip.q;
  const int c = cbeg + blockIdx.x;
  if (tr.son0[c] < 0) return;
  int *aidx = reinterpret_cast<int *>(smem_raw);
  T *A = reinterpret_cast<T *>(aidx + ((q + 3) & ~3));
  T *B = A + q * NRC;
  T *O = B + q * NRC;
  T *Tm = O + q * NRC;  // D x P1 x P1
  fill_moment_index<D>(aidx, q, P1);
  for (int rc = blockIdx.y * NRC; rc < nrhs; rc += gridDim.y * NRC) {
    const int nr = min(NRC, nrhs - rc);
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) O[e] = 0;
    for (int si = 0; si < 2; ++si) {
      const int s = si ? tr.son1[c] : tr.son0[c];
      __syncthreads();
      for (int e = threadIdx.x; e < D * P1 * P1; e += blockDim.x)
        Tm[e] = tr.transfer[(int64_t)s * D * P1 * P1 + e];
      for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
        const int lin = e / NRC, r = e % NRC;
        A[e] = r < nr ? mom[((int64_t)s * q + lin) * nrhs + rc + r] : T(0);
      }
      __syncthreads();
      T *cur = A, *nxt = B;
      int stride = 1;
#pragma unroll
      for (int i = 0; i < D; ++i) {
        mode_product<T, NRC>(cur, nxt, Tm + i * P1 * P1, aidx, i, P1, q, stride, false);
        __syncthreads();
        T *tmp = cur;
        cur = nxt;
        nxt = tmp;
        stride *= P1;
      }
      for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) O[e] += cur[e];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int lin = e / NRC, r = e % NRC;
      if (r < nr) mom[((int64_t)c * q + lin) * nrhs + rc + r] = O[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 4. coupling: tlhs[t] = sum_{s in far(t)} K(xi_t, xi_s) mom[s]
// (unsymmetricNystromEvaluator.h:42-58 with invV = I, H2kernelActor.h:64-66).
// One workgroup per target cluster; thread <-> target interpolation node.
// GEO: geodesic entries (KernelParams::geodesic), one source node at a time.
template <typename T, int D, int KT, int NR, bool GEO = false>
__global__ __launch_bounds__(kBlock) void k_coupling(DevTreeView<T> rows, DevTreeView<T> cols,
                                                     DevPatternView pat, InterpParams<T> ip,
                                                     KernelParams<T> kp, int nrhs,
                                                     const T *__restrict__ mom,
                                                     T *__restrict__ tlhs) {
  extern __shared__ unsigned char smem_raw[];
  const int P1 = ip.P1, q = ip.q;
  // source nodes: [D][kMaxP1] per-axis values (tensor) or [D][q] (total degree)
  const int ys = ip.td ? q : kMaxP1;
  T *Y = reinterpret_cast<T *>(smem_raw);
  T *Mv = Y + D * ys;  // [q][NR] source moments
  const int task = pat.far_order[blockIdx.x];
  if (task < 0) return;
  const int t = pat.far_task[4 * task], f0 = pat.far_task[4 * task + 1],
            f1 = pat.far_task[4 * task + 2];
  const bool split = (pat.far_task[4 * task + 3] & kTaskSplit) != 0;
  for (int rc = 0; rc < nrhs; rc += NR) {
    const int nr = min(NR, nrhs - rc);
    for (int abase = 0; abase < q; abase += blockDim.x) {
      const int alpha = abase + threadIdx.x;
      T x[D];
      {
        const int al = alpha < q ? alpha : 0;
        const int pk = basis_index<T, D>(al, ip);
#pragma unroll
        for (int i = 0; i < D; ++i) {
          const T u = ip.td ? ip.unodes[i * q + al] : ip.xi[(pk >> (8 * i)) & 0xff];
          x[i] = u * rows.bbsize[t * D + i] + rows.bbmin[t * D + i];
        }
      }
      T acc[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) acc[r] = 0;
      for (int f = f0; f < f1; ++f) {
        const int s = pat.far_src[f];
        __syncthreads();
        if (ip.td) {
          for (int e = threadIdx.x; e < D * q; e += blockDim.x) {
            const int i = e / q;
            Y[e] = ip.unodes[e] * cols.bbsize[s * D + i] + cols.bbmin[s * D + i];
          }
        } else {
          for (int e = threadIdx.x; e < D * P1; e += blockDim.x) {
            const int i = e / P1, b = e % P1;
            Y[i * kMaxP1 + b] = ip.xi[b] * cols.bbsize[s * D + i] + cols.bbmin[s * D + i];
          }
        }
        for (int e = threadIdx.x; e < q * NR; e += blockDim.x) {
          const int beta = e / NR, r = e % NR;
          Mv[e] = r < nr ? mom[((int64_t)s * q + beta) * nrhs + rc + r] : T(0);
        }
        __syncthreads();
        if constexpr (GEO) {
          // source node beta: axis i at Y[i * ys + b_i], b_0 fastest in the
          // tensor case
          int b[D];
#pragma unroll
          for (int i = 0; i < D; ++i) b[i] = 0;
          for (int beta = 0; alpha < q && beta < q; ++beta) {
            T y[D];
#pragma unroll
            for (int i = 0; i < D; ++i) y[i] = Y[i * ys + (ip.td ? beta : b[i])];
            const T kv = kernel_r2<KT, T>(geodesic_r2<T, D>(x, y, 1), kp);
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += kv * Mv[beta * NR + r];
#pragma unroll
            for (int i = 0; i < D; ++i) {
              if (++b[i] < P1) break;
              b[i] = 0;
            }
          }
        } else if (alpha < q && ip.td) {
          // total degree: nodes without tensor structure
          for (int beta = 0; beta < q; ++beta) {
            T r2 = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) {
              const T di = x[i] - Y[i * q + beta];
              r2 = fmat(di, di, r2);
            }
            const T kv = kernel_r2<KT, T>(r2, kp);
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += kv * Mv[beta * NR + r];
          }
        } else if (alpha < q) {
          // tensor loop over the source nodes: the squared distance is
          // accumulated dimension by dimension, one fma per entry innermost
          int beta = 0;
          if constexpr (D == 1) {
            for (int b0 = 0; b0 < P1; ++b0, ++beta) {
              const T d0 = x[0] - Y[b0];
              const T kv = kernel_r2<KT, T>(d0 * d0, kp);
#pragma unroll
              for (int r = 0; r < NR; ++r) acc[r] += kv * Mv[beta * NR + r];
            }
          } else if constexpr (D == 2) {
            for (int b1 = 0; b1 < P1; ++b1) {
              const T d1 = x[1] - Y[kMaxP1 + b1];
              const T t1 = d1 * d1;
              for (int b0 = 0; b0 < P1; ++b0, ++beta) {
                const T d0 = x[0] - Y[b0];
                const T kv = kernel_r2<KT, T>(fmat(d0, d0, t1), kp);
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] += kv * Mv[beta * NR + r];
              }
            }
          } else {
            for (int b2 = 0; b2 < P1; ++b2) {
              const T d2 = x[2] - Y[2 * kMaxP1 + b2];
              const T t2 = d2 * d2;
              for (int b1 = 0; b1 < P1; ++b1) {
                const T d1 = x[1] - Y[kMaxP1 + b1];
                const T t1 = fmat(d1, d1, t2);
                for (int b0 = 0; b0 < P1; ++b0, ++beta) {
                  const T d0 = x[0] - Y[b0];
                  const T kv = kernel_r2<KT, T>(fmat(d0, d0, t1), kp);
#pragma unroll
                  for (int r = 0; r < NR; ++r) acc[r] += kv * Mv[beta * NR + r];
                }
              }
            }
          }
        }
      }
      if (alpha < q) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (r < nr) {
            T *p = tlhs + ((int64_t)t * q + alpha) * nrhs + rc + r;
            if (split)
              atomic_accumulate(p, acc[r]);
            else
              *p = acc[r];
          }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 5. downward transfer for the clusters [cbeg, cend) of one level >= 1:
//    tlhs[c] += (T_c0 (x) ...)^T tlhs[parent(c)]
template <typename T, int D, int NRC>
__global__ __launch_bounds__(kBlock) void k_downward_transfer(DevTreeView<T> tr,
                                                              InterpParams<T> ip, int cbeg,
                                                              int nrhs, T *__restrict__ tlhs) {
  extern __shared__ unsigned char smem_raw[];
  const int P1 = ip.P1, q = ip.q;
  const int c = cbeg + blockIdx.x;
  const int p = tr.parent[c];
  int *aidx = reinterpret_cast<int *>(smem_raw);
  T *A = reinterpret_cast<T *>(aidx + ((q + 3) & ~3));
  T *B = A + q * NRC;
  T *Tm = B + q * NRC;
  fill_moment_index<D>(aidx, q, P1);
  for (int e = threadIdx.x; e < D * P1 * P1; e += blockDim.x)
    Tm[e] = tr.transfer[(int64_t)c * D * P1 * P1 + e];
  for (int rc = blockIdx.y * NRC; rc < nrhs; rc += gridDim.y * NRC) {
    const int nr = min(NRC, nrhs - rc);
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int lin = e / NRC, r = e % NRC;
      A[e] = r < nr ? tlhs[((int64_t)p * q + lin) * nrhs + rc + r] : T(0);
    }
    __syncthreads();
    T *cur = A, *nxt = B;
    int stride = 1;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      mode_product<T, NRC>(cur, nxt, Tm + i * P1 * P1, aidx, i, P1, q, stride, true);
      __syncthreads();
      T *tmp = cur;
      cur = nxt;
      nxt = tmp;
      stride *= P1;
    }
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int lin = e / NRC, r = e % NRC;
      if (r < nr) tlhs[((int64_t)c * q + lin) * nrhs + rc + r] += cur[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Total-degree scheme (TotalDegreeInterpolator).  The transfers are dense
// q x q matrices per son, E_c[a * q + b] (compute_cluster_bases_impl.h:36-51),
// and the coupling matrices invV K(xi_t, xi_s) invV^T
// (unsymmetricNystromEvaluator.h:56) are applied as
//   tlhs[t] = invV (sum_s K(xi_t, xi_s) (invV^T mom[s])),
// i.e. the source moments are changed to the nodal basis once per cluster
// (k_cluster_basis_change, transposed) before the coupling and the sums back
// to the polynomial basis once per cluster after it.

// 3'. upward: mom[t] = E_s0 mom[s0] + E_s1 mom[s1] for the inner clusters of
// one level; the sons' moments staged [2][q][NRC] in LDS, the two E matrices
// too when stage_e (else read through L2).
template <typename T, int NRC>
__global__ __launch_bounds__(kBlock) void k_upward_transfer_dense(DevTreeView<T> tr, int q,
                                                                  int cbeg, int nrhs, int stage_e,
                                                                  T *__restrict__ mom) {
  extern __shared__ unsigned char smem_raw[];
  T *S = reinterpret_cast<T *>(smem_raw);  // [2][q][NRC]
  T *Es = S + 2 * q * NRC;                 // [2][q][q] when stage_e
  const int c = cbeg + blockIdx.x;
  const int s0 = tr.son0[c], s1 = tr.son1[c];
  if (s0 < 0) return;
  const T *E0 = tr.transfer + (int64_t)s0 * q * q;
  const T *E1 = tr.transfer + (int64_t)s1 * q * q;
  if (stage_e) {
    for (int e = threadIdx.x; e < q * q; e += blockDim.x) {
      Es[e] = E0[e];
      Es[q * q + e] = E1[e];
    }
    E0 = Es;
    E1 = Es + q * q;
  }
  for (int rc = blockIdx.y * NRC; rc < nrhs; rc += gridDim.y * NRC) {
    const int nr = min(NRC, nrhs - rc);
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * q * NRC; e += blockDim.x) {
      const int si = e / (q * NRC), b = (e / NRC) % q, r = e % NRC;
      S[e] = r < nr ? mom[((int64_t)(si ? s1 : s0) * q + b) * nrhs + rc + r] : T(0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int a = e / NRC, r = e % NRC;
      if (r >= nr) continue;
      T acc = 0;
      for (int b = 0; b < q; ++b) acc += E0[a * q + b] * S[b * NRC + r];
      for (int b = 0; b < q; ++b) acc += E1[a * q + b] * S[(q + b) * NRC + r];
      mom[((int64_t)c * q + a) * nrhs + rc + r] = acc;
    }
  }
}

// 5'. downward: tlhs[c] += E_c^T tlhs[parent(c)] for the clusters of one
// level >= 1, the father's coefficients staged [q][NRC] in LDS, E_c too when
// stage_e.
template <typename T, int NRC>
__global__ __launch_bounds__(kBlock) void k_downward_transfer_dense(DevTreeView<T> tr, int q,
                                                                    int cbeg, int nrhs, int stage_e,
                                                                    T *__restrict__ tlhs) {
  extern __shared__ unsigned char smem_raw[];
  T *S = reinterpret_cast<T *>(smem_raw);  // [q][NRC]
  T *Es = S + q * NRC;                     // [q][q] when stage_e
  const int c = cbeg + blockIdx.x;
  const int p = tr.parent[c];
  const T *E = tr.transfer + (int64_t)c * q * q;
  if (stage_e) {
    for (int e = threadIdx.x; e < q * q; e += blockDim.x) Es[e] = E[e];
    E = Es;
  }
  for (int rc = blockIdx.y * NRC; rc < nrhs; rc += gridDim.y * NRC) {
    const int nr = min(NRC, nrhs - rc);
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int a = e / NRC, r = e % NRC;
      S[e] = r < nr ? tlhs[((int64_t)p * q + a) * nrhs + rc + r] : T(0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int b = e / NRC, r = e % NRC;
      if (r >= nr) continue;
      T acc = 0;
      for (int a = 0; a < q; ++a) acc += E[a * q + b] * S[a * NRC + r];
      tlhs[((int64_t)c * q + b) * nrhs + rc + r] += acc;
    }
  }
}

// Basis change of every cluster's q x nrhs block, in place:
//   transpose: x[c] <- invV^T x[c]  (polynomial moments -> nodal moments)
//   else:      x[c] <- invV x[c]    (nodal sums -> polynomial coefficients)
// invV is q x q column-major.  One workgroup per cluster.
template <typename T, int NRC>
__global__ __launch_bounds__(kBlock) void k_cluster_basis_change(int q, const T *__restrict__ invV,
                                                                 int transpose, int nrhs,
                                                                 int stage_e, T *__restrict__ x) {
  extern __shared__ unsigned char smem_raw[];
  T *S = reinterpret_cast<T *>(smem_raw);  // [q][NRC]
  T *Vs = S + q * NRC;                     // [q][q] copy of invV when stage_e
  const int c = blockIdx.x;
  const int64_t ia = transpose ? q : 1, ib = transpose ? 1 : q;  // M(a, b) = invV[a ia + b ib]
  if (stage_e) {
    for (int e = threadIdx.x; e < q * q; e += blockDim.x) Vs[e] = invV[e];
    invV = Vs;
  }
  for (int rc = blockIdx.y * NRC; rc < nrhs; rc += gridDim.y * NRC) {
    const int nr = min(NRC, nrhs - rc);
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int b = e / NRC, r = e % NRC;
      S[e] = r < nr ? x[((int64_t)c * q + b) * nrhs + rc + r] : T(0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < q * NRC; e += blockDim.x) {
      const int a = e / NRC, r = e % NRC;
      if (r >= nr) continue;
      T acc = 0;
      for (int b = 0; b < q; ++b) acc += invV[a * ia + b * ib] * S[b * NRC + r];
      x[((int64_t)c * q + a) * nrhs + rc + r] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// 6. leaf output: y_t = V_t^T tlhs[t] + sum_{s in near(t)} K(P_t, P_s) xs_s,
// written to out (natural order through perm, or cluster order).
// One workgroup per target leaf, thread <-> target point.  LDS holds the
// thread's Lagrange factors plus one union region used first for tiles of
// far-field coefficients, then for tiles of source points.
constexpr int kAlphaTile = 128;
constexpr int kSrcTile = 64;

template <typename T, int NR>
__host__ __device__ constexpr size_t leaf_output_union_elems(int D) {
  return (size_t)kAlphaTile * NR > (size_t)D * kSrcTile + (size_t)kSrcTile * NR
             ? (size_t)kAlphaTile * NR
             : (size_t)D * kSrcTile + (size_t)kSrcTile * NR;
}

template <typename T, int D, int KT, int NR, bool GEO = false>
__global__ __launch_bounds__(kBlock) void k_leaf_output(DevTreeView<T> rows, DevTreeView<T> cols,
                                                        DevPatternView pat, InterpParams<T> ip,
                                                        KernelParams<T> kp, int nrhs,
                                                        const T *__restrict__ xs,
                                                        const T *__restrict__ tlhs,
                                                        double *__restrict__ out, int64_t ldo,
                                                        int natural) {
  extern __shared__ unsigned char smem_raw[];
  const int P1 = ip.P1, q = ip.q;
  T *F = reinterpret_cast<T *>(smem_raw);  // [D][P1][kBlock] Lagrange factors
  T *U = F + D * P1 * kBlock;              // union region
  T *Mv = U;                               // [kAlphaTile][NR] far-field coefficients
  T *Ys = U;                               // [D][kSrcTile] source points
  T *Xs = U + D * kSrcTile;                // [kSrcTile][NR] source rhs
  const int task = pat.near_order[blockIdx.x];
  if (task < 0) return;
  const int li = pat.near_task[4 * task], n0 = pat.near_task[4 * task + 1],
            n1 = pat.near_task[4 * task + 2], flags = pat.near_task[4 * task + 3];
  const bool split = (flags & kTaskSplit) != 0, do_far = (flags & kTaskFar) != 0;
  const int c = rows.leaves[li];
  const int b0 = rows.begin[c], sz = rows.size[c];
  for (int rc = 0; rc < nrhs; rc += NR) {
    const int nr = min(NR, nrhs - rc);
    for (int pbase = 0; pbase < sz; pbase += kBlock) {
      const int i = pbase + threadIdx.x;
      const bool act = i < sz;
      T x[D];
#pragma unroll
      for (int d = 0; d < D; ++d)
        x[d] = act ? rows.pts[(int64_t)d * rows.npts + b0 + i] : T(0);
      T acc[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) acc[r] = 0;
      // far field through the leaf basis (backward_transform_impl.h:30-31),
      // applied by the first task of the leaf
      __syncthreads();
      if (act && do_far) {
#pragma unroll
        for (int d = 0; d < D; ++d)
          basis_1d<T>(rows.unit[(int64_t)d * rows.npts + b0 + i], ip,
                      F + (d * P1) * kBlock + threadIdx.x, kBlock);
      }
      int a[D];
#pragma unroll
      for (int d = 0; d < D; ++d) a[d] = 0;
      for (int at = 0; do_far && at < q; at += kAlphaTile) {
        const int na = min(kAlphaTile, q - at);
        __syncthreads();
        for (int e = threadIdx.x; e < kAlphaTile * NR; e += blockDim.x) {
          const int al = e / NR, r = e % NR;
          Mv[e] = (al < na && r < nr) ? tlhs[((int64_t)c * q + at + al) * nrhs + rc + r] : T(0);
        }
        __syncthreads();
        if (act) {
          for (int al = 0; al < na; ++al) {
            if (ip.td) {
              const int pk = ip.midx[at + al];
#pragma unroll
              for (int d = 0; d < D; ++d) a[d] = (pk >> (8 * d)) & 0xff;
            }
            T v = F[(0 * P1 + a[0]) * kBlock + threadIdx.x];
#pragma unroll
            for (int d = 1; d < D; ++d) v *= F[(d * P1 + a[d]) * kBlock + threadIdx.x];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += v * Mv[al * NR + r];
#pragma unroll
            for (int d = 0; d < D; ++d) {
              if (++a[d] < P1) break;
              a[d] = 0;
            }
          }
        }
      }
      // near field (unsymmetricNystromEvaluator.h:63-73, H2kernelActor.h:68-69);
      // a run of sources with adjacent point ranges is one source here
      for (int nb = n0; nb < n1;) {
        const int run = pat.near_run[nb];
        const int s = pat.near_src[nb], sl = pat.near_src[nb + run - 1];
        const int sb = cols.begin[s], ss = cols.begin[sl] + cols.size[sl] - sb;
        nb += run;
        for (int j0 = 0; j0 < ss; j0 += kSrcTile) {
          const int nj = min(kSrcTile, ss - j0);
          __syncthreads();
          for (int e = threadIdx.x; e < D * kSrcTile; e += blockDim.x) {
            const int d = e / kSrcTile, jj = e % kSrcTile;
            Ys[e] = jj < nj ? cols.pts[(int64_t)d * cols.npts + sb + j0 + jj] : T(0);
          }
          for (int e = threadIdx.x; e < kSrcTile * NR; e += blockDim.x) {
            const int jj = e / NR, r = e % NR;
            Xs[e] = (jj < nj && r < nr) ? xs[(int64_t)(sb + j0 + jj) * nrhs + rc + r] : T(0);
          }
          __syncthreads();
          if (act) {
            for (int jj = 0; jj < nj; ++jj) {
              T r2 = 0;
              if constexpr (GEO) {
                r2 = geodesic_r2<T, D>(x, Ys + jj, kSrcTile);
              } else {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                  const T dd = x[d] - Ys[d * kSrcTile + jj];
                  r2 += dd * dd;
                }
              }
              const T kv = kernel_r2<KT, T>(r2, kp);
#pragma unroll
              for (int r = 0; r < NR; ++r) acc[r] += kv * Xs[jj * NR + r];
            }
          }
        }
      }
      if (act) {
        const int64_t row = natural ? (int64_t)rows.perm[b0 + i] : (int64_t)(b0 + i);
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (r < nr) {
            double *p = out + row + (int64_t)(rc + r) * ldo;
            if (split)
              atomic_accumulate(p, (double)acc[r]);
            else
              *p = (double)acc[r];
          }
      }
    }
  }
}

// ---------------------------------------------------------------------------
template <typename T, int NR>
size_t output_lds(int D, int P1) {
  return ((size_t)D * P1 * kBlock + leaf_output_union_elems<T, NR>(D)) * sizeof(T);
}

template <typename T, int NR>
size_t coupling_lds(int D, int q, bool td) {
  return ((size_t)D * (td ? q : kMaxP1) + (size_t)q * NR) * sizeof(T);
}

template <typename T, int D, int KT, int NR, bool GEO = false>
hipError_t launch_far_near(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                           const DevPatternView &pat, const InterpParams<T> &ip,
                           const KernelParams<T> &kp, int nrhs, ActionBuffers<T> buf,
                           double *out, int64_t ldo, bool natural, hipStream_t stream) {
  const int q = ip.q;
  const int threads = q <= 64 ? 64 : (q <= 128 ? 128 : kBlock);
  const size_t lds = coupling_lds<T, NR>(D, q, ip.td != 0);
  hipLaunchKernelGGL((k_coupling<T, D, KT, NR, GEO>), dim3(pat.nfar_order), dim3(threads), lds, stream, rows,
                     cols, pat, ip, kp, nrhs, buf.mom, buf.tlhs);
  return hipGetLastError();
}

template <typename T, int D, int KT, int NR, bool GEO = false>
hipError_t launch_output(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const DevPatternView &pat, const InterpParams<T> &ip,
                         const KernelParams<T> &kp, int nrhs, ActionBuffers<T> buf, double *out,
                         int64_t ldo, bool natural, hipStream_t stream) {
  const size_t lds = output_lds<T, NR>(D, ip.P1);
  hipLaunchKernelGGL((k_leaf_output<T, D, KT, NR, GEO>), dim3(pat.nnear_order), dim3(kBlock), lds,
                     stream, rows, cols, pat, ip, kp, nrhs, buf.xs, buf.tlhs, out, ldo,
                     natural ? 1 : 0);
  return hipGetLastError();
}

template <typename T, int D, int NR>
hipError_t dispatch_kt(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                       const DevPatternView &pat, const InterpParams<T> &ip,
                       const KernelParams<T> &kp, int nrhs, ActionBuffers<T> buf, double *out,
                       int64_t ldo, bool natural, hipStream_t stream) {
#define FMCA_KT_CASE(K)                                                                 \
  case K:                                                                               \
    return output ? launch_output<T, D, K, NR>(rows, cols, pat, ip, kp, nrhs, buf, out, \
                                               ldo, natural, stream)                    \
                  : launch_far_near<T, D, K, NR>(rows, cols, pat, ip, kp, nrhs, buf, out, \
                                                 ldo, natural, stream);
  if (kp.geodesic) {
    // geodesic entries: FP64 only, the kernel type switched at run time
    if constexpr (sizeof(T) == 8)
      return output ? launch_output<T, D, -1, NR, true>(rows, cols, pat, ip, kp, nrhs, buf, out,
                                                        ldo, natural, stream)
                    : launch_far_near<T, D, -1, NR, true>(rows, cols, pat, ip, kp, nrhs, buf,
                                                          out, ldo, natural, stream);
    return hipErrorInvalidValue;
  }
  switch (kp.kt) {
    FMCA_KT_CASE(KT_EXPONENTIAL)
    FMCA_KT_CASE(KT_MATERN32)
    FMCA_KT_CASE(KT_MATERN52)
    FMCA_KT_CASE(KT_GAUSSIAN)
    default:
      return output ? launch_output<T, D, -1, NR>(rows, cols, pat, ip, kp, nrhs, buf, out, ldo,
                                                  natural, stream)
                    : launch_far_near<T, D, -1, NR>(rows, cols, pat, ip, kp, nrhs, buf, out,
                                                    ldo, natural, stream);
  }
#undef FMCA_KT_CASE
}

constexpr size_t kLdsLimit = 64 * 1024;

// Largest rhs tile (32, 8 or 1) whose LDS footprint fits one workgroup.
template <typename T>
int pick_nr(bool output, int D, int P1, int q, int nrhs, bool td) {
  auto fits = [&](int nr) {
    const size_t b = output ? (nr == 32   ? output_lds<T, 32>(D, P1)
                               : nr == 8 ? output_lds<T, 8>(D, P1)
                                         : output_lds<T, 1>(D, P1))
                            : (nr == 32   ? coupling_lds<T, 32>(D, q, td)
                               : nr == 8 ? coupling_lds<T, 8>(D, q, td)
                                         : coupling_lds<T, 1>(D, q, td));
    return b <= kLdsLimit;
  };
  if (nrhs > 8 && fits(32)) return 32;
  if (nrhs > 1 && fits(8)) return 8;
  return 1;
}

template <typename T, int D>
hipError_t dispatch_nr(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                       const DevPatternView &pat, const InterpParams<T> &ip,
                       const KernelParams<T> &kp, int nrhs, ActionBuffers<T> buf, double *out,
                       int64_t ldo, bool natural, hipStream_t stream) {
  if constexpr (sizeof(T) == 4) {
    if (kp.geodesic) return hipErrorInvalidValue;
    if (kp.split_bf16 && bf16_supported(output, D, ip.P1, ip.q, nrhs, kp.kt, kp.vmask))
      return launch_bf16(output, rows, cols, pat, ip, kp, nrhs, buf.xs, buf.mom, buf.tlhs, out,
                         ldo, natural, stream);
  }
  if (!kp.geodesic && mfma_supported<T>(output, D, ip.P1, ip.q, nrhs, kp.kt, kp.vmask, ip.td))
    return launch_mfma<T>(output, rows, cols, pat, ip, kp, nrhs, buf.xs, buf.mom, buf.tlhs, out,
                          ldo, natural, stream);
  // the VALU forms evaluate one variant per launch (kp.kt)
  if ((kp.vmask & (kp.vmask - 1)) != 0) return hipErrorInvalidValue;
  const bool td = ip.td != 0;
  const int nr = pick_nr<T>(output, D, ip.P1, ip.q, nrhs, td);
  if (output ? output_lds<T, 1>(D, ip.P1) > kLdsLimit : coupling_lds<T, 1>(D, ip.q, td) > kLdsLimit)
    return hipErrorInvalidConfiguration;
  if (nr == 1)
    return dispatch_kt<T, D, 1>(output, rows, cols, pat, ip, kp, nrhs, buf, out, ldo, natural,
                                stream);
  if (nr == 8)
    return dispatch_kt<T, D, 8>(output, rows, cols, pat, ip, kp, nrhs, buf, out, ldo, natural,
                                stream);
  return dispatch_kt<T, D, 32>(output, rows, cols, pat, ip, kp, nrhs, buf, out, ldo, natural,
                               stream);
}

template <typename T, int D>
hipError_t action_d(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                    const HostLevels &lrows, const HostLevels &lcols, const DevPatternView &pat,
                    const InterpParams<T> &ip, const KernelParams<T> &kp, const double *rhs,
                    int64_t ld_rhs, int nrhs, double *out, int64_t ld_out, bool natural,
                    ActionBuffers<T> buf, hipStream_t stream, const PhaseEvents *ev,
                    const StoredBlocks<T> *stored, const GradJob *grad) {
  const int q = ip.q, P1 = ip.P1;
  hipError_t err;
  auto mark = [&](int k) {
    if (ev) (void)hipEventRecord(ev->ev[k], stream);
  };
  mark(PH_PERMUTE);
  // 1. permute into cluster order
  {
    const int64_t grid = (cols.npts + kPermRows - 1) / kPermRows;
    hipLaunchKernelGGL((k_permute_in<T>), dim3((unsigned)grid), dim3(kBlock), 0, stream, rhs,
                       ld_rhs, nrhs, cols.npts, natural ? cols.pinv : nullptr, buf.xs);
  }
  // 2. leaf moments
  mark(PH_UPWARD_LEAF);
  {
    // FP64: column groups, as many as leave kBlock / G >= q threads per group
    // (n = 1e6, q = 125, 32 rhs: 0.79 -> 0.64 ms); FP32 is faster with one
    // group (0.36 against 0.40 ms: the shared basis values then cover twice
    // the columns)
    const int G = sizeof(T) == 4 ? 1 : (q <= kBlock / 4 ? 4 : (q <= kBlock / 2 ? 2 : 1));
#define FMCA_UPWARD_LEAF(NRV, GV)                                                              \
  hipLaunchKernelGGL((k_upward_leaf<T, D, NRV, GV>), dim3(cols.nleaves), dim3(kBlock),        \
                     ((size_t)D * P1 * 64 + 64 * NRV) * sizeof(T), stream, cols, ip, buf.xs,  \
                     nrhs, buf.mom)
    bool done = false;
    if constexpr (sizeof(T) == 4) {
      if (upward_leaf_mfma_supported(D, P1, q, nrhs)) {
        if ((err = launch_upward_leaf_mfma(cols, ip, buf.xs, nrhs, buf.mom, stream)) !=
            hipSuccess)
          return err;
        done = true;
      }
    }
    if (done) {
    } else if (nrhs == 1) {
      FMCA_UPWARD_LEAF(1, 1);
    } else if (nrhs <= 8) {
      if (G >= 2)
        FMCA_UPWARD_LEAF(8, 2);
      else
        FMCA_UPWARD_LEAF(8, 1);
    } else {
      if (G == 4)
        FMCA_UPWARD_LEAF(32, 4);
      else if (G == 2)
        FMCA_UPWARD_LEAF(32, 2);
      else
        FMCA_UPWARD_LEAF(32, 1);
    }
#undef FMCA_UPWARD_LEAF
  }
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // 3. upward transfers, deepest inner level first, up to the level of the
  // first far source: the moments above it are never read (zero-filled, so
  // that every cluster's moments stay finite).  Columns per workgroup: 8 (or
  // 1) with the column groups in grid y, so that a level's latency is that of
  // one 8-column group, not of all columns in one workgroup.
  mark(PH_UPWARD_TRANSFER);
  auto first_level = [](const HostLevels &lv, int id) {
    int L = 0;
    while (L <= lv.depth && lv.level_ptr[L + 1] <= id) ++L;
    return L;  // depth + 1: no cluster from id on
  };
  const int top_src = first_level(lcols, pat.far_top_source);
  const int top_tgt = first_level(lrows, pat.far_top_target);
  if (top_src > 0) {
    const int64_t count = (int64_t)lcols.level_ptr[std::min(top_src, lcols.depth + 1)] * q * nrhs;
    hipLaunchKernelGGL((k_zero_range<T>), dim3((unsigned)std::min<int64_t>(
                           1024, std::max<int64_t>(1, (count + kBlock - 1) / kBlock))),
                       dim3(kBlock), 0, stream, count, buf.mom);
  }
  auto transfer_lds = [&](int arrays, int nrc) {
    return (size_t)((q + 3) & ~3) * sizeof(int) +
           ((size_t)arrays * q * nrc + (size_t)D * P1 * P1) * sizeof(T);
  };
  auto transfer_nrc = [&](int arrays) {
    if (nrhs > 1 && transfer_lds(arrays, 8) <= kLdsLimit) return 8;
    return 1;
  };
  // total degree: dense transfers and basis changes, [arrays][q][nrc] in LDS
  auto dense_nrc = [&](int arrays) {
    if (nrhs > 1 && (size_t)arrays * q * 8 * sizeof(T) <= kLdsLimit) return 8;
    return 1;
  };
  auto groups = [&](int nrc) { return (unsigned)((nrhs + nrc - 1) / nrc); };
  auto dense_lds = [&](int arrays, int nrc) { return (size_t)arrays * q * nrc * sizeof(T); };
  // the q x q matrices (E, invV) go to LDS too where that fits
  auto mats_fit = [&](int arrays, int nrc) {
    return dense_lds(arrays, nrc) + (size_t)arrays * q * q * sizeof(T) <= kLdsLimit;
  };
  auto with_mats = [&](int arrays, int nrc) {
    return dense_lds(arrays, nrc) +
           (mats_fit(arrays, nrc) ? (size_t)arrays * q * q * sizeof(T) : 0);
  };
  if (ip.td) {
    const int nu = dense_nrc(2);
    if (dense_lds(2, nu) > kLdsLimit) return hipErrorInvalidConfiguration;
    const int se = mats_fit(2, nu) ? 1 : 0;
    for (int L = lcols.depth - 1; L >= top_src; --L) {
      const int cb = lcols.level_ptr[L], nc = lcols.level_ptr[L + 1] - cb;
      if (nu == 8)
        hipLaunchKernelGGL((k_upward_transfer_dense<T, 8>), dim3(nc, groups(8)), dim3(kBlock),
                           with_mats(2, 8), stream, cols, q, cb, nrhs, se, buf.mom);
      else
        hipLaunchKernelGGL((k_upward_transfer_dense<T, 1>), dim3(nc, groups(1)), dim3(kBlock),
                           with_mats(2, 1), stream, cols, q, cb, nrhs, se, buf.mom);
    }
  }
  auto basis_change = [&](int ncl, int transpose, T *x) {
    const int nb = dense_nrc(1);
    const int se = mats_fit(1, nb) ? 1 : 0;
    if (nb == 8)
      hipLaunchKernelGGL((k_cluster_basis_change<T, 8>), dim3(ncl, groups(8)), dim3(kBlock),
                         with_mats(1, 8), stream, q, ip.invV, transpose, nrhs, se, x);
    else
      hipLaunchKernelGGL((k_cluster_basis_change<T, 1>), dim3(ncl, groups(1)), dim3(kBlock),
                         with_mats(1, 1), stream, q, ip.invV, transpose, nrhs, se, x);
  };
  const int nrc_up = transfer_nrc(3), nrc_dn = transfer_nrc(2);
  if (!ip.td && (transfer_lds(3, nrc_up) > kLdsLimit || transfer_lds(2, nrc_dn) > kLdsLimit))
    return hipErrorInvalidConfiguration;
  for (int L = lcols.depth - 1; !ip.td && L >= top_src; --L) {
    const int cb = lcols.level_ptr[L], nc = lcols.level_ptr[L + 1] - cb;
    const size_t lds = transfer_lds(3, nrc_up);
    if (nrc_up == 8)
      hipLaunchKernelGGL((k_upward_transfer<T, D, 8>), dim3(nc, groups(8)), dim3(kBlock), lds,
                         stream, cols, ip, cb, nrhs, buf.mom);
    else
      hipLaunchKernelGGL((k_upward_transfer<T, D, 1>), dim3(nc, groups(1)), dim3(kBlock), lds,
                         stream, cols, ip, cb, nrhs, buf.mom);
  }
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // 4. coupling (every target cluster is written, empty far lists give zeros)
  mark(PH_COUPLING);
  if (ip.td) basis_change(cols.ncl, 1, buf.mom);
  if (stored) {
    // stored coupling matrices; one workgroup per target writes all of it
    if ((err = launch_stored<T>(false, rows, cols, pat, *stored, q, nrhs, buf.xs, buf.mom,
                                buf.tlhs, out, ld_out, natural, stream)) != hipSuccess)
      return err;
  } else {
    if (pat.nfar_split > 0)
      hipLaunchKernelGGL((k_zero_clusters<T>), dim3(pat.nfar_split), dim3(kBlock), 0, stream,
                         pat.far_split, (int64_t)q * nrhs, buf.tlhs);
    if ((err = dispatch_nr<T, D>(false, rows, cols, pat, ip, kp, nrhs, buf, out, ld_out, natural,
                                 stream)) != hipSuccess)
      return err;
  }
  // gradient job: the far part of dL/dl from the coupling sums of the dk/dl group
  if (grad) {
    if ((err = launch_grad_far_dot<T>(*grad, (int64_t)cols.ncl * q, buf.mom, buf.tlhs,
                                      stream)) != hipSuccess)
      return err;
  }
  if (ip.td) basis_change(rows.ncl, 0, buf.tlhs);
  // 5. downward transfers, top level first, from the level below the first
  // far target on (every cluster above it holds zeros after the coupling)
  mark(PH_DOWNWARD_TRANSFER);
  const int first_dn = std::max(1, top_tgt + 1);
  if (ip.td) {
    const int nd = dense_nrc(1);
    const int se = mats_fit(1, nd) ? 1 : 0;
    for (int L = first_dn; L <= lrows.depth; ++L) {
      const int cb = lrows.level_ptr[L], nc = lrows.level_ptr[L + 1] - cb;
      if (nd == 8)
        hipLaunchKernelGGL((k_downward_transfer_dense<T, 8>), dim3(nc, groups(8)), dim3(kBlock),
                           with_mats(1, 8), stream, rows, q, cb, nrhs, se, buf.tlhs);
      else
        hipLaunchKernelGGL((k_downward_transfer_dense<T, 1>), dim3(nc, groups(1)), dim3(kBlock),
                           with_mats(1, 1), stream, rows, q, cb, nrhs, se, buf.tlhs);
    }
  }
  for (int L = first_dn; !ip.td && L <= lrows.depth; ++L) {
    const int cb = lrows.level_ptr[L], nc = lrows.level_ptr[L + 1] - cb;
    const size_t lds = transfer_lds(2, nrc_dn);
    if (nrc_dn == 8)
      hipLaunchKernelGGL((k_downward_transfer<T, D, 8>), dim3(nc, groups(8)), dim3(kBlock), lds,
                         stream, rows, ip, cb, nrhs, buf.tlhs);
    else
      hipLaunchKernelGGL((k_downward_transfer<T, D, 1>), dim3(nc, groups(1)), dim3(kBlock), lds,
                         stream, rows, ip, cb, nrhs, buf.tlhs);
  }
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // 6. leaf output: far field through V^T plus the near field
  mark(PH_LEAF_OUTPUT);
  if (grad) {
    KernelParams<T> kg = kp;  // value, dk/dl and phi of every pair
    kg.kt = kp.kt % KT_COUNT;
    kg.vmask = (1 << KV_VALUE) | (1 << KV_DL) | (1 << KV_PHI);
    err = launch_grad_leaves<T>(rows, pat, ip, kg, *grad, buf.xs, buf.tlhs, stream);
    mark(PH_COUNT);
    return err;
  }
  if (stored) {
    err = launch_stored<T>(true, rows, cols, pat, *stored, q, nrhs, buf.xs, buf.mom, buf.tlhs,
                           out, ld_out, natural, stream);
    mark(PH_COUNT);
    return err;
  }
  if (pat.nnear_split > 0)
    hipLaunchKernelGGL((k_zero_leaf_rows<T>), dim3(pat.nnear_split), dim3(kBlock), 0, stream,
                       rows, pat.near_split, nrhs, out, ld_out, natural ? 1 : 0);
  err = dispatch_nr<T, D>(true, rows, cols, pat, ip, kp, nrhs, buf, out, ld_out, natural, stream);
  mark(PH_COUNT);
  return err;
}

}  // namespace

}  // namespace fmca_amd

namespace fmca_amd {

namespace {
template <typename T>
__global__ void k_math_probe(int fn, const double *__restrict__ x, double *__restrict__ y,
                             int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = (T)x[i];
  y[i] = (double)(fn == 0 ? Math<T>::exp_(v) : Math<T>::sqrt_(v));
}
}  // namespace

hipError_t launch_math_probe(int fn, const double *x, double *y, int64_t n, hipStream_t stream) {
  const int grid = (int)((n + kBlock - 1) / kBlock);
  if (n <= 0) return hipSuccess;
  if (fn == 0 || fn == 1)
    hipLaunchKernelGGL((k_math_probe<double>), dim3(grid), dim3(kBlock), 0, stream, fn, x, y, n);
  else if (fn == 2 || fn == 3)
    hipLaunchKernelGGL((k_math_probe<float>), dim3(grid), dim3(kBlock), 0, stream, fn - 2, x, y, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_action(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const HostLevels &lrows, const HostLevels &lcols,
                         const DevPatternView &pat, const InterpParams<T> &ip,
                         const KernelParams<T> &kp, const double *rhs, int64_t ld_rhs, int nrhs,
                         double *out, int64_t ld_out, bool natural_order, ActionBuffers<T> buf,
                         hipStream_t stream, const PhaseEvents *events,
                         const StoredBlocks<T> *stored, const GradJob *grad) {
  switch (rows.dim) {
    case 1:
      return action_d<T, 1>(rows, cols, lrows, lcols, pat, ip, kp, rhs, ld_rhs, nrhs, out,
                            ld_out, natural_order, buf, stream, events, stored, grad);
    case 2:
      return action_d<T, 2>(rows, cols, lrows, lcols, pat, ip, kp, rhs, ld_rhs, nrhs, out,
                            ld_out, natural_order, buf, stream, events, stored, grad);
    case 3:
      return action_d<T, 3>(rows, cols, lrows, lcols, pat, ip, kp, rhs, ld_rhs, nrhs, out,
                            ld_out, natural_order, buf, stream, events, stored, grad);
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_action<double>(const DevTreeView<double> &, const DevTreeView<double> &,
                                          const HostLevels &, const HostLevels &,
                                          const DevPatternView &, const InterpParams<double> &,
                                          const KernelParams<double> &, const double *, int64_t,
                                          int, double *, int64_t, bool, ActionBuffers<double>,
                                          hipStream_t, const PhaseEvents *,
                                          const StoredBlocks<double> *, const GradJob *);
template hipError_t launch_action<float>(const DevTreeView<float> &, const DevTreeView<float> &,
                                         const HostLevels &, const HostLevels &,
                                         const DevPatternView &, const InterpParams<float> &,
                                         const KernelParams<float> &, const double *, int64_t, int,
                                         double *, int64_t, bool, ActionBuffers<float>,
                                         hipStream_t, const PhaseEvents *,
                                         const StoredBlocks<float> *, const GradJob *);

}  // namespace fmca_amd
