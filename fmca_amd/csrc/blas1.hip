// Column-batched FP64 vector kernels for the CG solvers: every right-hand side
// is one column of an n x nrhs column-major array with its own scalars.
// Reductions are two-pass with a fixed grid, so results do not depend on
// scheduling (bitwise reproducible).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace fmca_amd {

namespace {

constexpr int kDotBlocks = 128;  // partial sums per column
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_dot_partial(int64_t n, const double *__restrict__ a,
                                                          int64_t lda, const double *__restrict__ b,
                                                          int64_t ldb,
                                                          double *__restrict__ partial) {
  const int col = blockIdx.y;
  const double *pa = a + col * lda;
  const double *pb = b + col * ldb;
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)kDotBlocks * kThreads)
    s = __builtin_fma(pa[i], pb[i], s);
  __shared__ double red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[col * kDotBlocks + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kDotBlocks) void k_dot_final(const double *__restrict__ partial,
                                                          double *__restrict__ out) {
  __shared__ double red[kDotBlocks];
  red[threadIdx.x] = partial[blockIdx.x * kDotBlocks + threadIdx.x];
  __syncthreads();
  for (int w = kDotBlocks / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// y_j = y_j + s * coef_j * x_j  (coef on the device, s = +1 or -1)
__global__ __launch_bounds__(kThreads) void k_axpy(int64_t n, int nrhs, double s,
                                                   const double *__restrict__ coef,
                                                   const double *__restrict__ x, int64_t ldx,
                                                   double *__restrict__ y, int64_t ldy) {
  const int col = blockIdx.y;
  const double a = s * coef[col];
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    y[col * ldy + i] = __builtin_fma(a, x[col * ldx + i], y[col * ldy + i]);
}

// p_j = z_j + coef_j * p_j
__global__ __launch_bounds__(kThreads) void k_xpby(int64_t n, int nrhs,
                                                   const double *__restrict__ coef,
                                                   const double *__restrict__ z, int64_t ldz,
                                                   double *__restrict__ p, int64_t ldp) {
  const int col = blockIdx.y;
  const double bcoef = coef[col];
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    p[col * ldp + i] = __builtin_fma(bcoef, p[col * ldp + i], z[col * ldz + i]);
}

// r_j = b_j - (y_j + ridge * x_j)
__global__ __launch_bounds__(kThreads) void k_residual(int64_t n, int nrhs, double ridge,
                                                       const double *__restrict__ b, int64_t ldb,
                                                       const double *__restrict__ y, int64_t ldy,
                                                       const double *__restrict__ x, int64_t ldx,
                                                       double *__restrict__ r, int64_t ldr) {
  const int col = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    r[col * ldr + i] = b[col * ldb + i] - __builtin_fma(ridge, x[col * ldx + i], y[col * ldy + i]);
}

// y_j += ridge * x_j
__global__ __launch_bounds__(kThreads) void k_add_ridge(int64_t n, int nrhs, double ridge,
                                                        const double *__restrict__ x, int64_t ldx,
                                                        double *__restrict__ y, int64_t ldy) {
  const int col = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    y[col * ldy + i] = __builtin_fma(ridge, x[col * ldx + i], y[col * ldy + i]);
}

// x_j = v (fill)
__global__ __launch_bounds__(kThreads) void k_fill(int64_t n, double v, double *__restrict__ x,
                                                   int64_t ldx) {
  const int col = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    x[col * ldx + i] = v;
}

__global__ __launch_bounds__(kThreads) void k_permute(int64_t n, const uint32_t *__restrict__ perm,
                                                      int gather, const double *__restrict__ in,
                                                      int64_t ldi, double *__restrict__ out,
                                                      int64_t ldo) {
  const int col = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const int64_t p = perm[i];
    if (gather)
      out[col * ldo + i] = in[col * ldi + p];
    else
      out[col * ldo + p] = in[col * ldi + i];
  }
}

// ---- Krylov basis of the GCR solve ------------------------------------------
// Pairs (p_j, q_j = (K + ridge I) p_j) live in the slots of a circular window:
// window position j is slot (first + j) % cap, and a slot holds n x nrhs values
// of type S (column-major, leading dimension n) at ptab[slot] / qtab[slot].
// Columns with active[col] == 0 are left untouched and report zero sums.

constexpr int kBasisGroup = 16;  // basis vectors per pass of a dot block

// Sum of s[k] over the block, k < N; thread k receives sum k.  Fixed order.
template <int N>
__device__ __forceinline__ double block_sums(double (&s)[N], double *lds /* [kThreads/64][N] */) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double v = s[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    s[k] = v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) lds[wave * N + k] = s[k];
  __syncthreads();
  double t = 0;
  if (threadIdx.x < N)
    for (int w = 0; w < kThreads / 64; ++w) t += lds[w * N + threadIdx.x];
  __syncthreads();
  return t;
}

// Rows of block b: a contiguous range, [b * chunk, min(n, (b + 1) * chunk)),
// chunk a multiple of 4 * kThreads; V = 4 reads four consecutive rows per
// thread (n % 4 == 0, so every column starts 16-byte aligned), V = 1 one.
template <int V>
__device__ __forceinline__ void block_rows(int64_t n, int64_t &lo, int64_t &hi) {
  const int64_t per = (n + kDotBlocks - 1) / kDotBlocks;
  const int64_t chunk = (per + 4 * kThreads - 1) / (4 * kThreads) * (4 * kThreads);
  lo = blockIdx.x * chunk;
  hi = lo + chunk < n ? lo + chunk : n;
}

template <typename T, int V>
__device__ __forceinline__ void load_v(const T *__restrict__ p, int64_t i, double (&out)[V]) {
  if constexpr (V == 4) {
    struct alignas(4 * sizeof(T)) T4 { T v[4]; };
    const T4 t = *reinterpret_cast<const T4 *>(p + i);
#pragma unroll
    for (int u = 0; u < 4; ++u) out[u] = (double)t.v[u];
  } else {
    out[0] = (double)p[i];
  }
}

template <typename T, int V>
__device__ __forceinline__ void store_v(T *__restrict__ p, int64_t i, const T (&in)[V]) {
  if constexpr (V == 4) {
    struct alignas(4 * sizeof(T)) T4 { T v[4]; };
    T4 t;
#pragma unroll
    for (int u = 0; u < 4; ++u) t.v[u] = in[u];
    *reinterpret_cast<T4 *>(p + i) = t;
  } else {
    p[i] = in[0];
  }
}

// partial[(col * count + j) * kDotBlocks + block] = sum over the block's rows of
// v * q_j (v: n x nrhs FP64)
template <typename S, int V>
__global__ __launch_bounds__(kThreads) void k_basis_dots(int64_t n, const double *__restrict__ v,
                                                         const S *const *__restrict__ qtab,
                                                         int first, int count, int cap,
                                                         const int *__restrict__ active,
                                                         double *__restrict__ partial) {
  const int col = blockIdx.y;
  const bool on = active[col] != 0;
  const double *vc = v + (size_t)col * n;
  int64_t lo, hi;
  block_rows<V>(n, lo, hi);
  __shared__ double lds[(kThreads / 64) * kBasisGroup];
  for (int g0 = 0; g0 < count; g0 += kBasisGroup) {
    const int ng = min(kBasisGroup, count - g0);
    const S *q[kBasisGroup];
#pragma unroll
    for (int k = 0; k < kBasisGroup; ++k)
      q[k] = qtab[(first + g0 + min(k, ng - 1)) % cap] + (size_t)col * n;
    double s[kBasisGroup];
#pragma unroll
    for (int k = 0; k < kBasisGroup; ++k) s[k] = 0;
    if (on)
      for (int64_t i = lo + (int64_t)threadIdx.x * V; i < hi; i += (int64_t)kThreads * V) {
        double vi[V];
        load_v<double, V>(vc, i, vi);
#pragma unroll
        for (int k = 0; k < kBasisGroup; ++k)
          if (k < ng) {
            double qi[V];
            load_v<S, V>(q[k], i, qi);
#pragma unroll
            for (int u = 0; u < V; ++u) s[k] = __builtin_fma(vi[u], qi[u], s[k]);
          }
      }
    const double t = block_sums(s, lds);
    if (threadIdx.x < ng)
      partial[((size_t)col * count + g0 + threadIdx.x) * kDotBlocks + blockIdx.x] = t;
  }
}

// coef[col * count + j] = (v . q_j) / (q_j . q_j), zero for inactive columns
// and empty pairs; qq[slot * nrhs + col] holds q_slot . q_slot
__global__ __launch_bounds__(kDotBlocks) void k_basis_coef(const double *__restrict__ partial,
                                                           int count, const double *__restrict__ qq,
                                                           int first, int cap, int nrhs,
                                                           const int *__restrict__ active,
                                                           double *__restrict__ coef) {
  const int j = blockIdx.x, col = blockIdx.y;
  __shared__ double red[kDotBlocks];
  red[threadIdx.x] = partial[((size_t)col * count + j) * kDotBlocks + threadIdx.x];
  __syncthreads();
  for (int w = kDotBlocks / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d = qq[(size_t)((first + j) % cap) * nrhs + col];
    coef[(size_t)col * count + j] = active[col] && d > 0 ? red[0] / d : 0.0;
  }
}

// The new pair from q0 = (K + ridge I) r (in y) and p0 = r, orthogonalised
// against the window: q = q0 - sum_j c_j q_j, p = r - sum_j c_j p_j, stored in
// slot (qk, pk); partial sums of q.q and r.q (of the stored q) per block.
template <typename S, int V>
__global__ __launch_bounds__(kThreads) void k_basis_new_pair(
    int64_t n, const double *__restrict__ y, const double *__restrict__ r,
    const double *__restrict__ coef, const S *const *__restrict__ qtab,
    const S *const *__restrict__ ptab, int first, int count, int cap, S *__restrict__ qk,
    S *__restrict__ pk, const int *__restrict__ active, double *__restrict__ partial) {
  const int col = blockIdx.y;
  __shared__ double c[kBasisMax];
  __shared__ const S *qs[kBasisMax];
  __shared__ const S *ps[kBasisMax];
  __shared__ double lds[(kThreads / 64) * 2];
  const bool on = active[col] != 0;
  for (int j = threadIdx.x; j < count; j += kThreads) {
    const int sl = (first + j) % cap;
    c[j] = coef[(size_t)col * count + j];
    qs[j] = qtab[sl] + (size_t)col * n;
    ps[j] = ptab[sl] + (size_t)col * n;
  }
  __syncthreads();
  int64_t lo, hi;
  block_rows<V>(n, lo, hi);
  double s[2] = {0, 0};
  if (on) {
    const double *yc = y + (size_t)col * n, *rc = r + (size_t)col * n;
    S *qc = qk + (size_t)col * n, *pc = pk + (size_t)col * n;
    for (int64_t i = lo + (int64_t)threadIdx.x * V; i < hi; i += (int64_t)kThreads * V) {
      double q[V], p[V];
      load_v<double, V>(yc, i, q);
      load_v<double, V>(rc, i, p);
      double ri[V];
#pragma unroll
      for (int u = 0; u < V; ++u) ri[u] = p[u];
      for (int j = 0; j < count; ++j) {
        double qj[V], pj[V];
        load_v<S, V>(qs[j], i, qj);
        load_v<S, V>(ps[j], i, pj);
#pragma unroll
        for (int u = 0; u < V; ++u) {
          q[u] = __builtin_fma(-c[j], qj[u], q[u]);
          p[u] = __builtin_fma(-c[j], pj[u], p[u]);
        }
      }
      S qo[V], po[V];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        qo[u] = (S)q[u];
        po[u] = (S)p[u];
        const double qd = (double)qo[u];
        s[0] = __builtin_fma(qd, qd, s[0]);
        s[1] = __builtin_fma(ri[u], qd, s[1]);
      }
      store_v<S, V>(qc, i, qo);
      store_v<S, V>(pc, i, po);
    }
  }
  const double t = block_sums(s, lds);
  if (threadIdx.x < 2) partial[((size_t)col * 2 + threadIdx.x) * kDotBlocks + blockIdx.x] = t;
}

// x += a p, r -= a q with a = (r.q) / (q.q) from qqrq[2 col + {0, 1}];
// records q.q of the new slot; partial sums of the new r.r per block
template <typename S, int V>
__global__ __launch_bounds__(kThreads) void k_basis_update(
    int64_t n, const S *__restrict__ pk, const S *__restrict__ qk,
    const double *__restrict__ qqrq, double *__restrict__ qq_slot, double *__restrict__ x,
    double *__restrict__ r, const int *__restrict__ active, double *__restrict__ partial) {
  const int col = blockIdx.y;
  __shared__ double lds[kThreads / 64];
  double s[1] = {0};
  int64_t lo, hi;
  block_rows<V>(n, lo, hi);
  if (active[col]) {
    const double qq = qqrq[2 * col], rq = qqrq[2 * col + 1];
    const double a = qq > 0 ? rq / qq : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) qq_slot[col] = qq;
    const S *pc = pk + (size_t)col * n, *qc = qk + (size_t)col * n;
    double *xc = x + (size_t)col * n, *rc = r + (size_t)col * n;
    for (int64_t i = lo + (int64_t)threadIdx.x * V; i < hi; i += (int64_t)kThreads * V) {
      double pv[V], qv[V], xv[V], rv[V];
      load_v<S, V>(pc, i, pv);
      load_v<S, V>(qc, i, qv);
      load_v<double, V>(xc, i, xv);
      load_v<double, V>(rc, i, rv);
#pragma unroll
      for (int u = 0; u < V; ++u) {
        xv[u] = __builtin_fma(a, pv[u], xv[u]);
        rv[u] = __builtin_fma(-a, qv[u], rv[u]);
        s[0] = __builtin_fma(rv[u], rv[u], s[0]);
      }
      store_v<double, V>(xc, i, xv);
      store_v<double, V>(rc, i, rv);
    }
  }
  const double t = block_sums(s, lds);
  if (threadIdx.x == 0) partial[(size_t)col * kDotBlocks + blockIdx.x] = t;
}

// x += sum_j c_j p_j, r -= sum_j c_j q_j; partial sums of the new r.r per block
template <typename S, int V>
__global__ __launch_bounds__(kThreads) void k_basis_project(
    int64_t n, const double *__restrict__ coef, const S *const *__restrict__ qtab,
    const S *const *__restrict__ ptab, int first, int count, int cap, double *__restrict__ x,
    double *__restrict__ r, const int *__restrict__ active, double *__restrict__ partial) {
  const int col = blockIdx.y;
  __shared__ double c[kBasisMax];
  __shared__ const S *qs[kBasisMax];
  __shared__ const S *ps[kBasisMax];
  __shared__ double lds[kThreads / 64];
  for (int j = threadIdx.x; j < count; j += kThreads) {
    const int sl = (first + j) % cap;
    c[j] = coef[(size_t)col * count + j];
    qs[j] = qtab[sl] + (size_t)col * n;
    ps[j] = ptab[sl] + (size_t)col * n;
  }
  __syncthreads();
  int64_t lo, hi;
  block_rows<V>(n, lo, hi);
  double s[1] = {0};
  if (active[col]) {
    double *xc = x + (size_t)col * n, *rc = r + (size_t)col * n;
    for (int64_t i = lo + (int64_t)threadIdx.x * V; i < hi; i += (int64_t)kThreads * V) {
      double xv[V], rv[V];
      load_v<double, V>(xc, i, xv);
      load_v<double, V>(rc, i, rv);
      for (int j = 0; j < count; ++j) {
        double qj[V], pj[V];
        load_v<S, V>(qs[j], i, qj);
        load_v<S, V>(ps[j], i, pj);
#pragma unroll
        for (int u = 0; u < V; ++u) {
          xv[u] = __builtin_fma(c[j], pj[u], xv[u]);
          rv[u] = __builtin_fma(-c[j], qj[u], rv[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < V; ++u) s[0] = __builtin_fma(rv[u], rv[u], s[0]);
      store_v<double, V>(xc, i, xv);
      store_v<double, V>(rc, i, rv);
    }
  }
  const double t = block_sums(s, lds);
  if (threadIdx.x == 0) partial[(size_t)col * kDotBlocks + blockIdx.x] = t;
}

dim3 grid_for(int64_t n, int nrhs) {
  const int64_t g = (n + kThreads - 1) / kThreads;
  return dim3((unsigned)(g < 1024 ? (g > 0 ? g : 1) : 1024), (unsigned)nrhs);
}

}  // namespace

size_t blas1_dot_scratch(int nrhs) { return (size_t)nrhs * kDotBlocks * sizeof(double); }

hipError_t blas1_dot(int64_t n, int nrhs, const double *a, int64_t lda, const double *b,
                     int64_t ldb, double *out, double *scratch, hipStream_t stream) {
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream, n, a, lda,
                     b, ldb, scratch);
  hipLaunchKernelGGL(k_dot_final, dim3(nrhs), dim3(kDotBlocks), 0, stream, scratch, out);
  return hipGetLastError();
}

hipError_t blas1_axpy(int64_t n, int nrhs, double s, const double *coef, const double *x,
                      int64_t ldx, double *y, int64_t ldy, hipStream_t stream) {
  hipLaunchKernelGGL(k_axpy, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, nrhs, s, coef, x,
                     ldx, y, ldy);
  return hipGetLastError();
}

hipError_t blas1_xpby(int64_t n, int nrhs, const double *coef, const double *z, int64_t ldz,
                      double *p, int64_t ldp, hipStream_t stream) {
  hipLaunchKernelGGL(k_xpby, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, nrhs, coef, z, ldz,
                     p, ldp);
  return hipGetLastError();
}

hipError_t blas1_residual(int64_t n, int nrhs, double ridge, const double *b, int64_t ldb,
                          const double *y, int64_t ldy, const double *x, int64_t ldx, double *r,
                          int64_t ldr, hipStream_t stream) {
  hipLaunchKernelGGL(k_residual, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, nrhs, ridge, b,
                     ldb, y, ldy, x, ldx, r, ldr);
  return hipGetLastError();
}

hipError_t blas1_add_ridge(int64_t n, int nrhs, double ridge, const double *x, int64_t ldx,
                           double *y, int64_t ldy, hipStream_t stream) {
  hipLaunchKernelGGL(k_add_ridge, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, nrhs, ridge, x,
                     ldx, y, ldy);
  return hipGetLastError();
}

hipError_t blas1_fill(int64_t n, int nrhs, double v, double *x, int64_t ldx, hipStream_t stream) {
  hipLaunchKernelGGL(k_fill, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, v, x, ldx);
  return hipGetLastError();
}

size_t basis_scratch(int nrhs, int count) {
  return (size_t)nrhs * (size_t)(count > 2 ? count : 2) * kDotBlocks * sizeof(double);
}

template <typename S>
hipError_t basis_dots(int64_t n, int nrhs, const double *v, const S *const *qtab, int first,
                      int count, int cap, const double *qq, const int *active, double *coef,
                      double *scratch, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  if (count > kBasisMax) return hipErrorInvalidValue;
  if (n % 4 == 0)
    hipLaunchKernelGGL((k_basis_dots<S, 4>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream,
                       n, v, qtab, first, count, cap, active, scratch);
  else
    hipLaunchKernelGGL((k_basis_dots<S, 1>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream,
                       n, v, qtab, first, count, cap, active, scratch);
  hipLaunchKernelGGL(k_basis_coef, dim3(count, nrhs), dim3(kDotBlocks), 0, stream, scratch, count,
                     qq, first, cap, nrhs, active, coef);
  return hipGetLastError();
}

template <typename S>
hipError_t basis_new_pair(int64_t n, int nrhs, const double *y, const double *r,
                          const double *coef, const S *const *qtab, const S *const *ptab,
                          int first, int count, int cap, S *qk, S *pk, const int *active,
                          double *qqrq, double *scratch, hipStream_t stream) {
  if (count < 0 || count > kBasisMax) return hipErrorInvalidValue;
  if (n % 4 == 0)
    hipLaunchKernelGGL((k_basis_new_pair<S, 4>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0,
                       stream, n, y, r, coef, qtab, ptab, first, count, cap, qk, pk, active,
                       scratch);
  else
    hipLaunchKernelGGL((k_basis_new_pair<S, 1>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0,
                       stream, n, y, r, coef, qtab, ptab, first, count, cap, qk, pk, active,
                       scratch);
  hipLaunchKernelGGL(k_dot_final, dim3(2 * nrhs), dim3(kDotBlocks), 0, stream, scratch, qqrq);
  return hipGetLastError();
}

template <typename S>
hipError_t basis_update(int64_t n, int nrhs, const S *pk, const S *qk, const double *qqrq,
                        double *qq_slot, double *x, double *r, const int *active, double *rr,
                        double *scratch, hipStream_t stream) {
  if (n % 4 == 0)
    hipLaunchKernelGGL((k_basis_update<S, 4>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream,
                       n, pk, qk, qqrq, qq_slot, x, r, active, scratch);
  else
    hipLaunchKernelGGL((k_basis_update<S, 1>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream,
                       n, pk, qk, qqrq, qq_slot, x, r, active, scratch);
  hipLaunchKernelGGL(k_dot_final, dim3(nrhs), dim3(kDotBlocks), 0, stream, scratch, rr);
  return hipGetLastError();
}

template <typename S>
hipError_t basis_project(int64_t n, int nrhs, const double *coef, const S *const *qtab,
                         const S *const *ptab, int first, int count, int cap, double *x, double *r,
                         const int *active, double *rr, double *scratch, hipStream_t stream) {
  if (count < 0 || count > kBasisMax) return hipErrorInvalidValue;
  if (n % 4 == 0)
    hipLaunchKernelGGL((k_basis_project<S, 4>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream,
                       n, coef, qtab, ptab, first, count, cap, x, r, active, scratch);
  else
    hipLaunchKernelGGL((k_basis_project<S, 1>), dim3(kDotBlocks, nrhs), dim3(kThreads), 0, stream,
                       n, coef, qtab, ptab, first, count, cap, x, r, active, scratch);
  hipLaunchKernelGGL(k_dot_final, dim3(nrhs), dim3(kDotBlocks), 0, stream, scratch, rr);
  return hipGetLastError();
}

#define FMCA_BASIS_INST(S)                                                                     \
  template hipError_t basis_dots<S>(int64_t, int, const double *, const S *const *, int, int, \
                                    int, const double *, const int *, double *, double *,     \
                                    hipStream_t);                                             \
  template hipError_t basis_new_pair<S>(int64_t, int, const double *, const double *,         \
                                        const double *, const S *const *, const S *const *,   \
                                        int, int, int, S *, S *, const int *, double *,       \
                                        double *, hipStream_t);                               \
  template hipError_t basis_update<S>(int64_t, int, const S *, const S *, const double *,     \
                                      double *, double *, double *, const int *, double *,    \
                                      double *, hipStream_t);                                 \
  template hipError_t basis_project<S>(int64_t, int, const double *, const S *const *,        \
                                       const S *const *, int, int, int, double *, double *,   \
                                       const int *, double *, double *, hipStream_t);
FMCA_BASIS_INST(double)
FMCA_BASIS_INST(float)
#undef FMCA_BASIS_INST

hipError_t blas1_permute(int64_t n, int nrhs, const uint32_t *perm, bool gather, const double *in,
                         int64_t ldi, double *out, int64_t ldo, hipStream_t stream) {
  hipLaunchKernelGGL(k_permute, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, perm,
                     gather ? 1 : 0, in, ldi, out, ldo);
  return hipGetLastError();
}

}  // namespace fmca_amd
