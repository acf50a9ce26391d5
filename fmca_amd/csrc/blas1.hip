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

hipError_t blas1_permute(int64_t n, int nrhs, const uint32_t *perm, bool gather, const double *in,
                         int64_t ldi, double *out, int64_t ldo, hipStream_t stream) {
  hipLaunchKernelGGL(k_permute, grid_for(n, nrhs), dim3(kThreads), 0, stream, n, perm,
                     gather ? 1 : 0, in, ldi, out, ldo);
  return hipGetLastError();
}

}  // namespace fmca_amd
