// GPU evaluation of the kernel blocks of the samplet compression
// (NystromEvaluator::compute_dense_block / interpolate_kernel,
// FMCA/src/MatrixEvaluator/NystromEvaluator.h:39-67, each followed by the
// cluster-basis products of recursivelyComputeBlock,
// FMCA/src/Samplets/samplet_matrix_compressor.h:336-345):
//   dense job (two leaves r, c):  B = Q_r^T K(P_r, P_c) Q_c
//   far job (admissible r, c):    B = W_r^T K(xi_r, xi_c) W_c,  W = invV^T V
// with xi the interpolation nodes in the cluster's box.  One workgroup per
// job; the kernel entries are generated once per output tile into LDS in
// 32 x 64 chunks, contracted with A_c into a 32-row slab of G = K A_c (LDS),
// which is folded into the output tile as A_r^T G.  FP64 throughout.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "device_math.hpp"
#include "samplets.hpp"

namespace fmca_amd {

namespace {

using namespace dev;

constexpr int kTile = 96;   // output tile (rows and columns of B)
constexpr int kRC = 32;     // rows of X per slab
constexpr int kKC = 64;     // columns of the kernel chunk
constexpr int kAcc = kTile * kTile / kBlock;  // 36 output entries per thread
constexpr int kG = kRC * kTile / kBlock;      // 12 slab entries per thread

__global__ __launch_bounds__(kBlock) void k_samplet_jobs(SampletDeviceData d,
                                                         KernelParams<double> kp,
                                                         double *__restrict__ out) {
  __shared__ double Xs[kMaxDim][kRC];
  __shared__ double Ys[kMaxDim][kKC];
  __shared__ double Ks[kRC][kKC + 1];
  __shared__ double Gs[kRC][kTile + 1];
  const SampletJob job = d.jobs[blockIdx.x];
  const bool far = job.kind == 1;
  const int r = job.r, c = job.c, D = d.dim;
  const int nx = far ? d.q : (int)d.size[r];
  const int ny = far ? d.q : (int)d.size[c];
  const int ar = d.qdim[r], ac = d.qdim[c];
  const double *Ar = far ? d.W + d.v_off[r] : d.Q + d.q_off[r];  // nx x ar
  const double *Ac = far ? d.W + d.v_off[c] : d.Q + d.q_off[c];  // ny x ac
  const int tid = threadIdx.x;
  // coordinate i of point / node j of cluster cl
  auto coord = [&](int cl, int j, int i, int64_t b0) -> double {
    if (far)
      return d.unodes[(int64_t)i * d.q + j] * d.bbsize[(int64_t)cl * D + i] +
             d.bbmin[(int64_t)cl * D + i];
    return d.pts[(int64_t)i * d.n + b0 + j];
  };
  const int64_t br = far ? 0 : d.begin[r], bc = far ? 0 : d.begin[c];
  for (int ta = 0; ta < ar; ta += kTile)
    for (int tc = 0; tc < ac; tc += kTile) {
      double acc[kAcc];
#pragma unroll
      for (int s = 0; s < kAcc; ++s) acc[s] = 0;
      for (int i0 = 0; i0 < nx; i0 += kRC) {
        const int nr = min(kRC, nx - i0);
        __syncthreads();
        for (int e = tid; e < D * kRC; e += kBlock) {
          const int i = e / kRC, ii = e % kRC;
          Xs[i][ii] = ii < nr ? coord(r, i0 + ii, i, br) : 0.0;
        }
        double g[kG];
#pragma unroll
        for (int s = 0; s < kG; ++s) g[s] = 0;
        for (int k0 = 0; k0 < ny; k0 += kKC) {
          const int nk = min(kKC, ny - k0);
          __syncthreads();
          for (int e = tid; e < D * kKC; e += kBlock) {
            const int i = e / kKC, kk = e % kKC;
            Ys[i][kk] = kk < nk ? coord(c, k0 + kk, i, bc) : 0.0;
          }
          __syncthreads();
          for (int e = tid; e < kRC * kKC; e += kBlock) {
            const int ii = e / kKC, kk = e % kKC;
            double v = 0;
            if (ii < nr && kk < nk) {
              double r2 = 0;
              for (int i = 0; i < D; ++i) {
                const double t = Xs[i][ii] - Ys[i][kk];
                r2 = __builtin_fma(t, t, r2);
              }
              v = kernel_r2<-1, double>(r2, kp);
            }
            Ks[ii][kk] = v;
          }
          __syncthreads();
#pragma unroll
          for (int s = 0; s < kG; ++s) {
            const int e = tid + kBlock * s;
            const int ii = e / kTile, jj = e % kTile;
            if (ii < nr && tc + jj < ac) {
              const double *col = Ac + (int64_t)(tc + jj) * ny + k0;
              double t = g[s];
              for (int kk = 0; kk < nk; ++kk) t = __builtin_fma(Ks[ii][kk], col[kk], t);
              g[s] = t;
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kG; ++s) {
          const int e = tid + kBlock * s;
          Gs[e / kTile][e % kTile] = g[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kAcc; ++s) {
          const int e = tid + kBlock * s;
          const int ai = e % kTile, cj = e / kTile;
          if (ta + ai < ar && tc + cj < ac) {
            const double *col = Ar + (int64_t)(ta + ai) * nx + i0;
            double t = acc[s];
            for (int ii = 0; ii < nr; ++ii) t = __builtin_fma(col[ii], Gs[ii][cj], t);
            acc[s] = t;
          }
        }
      }
#pragma unroll
      for (int s = 0; s < kAcc; ++s) {
        const int e = tid + kBlock * s;
        const int ai = e % kTile, cj = e / kTile;
        if (ta + ai < ar && tc + cj < ac) out[job.off - d.base + (ta + ai) + (int64_t)(tc + cj) * ar] = acc[s];
      }
    }
}

}  // namespace

int launch_samplet_jobs(const SampletDeviceData &d, const SampletKernel &k, int64_t njobs,
                        double *d_out, void *stream) {
  if (njobs <= 0) return (int)hipSuccess;
  KernelParams<double> kp;
  kp.kt = k.kt;
  kp.l = k.l;
  kp.c = k.c;
  kp.nu = k.nu;
  kp.s = kernel_radial_scale(k.kt, k.l);
  kp.sl2 = kp.s * 1.4426950408889634074;
  kp.inv_l = 1.0 / k.l;
  kp.c2 = k.c * k.c;
  kp.vmask = 1;
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // grid.x is limited to 2^31 - 1 workgroups; launch in pieces
  const int64_t piece = 1 << 30;
  for (int64_t j0 = 0; j0 < njobs; j0 += piece) {
    SampletDeviceData dj = d;
    dj.jobs = d.jobs + j0;
    const unsigned nb = (unsigned)std::min<int64_t>(piece, njobs - j0);
    hipLaunchKernelGGL(k_samplet_jobs, dim3(nb), dim3(kBlock), 0, s, dj, kp, d_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
  }
  return (int)hipSuccess;
}

}  // namespace fmca_amd
