// Gradient of L(alpha, l, points) = <G, K~(points; l) alpha> for one shared
// point set (targets = sources, one tree), the backward of the
// differentiable H2 matvec y = K~ alpha with dL/dy = G.  K~ is FMCA's H2
// matrix: exact kernel entries on the near blocks, V_t S_ts V_s^T on the far
// blocks, with boxes (hence S and the interpolation nodes) held fixed.  Then
//   dL/dalpha = K~^T G = K~ G
//   dL/dl     = <G, dK~/dl alpha>  (the same blocks with dk/dl entries)
//   dL/dx_p   = sum_{j in near(p)} phi(r_pj) (x_p - x_j) W_pj
//             + sum_k [G_pk dV(x_p)^T l^alpha_t,k + alpha_pk dV(x_p)^T l^G_t,k]
// with phi = k'(r) / r, W_pj = G_p . alpha_j + alpha_p . G_j (the two ends of
// the symmetric pair), and l^alpha_t, l^G_t the far-field sums of the target's
// leaf after the downward pass (the far field of K~ alpha and K~ G): the
// derivative of the interpolated far field, not an interpolation of the
// derivative.  These kernels run after the upward pass, the coupling and the
// downward pass of an action on the columns [alpha | G | alpha], the last
// group with the dk/dl kernel (its coupling sums, dotted with G's moments,
// give the far part of dL/dl).  No reference counterpart: FMCA has no
// derivatives; the forms are checked against finite differences of the H2
// forward and against dense FP64 gradients.
//
// Column layout of the action's rhs (R = 8 or 16, r <= R real columns each):
//   [0, R) alpha, [R, 2R) G, [2R, 2R + 16) alpha under dk/dl; zero padding.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"
#include "device_math.hpp"

namespace fmca_amd {

namespace {

using namespace dev;

template <typename T>
struct GMfma;

template <>
struct GMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int crow(int lk, int v) { return 4 * lk + v; }
};

template <>
struct GMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int crow(int lk, int v) { return lk + 4 * v; }
};

constexpr int kWaves = 4;
constexpr int kRT = 4;
constexpr int kRows = kWaves * kRT * 16;
constexpr int kGChunk = 64;    // near sources staged per chunk
constexpr int kFarRows = 64;   // rows per pass of the far kernel (one wave)
constexpr int kFarMom = 64;    // far sums staged per chunk

__device__ __forceinline__ int unpack(int p, int d) { return (p >> (8 * d)) & 0xff; }

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m) {
  if constexpr (sizeof(T) == 8) {
    const int lo = __shfl_xor((int)(__double_as_longlong(v) & 0xffffffff), m);
    const int hi = __shfl_xor((int)(__double_as_longlong(v) >> 32), m);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  } else {
    return __shfl_xor(v, m);
  }
}

// ---------------------------------------------------------------------------
// Near field: one workgroup per leaf-output task (target leaf, range of its
// near list).  Per 16 x 16 tile of (target row p, source j):
//   W^T (j x p) = [alpha_j | G_j] [G_p | alpha_p]^T on the matrix cores
//   (K = 2R), whose accumulator layout (row crow(lk, v), column lr) is also
//   the A-operand layout of the K-step v of the d-alpha product, source
//   j = crow(lk, v);
//   per pair: k, dk/dl, phi from one evaluation; phi W (x_p - x_j) into the
//   point sums, dk/dl W into the lengthscale sum;
//   d-alpha: A = k, B = G_j on the matrix cores.
template <typename T, int D, int KT, int R>
__global__ __launch_bounds__(256) void k_grad_near(
    DevTreeView<T> tree, DevPatternView pat, KernelParams<T> kp, int r, int ncol,
    const T *__restrict__ xs, double *__restrict__ dalpha, int64_t ld_dalpha,
    double *__restrict__ dpoints, int64_t ld_points, double *__restrict__ dl_part) {
  using M = GMfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int XSP = 2 * R + 1;   // LDS stride of a staged source row (odd: no bank repeats)
  constexpr int KW = 2 * R / 4;    // K-steps of the pair weights
  __shared__ T Ys[D][kGChunk];
  __shared__ T Xs[kGChunk * XSP];
  __shared__ T red[kWaves];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  T dls = 0;  // this lane's share of sum dk/dl W
  const int task = pat.near_order[blockIdx.x];
  if (task >= 0) {
    const int li = pat.near_task[4 * task], n0 = pat.near_task[4 * task + 1],
              n1 = pat.near_task[4 * task + 2];
    const int c = tree.leaves[li];
    const int b0 = tree.begin[c], sz = tree.size[c];
    for (int pbase = 0; pbase < sz; pbase += kRows) {
      T x[kRT][D];
      T bq[kRT][KW];  // [G_p | alpha_p] at columns 4 s + lk of this lane's row
      bool tile_on[kRT];
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        const int tile_row0 = pbase + (wave + kWaves * rt) * 16;
        tile_on[rt] = tile_row0 < sz;
        const int i = tile_row0 + lr;
#pragma unroll
        for (int d = 0; d < D; ++d) x[rt][d] = i < sz ? tree.pts[(int64_t)d * tree.npts + b0 + i] : T(0);
#pragma unroll
        for (int s = 0; s < KW; ++s) {
          const int cc = 4 * s + lk;                 // column of [G | alpha]
          const int src = cc < R ? R + cc : cc - R;  // its column of [alpha | G]
          bq[rt][s] = i < sz ? xs[(int64_t)(b0 + i) * ncol + src] : T(0);
        }
      }
      acc_t da[kRT];
      T sd[kRT][D];  // sum_j phi W (x_p - x_j) of this lane's pairs
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        da[rt] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d) sd[rt][d] = 0;
      }
      int nb = n0, j0 = 0;
      while (nb < n1) {
        const int run = pat.near_run[nb];
        const int s0 = pat.near_src[nb], s1 = pat.near_src[nb + run - 1];
        const int sb = tree.begin[s0], ss = tree.begin[s1] + tree.size[s1] - sb;
        const int nj = min(kGChunk, ss - j0);
        __syncthreads();
        for (int e = threadIdx.x; e < D * kGChunk; e += blockDim.x) {
          const int d = e / kGChunk, jj = e % kGChunk;
          Ys[d][jj] = jj < nj ? tree.pts[(int64_t)d * tree.npts + sb + j0 + jj] : T(0);
        }
        for (int e = threadIdx.x; e < kGChunk * 2 * R; e += blockDim.x) {
          const int jj = e / (2 * R), cc = e % (2 * R);
          Xs[jj * XSP + cc] = jj < nj ? xs[(int64_t)(sb + j0 + jj) * ncol + cc] : T(0);
        }
        __syncthreads();
        j0 += kGChunk;
        if (j0 >= ss) {
          nb += run;
          j0 = 0;
        }
        for (int jb = 0; jb < nj; jb += 16) {
#pragma unroll
          for (int rt = 0; rt < kRT; ++rt) {
            if (!tile_on[rt]) continue;
            acc_t w = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < KW; ++s)
              w = M::mma(Xs[(jb + lr) * XSP + 4 * s + lk], bq[rt][s], w);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int j = jb + M::crow(lk, v);
              T r2 = 0, dd[D];
#pragma unroll
              for (int d = 0; d < D; ++d) {
                dd[d] = x[rt][d] - Ys[d][j];
                r2 = fmat(dd[d], dd[d], r2);
              }
              T ev[3];
              kernel_r2_all<KT, T>(r2, kp, ev);
              const T wv = w[v];
              dls = fmat(ev[KV_DL], wv, dls);
              const T u = ev[KV_PHI] * wv;
#pragma unroll
              for (int d = 0; d < D; ++d) sd[rt][d] = fmat(u, dd[d], sd[rt][d]);
              const T b = lr < R ? Xs[j * XSP + R + lr] : T(0);
              da[rt] = M::mma(ev[KV_VALUE], b, da[rt]);
            }
          }
        }
      }
      // point sums: add the four lane groups that share a row
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        T g[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          T b = sd[rt][d];
          b += shfl_xor(b, 16);
          b += shfl_xor(b, 32);
          g[d] = b;
        }
        const int i = pbase + (wave + kWaves * rt) * 16 + lr;
        if (tile_on[rt] && lk == 0 && i < sz) {
          const int64_t p = tree.perm[b0 + i];
#pragma unroll
          for (int d = 0; d < D; ++d) atomic_accumulate(dpoints + d * ld_points + p, (double)g[d]);
        }
        if (!tile_on[rt]) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ii = pbase + (wave + kWaves * rt) * 16 + M::crow(lk, v);
          if (ii < sz && lr < r)
            atomic_accumulate(dalpha + (int64_t)lr * ld_dalpha + tree.perm[b0 + ii],
                              (double)da[rt][v]);
        }
      }
    }
  }
  // the workgroup's sum of dk/dl W, in a fixed order
  T a = dls;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a += shfl_xor(a, m);
  if (lane == 0) red[wave] = a;
  __syncthreads();
  if (threadIdx.x == 0) dl_part[blockIdx.x] = (double)((red[0] + red[1]) + (red[2] + red[3]));
}

// ---------------------------------------------------------------------------
// Far field: one wave per leaf, a thread per row.  With the leaf's far sums
// l (q x [alpha | G] columns after the downward pass) and the basis V_a(x_p)
// and its gradient (the tensor basis: products of 1-D Lagrange factors and
// their derivatives, scaled by 1 / bbsize for the unit coordinates):
//   d-alpha_p += V(x_p)^T l^G
//   dx_p      += sum_a grad V_a(x_p) ([G_p | alpha_p] . l_a)
template <typename T, int D, int R>
__global__ __launch_bounds__(kFarRows) void k_grad_far(
    DevTreeView<T> tree, InterpParams<T> ip, int r, int ncol, const T *__restrict__ xs,
    const T *__restrict__ tlhs, double *__restrict__ dalpha, int64_t ld_dalpha,
    double *__restrict__ dpoints, int64_t ld_points) {
  constexpr int KP = kMaxP1;
  __shared__ T F[D][KP][kFarRows];
  __shared__ T dF[D][KP][kFarRows];
  __shared__ T L[kFarMom][2 * R + 1];
  __shared__ int aidx[kFarMom];
  const int leaf = blockIdx.x;
  const int c = tree.leaves[leaf];
  const int b0 = tree.begin[c], sz = tree.size[c];
  const int q = ip.q, P1 = ip.P1;
  T inv_size[D];
#pragma unroll
  for (int d = 0; d < D; ++d) inv_size[d] = T(1) / tree.bbsize[c * D + d];
  for (int pbase = 0; pbase < sz; pbase += kFarRows) {
    const int i = pbase + threadIdx.x;
    const bool on = i < sz;
    __syncthreads();
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T u = on ? tree.unit[(int64_t)d * tree.npts + b0 + i] : T(0.5);
      lagrange_1d<T>(u, ip, &F[d][0][threadIdx.x], kFarRows);
      lagrange_1d_deriv<T>(u, ip, &dF[d][0][threadIdx.x], kFarRows);
    }
    T wp[2 * R];  // [G_p | alpha_p]
#pragma unroll
    for (int cc = 0; cc < 2 * R; ++cc) {
      const int src = cc < R ? R + cc : cc - R;
      wp[cc] = on ? xs[(int64_t)(b0 + i) * ncol + src] : T(0);
    }
    T dav[R], g[D];
#pragma unroll
    for (int k = 0; k < R; ++k) dav[k] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = 0;
    for (int a0 = 0; a0 < q; a0 += kFarMom) {
      const int na = min(kFarMom, q - a0);
      __syncthreads();
      for (int e = threadIdx.x; e < kFarMom * 2 * R; e += blockDim.x) {
        const int al = e / (2 * R), cc = e % (2 * R);
        L[al][cc] = al < na ? tlhs[((int64_t)c * q + a0 + al) * ncol + cc] : T(0);
      }
      for (int e = threadIdx.x; e < kFarMom; e += blockDim.x)
        aidx[e] = e < na ? basis_index<T, D>(a0 + e, ip) : 0;
      __syncthreads();
      for (int al = 0; al < na; ++al) {
        const int pk = aidx[al];
        T f[D], df[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          f[d] = F[d][unpack(pk, d)][threadIdx.x];
          df[d] = dF[d][unpack(pk, d)][threadIdx.x];
        }
        T v = f[0];
#pragma unroll
        for (int d = 1; d < D; ++d) v *= f[d];
        T s = 0;
#pragma unroll
        for (int cc = 0; cc < 2 * R; ++cc) s = fmat(wp[cc], L[al][cc], s);
#pragma unroll
        for (int k = 0; k < R; ++k) dav[k] = fmat(v, L[al][R + k], dav[k]);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          T gv = df[d];
#pragma unroll
          for (int e = 0; e < D; ++e)
            if (e != d) gv *= f[e];
          g[d] = fmat(gv, s, g[d]);
        }
      }
    }
    if (on) {
      const int64_t p = tree.perm[b0 + i];
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (k < r) atomic_accumulate(dalpha + (int64_t)k * ld_dalpha + p, (double)dav[k]);
#pragma unroll
      for (int d = 0; d < D; ++d)
        atomic_accumulate(dpoints + d * ld_points + p, (double)(g[d] * inv_size[d]));
    }
  }
}

// Far part of dL/dl: sum over clusters and nodes of G's moments times the
// coupling sums of the dk/dl group (before the downward pass), one partial
// per workgroup over a contiguous range of moments, in a fixed order.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_grad_far_dot(int64_t nmom, int ncol, int r,
                                                      const T *__restrict__ mom,
                                                      const T *__restrict__ tlhs,
                                                      double *__restrict__ part) {
  __shared__ double red[256];
  const int64_t per = (nmom + gridDim.x - 1) / gridDim.x;
  const int64_t m0 = per * blockIdx.x, m1 = min(nmom, m0 + per);
  double acc = 0;
  for (int64_t e = m0 * R + threadIdx.x; e < m1 * R; e += blockDim.x) {
    const int64_t m = e / R;
    const int k = (int)(e % R);
    if (k < r) acc += (double)mom[m * ncol + R + k] * (double)tlhs[m * ncol + 2 * R + k];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dL/dl = 1/2 (sum of the near partials: every unordered pair is met from
// both ends) + sum of the far partials; one workgroup, fixed order.
__global__ __launch_bounds__(256) void k_grad_dl_final(int nnear, int nfar,
                                                       const double *__restrict__ part,
                                                       double *__restrict__ out) {
  __shared__ double red[2][256];
  double a = 0, b = 0;
  for (int e = threadIdx.x; e < nnear; e += blockDim.x) a += part[e];
  for (int e = threadIdx.x; e < nfar; e += blockDim.x) b += part[nnear + e];
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = 0.5 * red[0][0] + red[1][0];
}

template <typename T>
__global__ void k_fill_zero(double *__restrict__ p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

template <typename T, int D, int R>
hipError_t grad_near_kt(const DevTreeView<T> &tree, const DevPatternView &pat,
                        const KernelParams<T> &kp, const GradJob &job, const T *xs,
                        hipStream_t stream) {
  const int base = kp.kt % KT_COUNT;
#define FMCA_GRAD_CASE(K)                                                                     \
  case K:                                                                                     \
    hipLaunchKernelGGL((k_grad_near<T, D, K, R>), dim3(pat.nnear_order), dim3(256), 0, stream, \
                       tree, pat, kp, job.r, job.ncol, xs, job.dalpha, job.ld_dalpha,          \
                       job.dpoints, job.ld_points, job.dl_part);                              \
    break;
  switch (base) {
    FMCA_GRAD_CASE(KT_EXPONENTIAL)
    FMCA_GRAD_CASE(KT_MATERN32)
    FMCA_GRAD_CASE(KT_MATERN52)
    FMCA_GRAD_CASE(KT_GAUSSIAN)
    default:
      hipLaunchKernelGGL((k_grad_near<T, D, -1, R>), dim3(pat.nnear_order), dim3(256), 0, stream,
                         tree, pat, kp, job.r, job.ncol, xs, job.dalpha, job.ld_dalpha,
                         job.dpoints, job.ld_points, job.dl_part);
  }
#undef FMCA_GRAD_CASE
  return hipGetLastError();
}

template <typename T, int D, int R>
hipError_t grad_leaves_dr(const DevTreeView<T> &tree, const DevPatternView &pat,
                          const InterpParams<T> &ip, const KernelParams<T> &kp,
                          const GradJob &job, const T *xs, const T *tlhs, hipStream_t stream) {
  hipLaunchKernelGGL((k_grad_far<T, D, R>), dim3(tree.nleaves), dim3(kFarRows), 0, stream, tree,
                     ip, job.r, job.ncol, xs, tlhs, job.dalpha, job.ld_dalpha, job.dpoints,
                     job.ld_points);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return grad_near_kt<T, D, R>(tree, pat, kp, job, xs, stream);
}

}  // namespace

int grad_far_dot_blocks() { return 256; }

template <typename T>
hipError_t launch_grad_far_dot(const GradJob &job, int64_t nmom, const T *mom, const T *tlhs,
                               hipStream_t stream) {
  const int nb = grad_far_dot_blocks();
  double *part = job.dl_part + job.nnear_parts;
  if (job.R == 8)
    hipLaunchKernelGGL((k_grad_far_dot<T, 8>), dim3(nb), dim3(256), 0, stream, nmom, job.ncol,
                       job.r, mom, tlhs, part);
  else
    hipLaunchKernelGGL((k_grad_far_dot<T, 16>), dim3(nb), dim3(256), 0, stream, nmom, job.ncol,
                       job.r, mom, tlhs, part);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_grad_leaves(const DevTreeView<T> &tree, const DevPatternView &pat,
                              const InterpParams<T> &ip, const KernelParams<T> &kp,
                              const GradJob &job, const T *xs, const T *tlhs,
                              hipStream_t stream) {
  hipError_t e;
#define FMCA_GRAD_DR(DD, RR) \
  e = grad_leaves_dr<T, DD, RR>(tree, pat, ip, kp, job, xs, tlhs, stream)
  if (tree.dim == 1) {
    if (job.R == 8) FMCA_GRAD_DR(1, 8); else FMCA_GRAD_DR(1, 16);
  } else if (tree.dim == 2) {
    if (job.R == 8) FMCA_GRAD_DR(2, 8); else FMCA_GRAD_DR(2, 16);
  } else if (tree.dim == 3) {
    if (job.R == 8) FMCA_GRAD_DR(3, 8); else FMCA_GRAD_DR(3, 16);
  } else {
    return hipErrorInvalidValue;
  }
#undef FMCA_GRAD_DR
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_grad_dl_final, dim3(1), dim3(256), 0, stream, job.nnear_parts,
                     grad_far_dot_blocks(), job.dl_part, job.dl_out);
  return hipGetLastError();
}

hipError_t launch_grad_zero(double *p, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_fill_zero<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     stream, p, n);
  return hipGetLastError();
}

template hipError_t launch_grad_far_dot<float>(const GradJob &, int64_t, const float *,
                                               const float *, hipStream_t);
template hipError_t launch_grad_far_dot<double>(const GradJob &, int64_t, const double *,
                                                const double *, hipStream_t);
template hipError_t launch_grad_leaves<float>(const DevTreeView<float> &, const DevPatternView &,
                                              const InterpParams<float> &,
                                              const KernelParams<float> &, const GradJob &,
                                              const float *, const float *, hipStream_t);
template hipError_t launch_grad_leaves<double>(const DevTreeView<double> &,
                                               const DevPatternView &,
                                               const InterpParams<double> &,
                                               const KernelParams<double> &, const GradJob &,
                                               const double *, const double *, hipStream_t);

}  // namespace fmca_amd
