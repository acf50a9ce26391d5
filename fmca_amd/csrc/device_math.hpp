// Device helpers shared by the HIP kernels: scalar math of the kernel-entry
// loops, kernel functions, barycentric Lagrange factors.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device.hpp"

namespace fmca_amd {
namespace dev {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Scalar helpers.
//
// FP64: exp and sqrt are written out for the kernel-entry loops, which only
// ever see finite arguments (exp of -s*r <= 0, sqrt of a squared distance):
//   exp(x)  = 2^k e^f with k = rint(x log2 e) and f = x - k ln2, ln2 split
//             into a high and a low part (Cody & Waite) so that f is exact to
//             working precision; e^f from its Taylor polynomial of degree 12
//             (|f| <= ln2/2, truncation error < 2e-16 relative), evaluated
//             with Estrin's scheme.
//   sqrt(x) = hardware reciprocal-square-root seed (v_rsq_f64) refined by a
//             Goldschmidt step and a Newton correction of the root; x = 0
//             returns 0.
// FP32 uses the hardware transcendental instructions (v_exp_f32, v_sqrt_f32).
template <typename T>
struct Math;

template <>
struct Math<double> {
  static __device__ __forceinline__ double exp_(double x) {
    const double k = __builtin_rint(x * 1.4426950408889634074);
    double f = __builtin_fma(-k, 6.93147180369123816490e-01, x);
    f = __builtin_fma(-k, 1.90821492927058770002e-10, f);
    // Estrin evaluation of sum_{j<=12} f^j / j! (fewer dependent steps than
    // Horner and no accumulator copies)
    const double f2 = f * f, f4 = f2 * f2, f8 = f4 * f4;
    const double q0 = __builtin_fma(f, 1.0, 1.0);
    const double q1 = __builtin_fma(f, 1.66666666666666666667e-01, 0.5);
    const double q2 = __builtin_fma(f, 8.33333333333333333333e-03, 4.16666666666666666667e-02);
    const double q3 = __builtin_fma(f, 1.98412698412698412698e-04, 1.38888888888888888889e-03);
    const double q4 = __builtin_fma(f, 2.75573192239858906526e-06, 2.48015873015873015873e-05);
    const double q5 = __builtin_fma(f, 2.50521083854417187751e-08, 2.75573192239858906526e-07);
    const double r0 = __builtin_fma(f2, q1, q0);
    const double r1 = __builtin_fma(f2, q3, q2);
    const double r2 = __builtin_fma(f2, q5, q4);
    const double s0 = __builtin_fma(f4, r1, r0);
    const double s1 = __builtin_fma(f4, 2.08767569878680989792e-09, r2);
    const double p = __builtin_fma(f8, s1, s0);
    const double r = __builtin_ldexp(p, (int)k);
    return x < -745.0 ? 0.0 : r;
  }
  static __device__ __forceinline__ double sqrt_(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x > 0.0 ? g : 0.0;
  }
  static __device__ __forceinline__ double div_(double a, double b) { return a / b; }
  static __device__ __forceinline__ double log_(double x) { return log(x); }
  static __device__ __forceinline__ double eps() { return 2.220446049250313e-16; }
};

template <>
struct Math<float> {
  static __device__ __forceinline__ float exp_(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  }
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float log_(float x) { return logf(x); }
  // a / b through the hardware reciprocal (about 1 ulp; the FP32 form's
  // entries carry more error than that already)
  static __device__ __forceinline__ float div_(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float eps() { return 1.1920928955078125e-07f; }
};

__device__ __forceinline__ double fmat(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Squared geodesic distance of FMCA's "GEODESIC" distance type
// (CovarianceKernel.h:130-133, SphereClusterTree.h:65-69): r = acos of the
// dot product clamped to [-1, 1], returned as r * r for kernel_r2.  The dot
// product is x[0] y[0] followed by one fma per further axis, in axis order,
// with `ys` the stride between the axes of y: near the diagonal acos
// amplifies a rounding of the dot product by 1 / r, so every evaluation of
// the product (and the oracle) forms it in this one order.
template <typename T, int D>
__device__ __forceinline__ T geodesic_r2(const T *x, const T *y, int ys) {
  T dot = x[0] * y[0];
#pragma unroll
  for (int d = 1; d < D; ++d) dot = fmat(x[d], y[d * ys], dot);
  const T c = dot > T(1) ? T(1) : (dot < T(-1) ? T(-1) : dot);
  const T r = acos(c);
  return r * r;
}

// exp(-u * s): in FP32 one hardware exp2 of u * (s log2 e), the product
// formed on the host (kp.sl2); in FP64 the exp above.
template <typename T>
__device__ __forceinline__ T exp_neg_scaled(T u, const KernelParams<T> &kp) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_exp2f(-u * kp.sl2);
  else
    return Math<T>::exp_(-u * kp.s);
}

// Kernel value from the squared distance (CovarianceKernel.h:56-116) with the
// scales of KernelParams.  KT < 0 selects the type at run time (uniform
// branch) for the less common kernels.
template <int KT, typename T>
__device__ __forceinline__ T kernel_r2(T r2, const KernelParams<T> &kp) {
  using M = Math<T>;
  const int kt = KT >= 0 ? KT : kp.kt;
  switch (kt) {
    case KT_EXPONENTIAL:
      return exp_neg_scaled(M::sqrt_(r2), kp);
    case KT_MATERN32: {
      const T r = M::sqrt_(r2), a = r * kp.s;
      return (T(1) + a) * exp_neg_scaled(r, kp);
    }
    case KT_MATERN52: {
      const T r = M::sqrt_(r2), a = r * kp.s;
      return (T(1) + a + a * a * T(1. / 3.)) * exp_neg_scaled(r, kp);
    }
    case KT_MATERN72: {
      const T x = M::sqrt_(r2) * kp.inv_l;
      return (T(1) + T(2.6457513110645905905) * x + T(14. / 5) * x * x +
              T(7. * 2.6457513110645905905 / 15.) * x * x * x) *
             M::exp_(-T(2.6457513110645905905) * x);
    }
    case KT_MATERN92: {
      const T x = M::sqrt_(r2) * kp.inv_l;
      return (T(1) + T(3) * x + T(27. / 7.) * x * x + T(18. / 7.) * x * x * x +
              T(27. / 35.) * x * x * x * x) *
             M::exp_(-T(3) * x);
    }
    case KT_GAUSSIAN:
      return exp_neg_scaled(-r2, kp);
    case KT_INVMULTIQUADRIC:
      return T(1) / M::sqrt_(r2 * kp.inv_l * kp.inv_l + kp.c2);
    case KT_MULTIQUADRIC:
      return M::sqrt_(r2 * kp.inv_l * kp.inv_l + kp.c2);
    case KT_BIHARMONIC2D: {
      const T r = M::sqrt_(r2), x = r * kp.inv_l;
      return r < M::eps() ? T(0) : M::log_(x) * x * x;
    }
    case KT_BIHARMONIC3D:
      return M::sqrt_(r2) * kp.inv_l;
    case KT_TRIHARMONIC3D: {
      const T x = M::sqrt_(r2) * kp.inv_l;
      return x * x * x;
    }
    case KT_TPS2D: {
      const T r = M::sqrt_(r2);
      return r > T(0) ? r * r * M::log_(r) : T(0);
    }
  }
  // derivatives: dk/dl (variant 1) and phi = k'(r)/r (variant 2).  For every
  // kernel of the form f(r/l): dk/dl = -(r^2 / l) phi(r).
  const int base = kt % KT_COUNT, var = kt / KT_COUNT;
  const T il = kp.inv_l, il2 = il * il;
  T phi = 0;
  switch (base) {
    case KT_EXPONENTIAL: {
      const T r = M::sqrt_(r2), e = M::exp_(-r * il);
      if (var == KV_DL) return r * il2 * e;
      phi = r > T(0) ? -M::div_(il, r) * e : T(0);
      break;
    }
    case KT_MATERN32: {
      const T x = M::sqrt_(r2) * kp.s;
      phi = -kp.s * kp.s * M::exp_(-x);
      break;
    }
    case KT_MATERN52: {
      const T x = M::sqrt_(r2) * kp.s;
      phi = -kp.s * kp.s * T(1. / 3.) * (T(1) + x) * M::exp_(-x);
      break;
    }
    case KT_MATERN72: {
      const T x = M::sqrt_(r2) * kp.s;
      phi = -kp.s * kp.s * T(1. / 15.) * (T(3) + T(3) * x + x * x) * M::exp_(-x);
      break;
    }
    case KT_MATERN92: {
      const T x = M::sqrt_(r2) * kp.s;
      phi = -kp.s * kp.s * T(1. / 105.) * (T(15) + x * (T(15) + x * (T(6) + x))) * M::exp_(-x);
      break;
    }
    case KT_GAUSSIAN:
      phi = -il2 * M::exp_(r2 * kp.s);
      break;
    case KT_INVMULTIQUADRIC: {
      const T u = r2 * il2 + kp.c2;
      phi = -il2 / (u * M::sqrt_(u));
      break;
    }
    case KT_MULTIQUADRIC:
      phi = il2 / M::sqrt_(r2 * il2 + kp.c2);
      break;
    case KT_BIHARMONIC2D: {
      const T r = M::sqrt_(r2);
      phi = r < M::eps() ? T(0) : il2 * (T(1) + T(2) * M::log_(r * il));
      break;
    }
    case KT_BIHARMONIC3D: {
      const T r = M::sqrt_(r2);
      if (var == KV_DL) return -r * il2;
      phi = r > T(0) ? il / r : T(0);
      break;
    }
    case KT_TRIHARMONIC3D:
      phi = T(3) * M::sqrt_(r2) * il2 * il;
      break;
    case KT_TPS2D: {
      if (var == KV_DL) return T(0);
      const T r = M::sqrt_(r2);
      phi = r > T(0) ? T(2) * M::log_(r) + T(1) : T(0);
      break;
    }
  }
  return var == KV_DL ? -r2 * il * phi : phi;
}

// Value, dk/dl and phi of one entry (ev[KV_VALUE], ev[KV_DL], ev[KV_PHI]) for
// the base kernel KT (kp.kt % KT_COUNT at run time when KT < 0).  Evaluated
// together, the three share the distance, square root and exponential.  A
// variant outside kp.vmask (uniform over the launch) is not needed by any
// column tile: where it costs extra work (a division, a separate evaluation)
// it is skipped and left 0.
template <int KT, typename T>
__device__ __forceinline__ void kernel_r2_all(T r2, const KernelParams<T> &kp, T ev[3]) {
  using M = Math<T>;
  if constexpr (KT == KT_EXPONENTIAL) {
    // s = 1 / l: value e, dk/dl = r e / l^2, phi = -e / (l r)
    const T r = M::sqrt_(r2), e = exp_neg_scaled(r, kp);
    ev[KV_VALUE] = e;
    ev[KV_DL] = r * (kp.s * kp.s) * e;
    ev[KV_PHI] = T(0);
    if (kp.vmask & (1 << KV_PHI)) ev[KV_PHI] = r > T(0) ? -M::div_(kp.s, r) * e : T(0);
  } else if constexpr (KT == KT_MATERN32 || KT == KT_MATERN52 || KT == KT_GAUSSIAN) {
    // the shared exponential; dk/dl = -(r^2 / l) phi for these f(r / l)
    T phi;
    if constexpr (KT == KT_GAUSSIAN) {
      const T e = exp_neg_scaled(-r2, kp);
      ev[KV_VALUE] = e;
      phi = -(kp.inv_l * kp.inv_l) * e;
    } else {
      const T r = M::sqrt_(r2), x = r * kp.s, e = exp_neg_scaled(r, kp);
      if constexpr (KT == KT_MATERN32) {
        ev[KV_VALUE] = (T(1) + x) * e;
        phi = -kp.s * kp.s * e;
      } else {
        ev[KV_VALUE] = (T(1) + x + x * x * T(1. / 3.)) * e;
        phi = -kp.s * kp.s * T(1. / 3.) * (T(1) + x) * e;
      }
    }
    ev[KV_PHI] = phi;
    ev[KV_DL] = -r2 * kp.inv_l * phi;
  } else if constexpr (KT >= 0) {
    ev[KV_VALUE] = ev[KV_DL] = ev[KV_PHI] = T(0);
    if (kp.vmask & (1 << KV_VALUE)) ev[KV_VALUE] = kernel_r2<KT, T>(r2, kp);
    if (kp.vmask & (1 << KV_DL)) ev[KV_DL] = kernel_r2<KT + KT_COUNT * KV_DL, T>(r2, kp);
    if (kp.vmask & (1 << KV_PHI)) ev[KV_PHI] = kernel_r2<KT + KT_COUNT * KV_PHI, T>(r2, kp);
  } else {
    KernelParams<T> k = kp;
    const int base = kp.kt % KT_COUNT;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      k.kt = base + KT_COUNT * v;
      ev[v] = (kp.vmask & (1 << v)) ? kernel_r2<-1, T>(r2, k) : T(0);
    }
  }
}

// Adds v to *p with a relaxed device-scope hardware atomic: the tasks of a
// split target accumulate into one (zeroed) output.
template <typename T>
__device__ __forceinline__ void atomic_accumulate(T *p, T v) {
  unsafeAtomicAdd(p, v);
}

// ev[v] for a uniform variant index v
template <typename T>
__device__ __forceinline__ T pick_variant(const T ev[3], int v) {
  return v == KV_VALUE ? ev[KV_VALUE] : (v == KV_DL ? ev[KV_DL] : ev[KV_PHI]);
}

// Barycentric Lagrange factors l_a(u), a < P1, written with a stride
// (TensorProductInterpolator.h:64-81; a factor within eps of its node is 1).
template <typename T>
__device__ __forceinline__ void lagrange_1d(T u, const InterpParams<T> &ip, T *out, int stride) {
  T s = 0;
  for (int j = 0; j < ip.P1; ++j) s += ip.w[j] / (u - ip.xi[j]);
  for (int a = 0; a < ip.P1; ++a) {
    const T d = u - ip.xi[a];
    out[a * stride] = fabs(d) > Math<T>::eps() ? ip.w[a] / d / s : T(1);
  }
}

// Derivatives l_a'(u), a < P1, of the barycentric Lagrange factors (the
// values stay lagrange_1d's).  Away from the nodes l_a' = l_a (S2 / S1 -
// 1 / (u - xi_a)) with S1 = sum_j w_j / (u - xi_j), S2 = sum_j w_j / (u - xi_j)^2
// (the derivative of the barycentric quotient).  Within eps of node m, where
// lagrange_1d takes the factors as delta_am, they are the entries of the
// barycentric differentiation matrix: l_a'(xi_m) = (w_a / w_m) / (xi_m - xi_a)
// for a != m and l_m'(xi_m) = -sum_{a != m} l_a'(xi_m).
template <typename T>
__device__ __forceinline__ void lagrange_1d_deriv(T u, const InterpParams<T> &ip, T *der,
                                                  int stride) {
  int hit = -1;
  T s1 = 0, s2 = 0;
  for (int j = 0; j < ip.P1; ++j) {
    const T d = u - ip.xi[j];
    if (fabs(d) <= Math<T>::eps()) hit = j;
    const T wd = ip.w[j] / d;
    s1 += wd;
    s2 += wd / d;
  }
  if (hit < 0) {
    const T rho = s2 / s1;
    for (int a = 0; a < ip.P1; ++a) {
      const T d = u - ip.xi[a];
      der[a * stride] = ip.w[a] / d / s1 * (rho - T(1) / d);
    }
  } else {
    T sum = 0;
    for (int a = 0; a < ip.P1; ++a) {
      if (a == hit) continue;
      const T da = (ip.w[a] / ip.w[hit]) / (ip.xi[hit] - ip.xi[a]);
      der[a * stride] = da;
      sum += da;
    }
    der[hit * stride] = -sum;
  }
}

// L2-normalised shifted Legendre polynomials sqrt(2a+1) P_a(2u-1), a < P1,
// by the three-term recurrence (evalPolynomials.h:19-40), written with a
// stride.
template <typename T>
__device__ __forceinline__ void legendre_1d(T u, int P1, T *out, int stride) {
  const T x = 2 * u - 1;
  T prev = 0, cur = 1;
  out[0] = T(1);
  for (int a = 1; a < P1; ++a) {
    const T next = T(2 * a - 1) / T(a) * x * cur - T(a - 1) / T(a) * prev;
    prev = cur;
    cur = next;
    out[a * stride] = next * sqrt(T(2 * a + 1));
  }
}

// 1-D factors of the leaf basis at unit coordinate u: Lagrange factors
// (tensor scheme) or Legendre polynomials (total-degree scheme).
template <typename T>
__device__ __forceinline__ void basis_1d(T u, const InterpParams<T> &ip, T *out, int stride) {
  if (ip.td)
    legendre_1d<T>(u, ip.P1, out, stride);
  else
    lagrange_1d<T>(u, ip, out, stride);
}

// Packed multi-index a_0 | a_1 << 8 | a_2 << 16 of basis function alpha.
template <typename T, int D>
__device__ __forceinline__ int basis_index(int alpha, const InterpParams<T> &ip) {
  if (ip.td) return ip.midx[alpha];
  int pk = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    pk |= (alpha % ip.P1) << (8 * d);
    alpha /= ip.P1;
  }
  return pk;
}

// Kernels smooth at r = 0 take their squared distances from the matrix
// cores: for 16 sources and 16 targets one v_mfma_f32_16x16x4_f32 (BF16 forms) of the
// source rows [-2 y', |y'|^2] (k-slots 0..D-1 and 3) with the target columns
// [x', 1], accumulated onto |x'|^2 (the accumulator input, one value per
// target column), gives r^2 = |y'|^2 - 2 x'.y' + |x'|^2 (x', y' centred on
// the target leaf) with no VALU work.  In FP32 the cancellation leaves an
// absolute error of a few eps |y'|^2 in r^2, of either sign; the entries take
// |r^2| (a free source modifier), which moves k by at most |dk/d(r^2)| times
// that error -- bounded for these kernels, unbounded at r -> 0 for kernels
// with a cusp there (EXPONENTIAL and the run-time kernels keep the direct
// differences in the near field).
template <int KT>
constexpr bool kGramNear = KT == KT_MATERN32 || KT == KT_MATERN52 || KT == KT_GAUSSIAN;

// The Gram distances of these kernels are taken in scaled coordinates
// sigma x, so that the exponential needs no multiply of its own:
//   MATERN32/52: FP32 sigma = s log2(e), r' = sigma r, s r = r' ln 2,
//                k = p(r' ln 2) exp2(-r');
//                FP64 sigma = s, k = p(r') exp(-r')
//   GAUSSIAN:    FP32 sigma^2 = log2(e) / (2 l^2), k = exp2(-r'^2);
//                FP64 sigma^2 = 1 / (2 l^2), k = exp(-r'^2)
// (the negation is a source modifier).  Other kernels take sigma = 1 and
// kernel_r2.
template <int KT, typename T>
__device__ __forceinline__ T gram_sigma(const KernelParams<T> &kp) {
  if constexpr (sizeof(T) == 4) {
    if constexpr (KT == KT_MATERN32 || KT == KT_MATERN52)
      return kp.sl2;
    else if constexpr (KT == KT_GAUSSIAN)
      return __builtin_sqrtf(-kp.sl2);
    else
      return 1.f;
  } else {
    if constexpr (KT == KT_MATERN32 || KT == KT_MATERN52)
      return kp.s;
    else if constexpr (KT == KT_GAUSSIAN)
      return __builtin_sqrt(-kp.s);
    else
      return 1.0;
  }
}

// Kernel value from the squared distance in the coordinates of gram_sigma.
template <int KT, typename T>
__device__ __forceinline__ T gram_kernel(T r2s, const KernelParams<T> &kp) {
  if constexpr (sizeof(T) == 4) {
    constexpr float kLn2 = 0.693147180559945309f;
    if constexpr (KT == KT_MATERN32) {
      const float r = __builtin_amdgcn_sqrtf(r2s);
      return fmat(r, kLn2, 1.f) * __builtin_amdgcn_exp2f(-r);
    } else if constexpr (KT == KT_MATERN52) {
      const float r = __builtin_amdgcn_sqrtf(r2s);
      // 1 + a + a^2 / 3 with a = r ln 2
      return fmat(r, fmat(r, kLn2 * kLn2 / 3.f, kLn2), 1.f) * __builtin_amdgcn_exp2f(-r);
    } else if constexpr (KT == KT_GAUSSIAN) {
      return __builtin_amdgcn_exp2f(-r2s);
    } else {
      return kernel_r2<KT, float>(r2s, kp);
    }
  } else {
    using M = Math<double>;
    if constexpr (KT == KT_MATERN32) {
      const double a = M::sqrt_(r2s);
      return (1.0 + a) * M::exp_(-a);
    } else if constexpr (KT == KT_MATERN52) {
      const double a = M::sqrt_(r2s);
      return fmat(a, fmat(a, 1. / 3., 1.0), 1.0) * M::exp_(-a);
    } else if constexpr (KT == KT_GAUSSIAN) {
      return M::exp_(-r2s);
    } else {
      return kernel_r2<KT, double>(r2s, kp);
    }
  }
}

}  // namespace dev
}  // namespace fmca_amd
