// The assembled H2 matrix: H2Matrix::computeH2Matrix (H2MatrixBase.h:107-109,
// 222-255) stores every coupling matrix and dense block, and operator*
// (matrix_vector_product_impl.h:17-98) applies them.  Here the coupling
// matrices S_ts = K(xi_t, xi_s), the near blocks K(P_t, P_s) and the leaf
// bases V_t^T are written once into HBM as v_mfma 16x16x4 left-operand tiles
// (StoredBlocks), and an action streams them through the matrix cores instead
// of generating the entries: the two entry phases become HBM-bound
// contractions.  The tree passes are those of the matrix-free action.
//
// Streaming: one workgroup per target (cluster for the coupling, leaf for the
// output); its waves split the target's list and its row tiles, each lane
// reading several k-steps of a tile with one 16-byte load; the right operand
// (moments, rhs rows, or the target's own far-field coefficients for V_t^T)
// is read from global memory (L2) in the instruction's lane order.  One
// workgroup covers a target's whole list and adds its waves' sums in LDS, so
// the result is written once, without atomics, in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"
#include "device_math.hpp"

namespace fmca_amd {
namespace {

using namespace dev;

template <typename T>
struct Mma;

template <>
struct Mma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // result row of accumulator item v in lane group lk (lane / 16)
  static __device__ __forceinline__ int crow(int lk, int v) { return 4 * lk + v; }
};

template <>
struct Mma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int crow(int lk, int v) { return lk + 4 * v; }
};

constexpr int kTB = 256;
constexpr int kWaves = kTB / 64;
constexpr int kMaxRT = 4;  // row tiles per wave

__host__ __device__ inline int pad16(int x) { return (x + 15) & ~15; }

// (i, j) of element e of a tiled R x K block (layout in device.hpp)
template <typename T>
__device__ __forceinline__ void tile_pos(int64_t e, int K, int &i, int &j) {
  constexpr int G = kStoredGroup<T>;
  const int u = (int)(e % G);
  const int64_t r = e / G;
  const int lane = (int)(r & 63);
  const int64_t sg = r >> 6;
  const int ng = K / (4 * G);
  const int ks = (int)(sg % ng) * G + u;
  i = (int)(sg / ng) * 16 + (lane & 15);
  j = ks * 4 + (lane >> 4);
}

// unit coordinate d of interpolation node a (Xi(d, a))
template <typename T, int D>
__device__ __forceinline__ T unit_node(int a, int d, const InterpParams<T> &ip) {
  if (ip.td) return ip.unodes[d * ip.q + a];
  return ip.xi[(basis_index<T, D>(a, ip) >> (8 * d)) & 255];
}

// Far blocks: one workgroup per block b, S[i][j] = K(xi_t(i), xi_s(j)).
template <typename T, int D>
__global__ __launch_bounds__(kTB) void k_assemble_far(DevTreeView<T> rows, DevTreeView<T> cols,
                                                      DevPatternView pat, InterpParams<T> ip,
                                                      KernelParams<T> kp,
                                                      const int32_t *__restrict__ far_tgt,
                                                      int Rq, int Kq, T *__restrict__ far) {
  const int64_t b = blockIdx.x;
  const int t = far_tgt[b], s = pat.far_src[b], q = ip.q;
  T *dst = far + b * (int64_t)Rq * Kq;
  for (int64_t e = threadIdx.x; e < (int64_t)Rq * Kq; e += kTB) {
    int i, j;
    tile_pos<T>(e, Kq, i, j);
    T v = T(0);
    if (i < q && j < q) {
      T r2 = T(0), x[D], y[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        x[d] = fmat(unit_node<T, D>(i, d, ip), rows.bbsize[t * D + d], rows.bbmin[t * D + d]);
        y[d] = fmat(unit_node<T, D>(j, d, ip), cols.bbsize[s * D + d], cols.bbmin[s * D + d]);
        const T dd = x[d] - y[d];
        r2 = fmat(dd, dd, r2);
      }
      if (kp.geodesic) r2 = geodesic_r2<T, D>(x, y, 1);
      v = kernel_r2<-1, T>(r2, kp);
    }
    dst[e] = v;
  }
}

// Leaf blocks: workgroup w < nleaves writes V_t^T of leaf w (row i = point,
// column a = basis function: prod_d basis_1d(u_d(i))[a_d]); workgroup
// nleaves + e writes the near block of near list entry e.
template <typename T, int D>
__global__ __launch_bounds__(kTB) void k_assemble_near(DevTreeView<T> rows, DevTreeView<T> cols,
                                                       DevPatternView pat, InterpParams<T> ip,
                                                       KernelParams<T> kp,
                                                       const int32_t *__restrict__ near_leaf,
                                                       const int64_t *__restrict__ near_off,
                                                       const int64_t *__restrict__ v_off, int Kq,
                                                       T *__restrict__ near) {
  const int64_t w = blockIdx.x;
  const int nl = rows.nleaves;
  if (w < nl) {
    const int t = rows.leaves[w];
    const int bt = rows.begin[t], st = rows.size[t], Rt = pad16(st);
    T *dst = near + v_off[w];
    for (int64_t e = threadIdx.x; e < (int64_t)Rt * Kq; e += kTB) {
      int i, a;
      tile_pos<T>(e, Kq, i, a);
      T v = T(0);
      if (i < st && a < ip.q) {
        const int pk = basis_index<T, D>(a, ip);
        v = T(1);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          T f[kMaxP1];
          basis_1d<T>(rows.unit[(int64_t)d * rows.npts + bt + i], ip, f, 1);
          v *= f[(pk >> (8 * d)) & 255];
        }
      }
      dst[e] = v;
    }
    return;
  }
  const int64_t en = w - nl;
  const int t = rows.leaves[near_leaf[en]], s = pat.near_src[en];
  const int bt = rows.begin[t], st = rows.size[t], Rt = pad16(st);
  const int bs = cols.begin[s], ss = cols.size[s], Ks = stored_kpad<T>(ss);
  T *dst = near + near_off[en];
  for (int64_t e = threadIdx.x; e < (int64_t)Rt * Ks; e += kTB) {
    int i, j;
    tile_pos<T>(e, Ks, i, j);
    T v = T(0);
    if (i < st && j < ss) {
      T r2 = T(0), x[D], y[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        x[d] = rows.pts[(int64_t)d * rows.npts + bt + i];
        y[d] = cols.pts[(int64_t)d * cols.npts + bs + j];
        const T dd = x[d] - y[d];
        r2 = fmat(dd, dd, r2);
      }
      if (kp.geodesic) r2 = geodesic_r2<T, D>(x, y, 1);
      v = kernel_r2<-1, T>(r2, kp);
    }
    dst[e] = v;
  }
}

// acc[m][nt] += A (row tiles rt = r0 + rs m of the block, K columns) *
// B (K x 16 NT): B rows beyond brows are zero, B row r at B + r * nrhs + c0.
// Each lane reads G k-steps of a tile with one 16-byte load; the loads (A and
// B) of the next group are issued before the matrix-core work of the current
// one (PF = 1) or two (PF = 2).
template <typename T, int NT, int PF>
__device__ __forceinline__ void stream_item(const T *__restrict__ A, int K, int nrt, int r0,
                                            int rs, const T *__restrict__ B, int brows,
                                            int nrhs, int c0,
                                            typename Mma<T>::acc_t (&acc)[kMaxRT][NT]) {
  constexpr int G = kStoredGroup<T>;
  typedef T vec_t __attribute__((ext_vector_type(G)));
  const int lane = threadIdx.x & 63;
  const int ng = K / (4 * G);
  const int br = lane >> 4, bc = lane & 15;
  const vec_t *Av = reinterpret_cast<const vec_t *>(A);
  vec_t a[kMaxRT], an[kMaxRT];
  T b[G][NT], bn[G][NT];
  auto load = [&](vec_t(&da)[kMaxRT], T(&db)[G][NT], int g) {
#pragma unroll
    for (int m = 0; m < kMaxRT; ++m) {
      const int rt = r0 + rs * m;
      if (rt < nrt) da[m] = Av[((int64_t)rt * ng + g) * 64 + lane];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int row = (g * G + u) * 4 + br;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int c = c0 + nt * 16 + bc;
        db[u][nt] = (row < brows && c < nrhs) ? B[(int64_t)row * nrhs + c] : T(0);
      }
    }
  };
  // three buffers: the loads run two groups ahead of the matrix-core work
  vec_t a2[kMaxRT];
  T b2[G][NT];
  auto mma = [&](const vec_t(&da)[kMaxRT], const T(&db)[G][NT]) {
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
      for (int m = 0; m < kMaxRT; ++m) {
        if (r0 + rs * m < nrt) {
#pragma unroll
          for (int nt = 0; nt < NT; ++nt)
            acc[m][nt] = Mma<T>::mma(da[m][u], db[u][nt], acc[m][nt]);
        }
      }
  };
  if (PF == 1) {  // one group ahead
    if (ng > 0) load(a, b, 0);
    for (int g = 0; g < ng; ++g) {
      if (g + 1 < ng) load(an, bn, g + 1);
      mma(a, b);
#pragma unroll
      for (int m = 0; m < kMaxRT; ++m) a[m] = an[m];
#pragma unroll
      for (int u = 0; u < G; ++u)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) b[u][nt] = bn[u][nt];
    }
    return;
  }
  if (ng > 0) load(a, b, 0);
  if (ng > 1) load(an, bn, 1);
  for (int g = 0; g < ng; ++g) {
    if (g + 2 < ng) load(a2, b2, g + 2);
    mma(a, b);
#pragma unroll
    for (int m = 0; m < kMaxRT; ++m) {
      a[m] = an[m];
      an[m] = a2[m];
    }
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        b[u][nt] = bn[u][nt];
        bn[u][nt] = b2[u][nt];
      }
  }
}

// One workgroup per target (in the order far_order / near_order, longest
// list first): OUTPUT = false, a target cluster t and its far
// list (tlhs_t = sum S_ts mom_s); OUTPUT = true, a target leaf and [V_t^T,
// then its near blocks] (y_t = V_t^T tlhs_t + sum K(P_t, P_s) x_s), written
// to out in natural (through perm) or cluster order.  The four waves form
// 4 / WS groups of WS waves: the groups take every (4 / WS)-th block of the
// list, the WS waves of a group split the block's row tiles.  The groups'
// sums are added in LDS in a fixed order, so the result is deterministic.
template <typename T, int NT, int WS, int PF, bool OUTPUT>
__global__ __launch_bounds__(kTB) void k_stream(DevTreeView<T> rows, DevTreeView<T> cols,
                                                DevPatternView pat, StoredBlocks<T> st, int q,
                                                int nrhs, const T *__restrict__ xs,
                                                const T *__restrict__ mom, T *__restrict__ tlhs,
                                                double *__restrict__ out, int64_t ldo,
                                                int natural) {
  constexpr int NG = kWaves / WS;  // groups
  constexpr int NC = 16 * NT;      // columns per pass
  extern __shared__ unsigned char smem[];
  T *red = reinterpret_cast<T *>(smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = wave / WS, wig = wave % WS;
  int t, nrt, nitems, bt = 0, sz = 0, k = 0;
  if (OUTPUT) {
    k = st.near_order[blockIdx.x];
    t = rows.leaves[k];
    bt = rows.begin[t];
    sz = rows.size[t];
    nrt = (sz + 15) / 16;
    nitems = 1 + pat.near_ptr[k + 1] - pat.near_ptr[k];
  } else {
    t = st.far_order[blockIdx.x];
    nrt = st.Rq / 16;
    nitems = pat.far_ptr[t + 1] - pat.far_ptr[t];
  }
  const int RP = OUTPUT ? kStoredMaxRows : st.Rq;  // rows of an LDS slot
  for (int c0 = 0; c0 < nrhs; c0 += NC) {
    typename Mma<T>::acc_t acc[kMaxRT][NT];
#pragma unroll
    for (int m = 0; m < kMaxRT; ++m)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[m][nt] = typename Mma<T>::acc_t{0, 0, 0, 0};
    for (int it = grp; it < nitems; it += NG) {
      const T *A, *B;
      int K, brows;
      if (OUTPUT) {
        if (it == 0) {  // V_t^T against the target's far-field coefficients
          A = st.near + st.v_off[k];
          K = st.Kq;
          B = tlhs + (int64_t)t * q * nrhs;
          brows = q;
        } else {
          const int e = pat.near_ptr[k] + it - 1;
          const int s = pat.near_src[e];
          brows = cols.size[s];
          A = st.near + st.near_off[e];
          K = stored_kpad<T>(brows);
          B = xs + (int64_t)cols.begin[s] * nrhs;
        }
      } else {
        const int b = pat.far_ptr[t] + it;
        A = st.far + (int64_t)b * st.Rq * st.Kq;
        K = st.Kq;
        B = mom + (int64_t)pat.far_src[b] * q * nrhs;
        brows = q;
      }
      stream_item<T, NT, PF>(A, K, nrt, wig, WS, B, brows, nrhs, c0, acc);
    }
    // groups' sums: slot s holds one group's partial tile values
    auto put = [&](int slot) {
#pragma unroll
      for (int m = 0; m < kMaxRT; ++m) {
        const int rt = wig + WS * m;
        if (rt >= nrt) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = rt * 16 + Mma<T>::crow(lane >> 4, v);
            red[((int64_t)slot * RP + r) * NC + nt * 16 + (lane & 15)] = acc[m][nt][v];
          }
      }
    };
    auto add = [&](int slot) {
#pragma unroll
      for (int m = 0; m < kMaxRT; ++m) {
        const int rt = wig + WS * m;
        if (rt >= nrt) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = rt * 16 + Mma<T>::crow(lane >> 4, v);
            acc[m][nt][v] += red[((int64_t)slot * RP + r) * NC + nt * 16 + (lane & 15)];
          }
      }
    };
    if (NG == 4) {
      if (grp >= 2) put(grp - 2);
      __syncthreads();
      if (grp < 2) add(grp);
      __syncthreads();
      if (grp == 1) put(0);
      __syncthreads();
      if (grp == 0) add(0);
    } else if (NG == 2) {
      if (grp == 1) put(0);
      __syncthreads();
      if (grp == 0) add(0);
    }
    if (grp == 0) {
#pragma unroll
      for (int m = 0; m < kMaxRT; ++m) {
        const int rt = wig + WS * m;
        if (rt >= nrt) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int c = c0 + nt * 16 + (lane & 15);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = rt * 16 + Mma<T>::crow(lane >> 4, v);
            if (c >= nrhs) continue;
            if (OUTPUT) {
              if (r < sz) {
                const int64_t row = natural ? (int64_t)rows.perm[bt + r] : (int64_t)(bt + r);
                out[row + (int64_t)c * ldo] = (double)acc[m][nt][v];
              }
            } else if (r < q) {
              tlhs[((int64_t)t * q + r) * nrhs + c] = acc[m][nt][v];
            }
          }
        }
      }
    }
    __syncthreads();  // the next pass reuses the LDS slots
  }
}

template <typename T, int D>
hipError_t assemble_d(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                      const DevPatternView &pat, const InterpParams<T> &ip,
                      const KernelParams<T> &kp, int64_t nfar, const int32_t *far_tgt,
                      int64_t nnear, const int32_t *near_leaf, T *far, T *near,
                      const int64_t *near_off, const int64_t *v_off, int Rq, int Kq,
                      hipStream_t stream) {
  if (nfar > 0)
    hipLaunchKernelGGL((k_assemble_far<T, D>), dim3((unsigned)nfar), dim3(kTB), 0, stream, rows,
                       cols, pat, ip, kp, far_tgt, Rq, Kq, far);
  const int64_t nw = rows.nleaves + nnear;
  hipLaunchKernelGGL((k_assemble_near<T, D>), dim3((unsigned)nw), dim3(kTB), 0, stream, rows,
                     cols, pat, ip, kp, near_leaf, near_off, v_off, Kq, near);
  return hipGetLastError();
}
}  // namespace

template <typename T>
hipError_t launch_assemble(const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                           const DevPatternView &pat, const InterpParams<T> &ip,
                           const KernelParams<T> &kp, int64_t nfar, const int32_t *far_tgt,
                           int64_t nnear, const int32_t *near_leaf, T *far, T *near,
                           const int64_t *near_off, const int64_t *v_off, int Rq, int Kq,
                           hipStream_t stream) {
  if (Rq > kStoredMaxRows || nfar > INT32_MAX || rows.nleaves + nnear > INT32_MAX)
    return hipErrorInvalidValue;
  switch (rows.dim) {
    case 1:
      return assemble_d<T, 1>(rows, cols, pat, ip, kp, nfar, far_tgt, nnear, near_leaf, far,
                              near, near_off, v_off, Rq, Kq, stream);
    case 2:
      return assemble_d<T, 2>(rows, cols, pat, ip, kp, nfar, far_tgt, nnear, near_leaf, far,
                              near, near_off, v_off, Rq, Kq, stream);
    case 3:
      return assemble_d<T, 3>(rows, cols, pat, ip, kp, nfar, far_tgt, nnear, near_leaf, far,
                              near, near_off, v_off, Rq, Kq, stream);
  }
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_stored(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const DevPatternView &pat, const StoredBlocks<T> &st, int q, int nrhs,
                         const T *xs, const T *mom, T *tlhs, double *out, int64_t ldo,
                         bool natural, hipStream_t stream) {
  if (st.Rq > kStoredMaxRows / 2) return hipErrorInvalidValue;
  const int nat = natural ? 1 : 0;
  auto lds = [&](int slots, int rp, int nt) { return (size_t)slots * rp * 16 * nt * sizeof(T); };
  // Two column tiles: loads two groups ahead, and the coupling (q <= 128: 8
  // row tiles) spreads its list over 2 two-wave groups.  One column tile:
  // one group ahead, all four waves on one block at a time (measured faster
  // there: fewer registers, more waves per SIMD).  The output keeps all four
  // waves on one block (16 row tiles).
#define FMCA_STREAM(NTV, WSC, PFV)                                                           \
  if (output)                                                                                \
    hipLaunchKernelGGL((k_stream<T, NTV, 4, PFV, true>), dim3(rows.nleaves), dim3(kTB), 0,   \
                       stream, rows, cols, pat, st, q, nrhs, xs, mom, tlhs, out, ldo, nat);  \
  else                                                                                       \
    hipLaunchKernelGGL((k_stream<T, NTV, WSC, PFV, false>), dim3(rows.ncl), dim3(kTB),       \
                       lds(WSC == 4 ? 0 : 1, st.Rq, NTV), stream, rows, cols, pat, st, q,    \
                       nrhs, xs, mom, tlhs, out, ldo, nat);
  if (nrhs <= 16) {
    FMCA_STREAM(1, 4, 1)
  } else {
    FMCA_STREAM(2, 2, 2)
  }
#undef FMCA_STREAM
  return hipGetLastError();
}

template hipError_t launch_assemble<double>(const DevTreeView<double> &,
                                            const DevTreeView<double> &, const DevPatternView &,
                                            const InterpParams<double> &,
                                            const KernelParams<double> &, int64_t,
                                            const int32_t *, int64_t, const int32_t *, double *,
                                            double *, const int64_t *, const int64_t *, int, int,
                                            hipStream_t);
template hipError_t launch_assemble<float>(const DevTreeView<float> &, const DevTreeView<float> &,
                                           const DevPatternView &, const InterpParams<float> &,
                                           const KernelParams<float> &, int64_t, const int32_t *,
                                           int64_t, const int32_t *, float *, float *,
                                           const int64_t *, const int64_t *, int, int,
                                           hipStream_t);
template hipError_t launch_stored<double>(bool, const DevTreeView<double> &,
                                          const DevTreeView<double> &, const DevPatternView &,
                                          const StoredBlocks<double> &, int, int, const double *,
                                          const double *, double *, double *, int64_t, bool,
                                          hipStream_t);
template hipError_t launch_stored<float>(bool, const DevTreeView<float> &,
                                         const DevTreeView<float> &, const DevPatternView &,
                                         const StoredBlocks<float> &, int, int, const float *,
                                         const float *, float *, double *, int64_t, bool,
                                         hipStream_t);

}  // namespace fmca_amd
