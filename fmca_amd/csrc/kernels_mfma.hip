// Matrix-core (MFMA) forms of the two kernel-entry phases for several rhs.
//
// Both phases are contractions whose left operand is generated on the fly:
//   leaf output   Y_t (rows x rhs) = V_t^T l_t + sum_near K(P_t, P_s) X_s
//   coupling      l_t (q x rhs)    = sum_far K(xi_t, xi_s) m_s
// They run on v_mfma_{f32,f64}_16x16x4 tiles.  In that instruction lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; its four accumulator values are
// C[4(l/16)+v][l%16] for f32 but C[(l/16)+4v][l%16] for f64 (v = 0..3; both
// measured on gfx950 with lane-coded operands).  Each lane therefore evaluates exactly the one kernel entry (or
// leaf-basis value) that its A slot needs -- entries go from VALU registers
// straight into the matrix core, with no LDS round trip and no shuffles --
// while the right operand (rhs rows / moments) is staged in LDS once per
// workgroup and read lane-parallel.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "device.hpp"
#include "device_math.hpp"

namespace fmca_amd {

namespace {

using namespace dev;

template <typename T>
struct Mfma;

template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of the 16x16 result tile held in accumulator item v of lane group lk
  static __device__ __forceinline__ int crow(int lk, int v) { return 4 * lk + v; }
};

template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int crow(int lk, int v) { return lk + 4 * v; }
};

constexpr int kWaves = 4;         // waves per workgroup
constexpr int kRT = 4;            // 16-row tiles per wave -> 256 rows per pass
constexpr int kRows = kWaves * kRT * 16;
// LDS row strides padded by 16 elements: the four 16-lane groups of a wave
// read rows k..k+3 of a staged operand, and a stride = 16 (mod 32) puts them
// on disjoint banks (for 32 and for 64 banks)
constexpr int kRowsP = kRows + 16;
// source points / nodes staged per LDS chunk: 128 for FP32 32-column passes
// (one chunk per far source at q <= 128, half the barriers; 1.3 % faster at
// n = 1e6, 32 rhs), 64 otherwise
template <typename T, int NT>
constexpr int kChunkOf = (sizeof(T) == 4 && NT == 2) ? 128 : 64;
// rows of the leaf output's staged source block: D coordinates, or in FP64
// the four k-slots of the distance rows
template <typename T, int D>
constexpr int kYsRows = (sizeof(T) == 8 && D < 4) ? 4 : D;

// Occupancy target (waves per SIMD) of the MFMA kernels: FP32 fits 4 waves
// without spilling; FP64 (twice the registers per value) and the
// three-variant run-time switch keep the default.
template <typename T, int NT, int KT, bool MULTI>
constexpr int kMinWaves = (sizeof(T) == 4 && NT <= 2 && !(MULTI && KT < 0)) ? 3 : 1;
// The FP64 kernels with distances from the matrix cores ask for 2 waves per
// SIMD (256 registers, a few spills at 32 columns): leaf output 45.3 ms at
// n = 1e6, 32 rhs, against 56.7 ms with the default target
template <typename T, int NT, int KT, bool MULTI>
constexpr int kMinWavesGram =
    (sizeof(T) == 8 && !MULTI && kGramNear<KT>) ? 2 : kMinWaves<T, NT, KT, MULTI>;

// Packed tensor index of every moment: a0 | a1 << 8 | a2 << 16.
__device__ __forceinline__ int unpack(int p, int d) { return (p >> (8 * d)) & 0xff; }

// ---------------------------------------------------------------------------
// Leaf output on MFMA: one workgroup per target leaf, 4 waves, each wave owns
// row tiles {wave, wave+4, wave+8, wave+12} of the current 256-row pass and
// 16*NT rhs columns.
// MULTI: the rhs columns carry different kernel variants (kp.tv per 16-column
// tile); every entry is evaluated once for all three variants.
template <typename T, int D, int KT, int NT, bool MULTI>
__global__ __launch_bounds__(256, (kMinWavesGram<T, NT, KT, MULTI>)) void k_leaf_output_mfma(
    DevTreeView<T> rows, DevTreeView<T> cols, DevPatternView pat, InterpParams<T> ip,
    KernelParams<T> kp, int nrhs, const T *__restrict__ xs, const T *__restrict__ tlhs,
    double *__restrict__ out, int64_t ldo, int natural) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int NC = 16 * NT;
  constexpr int NCP = NC + 16;  // padded LDS row stride
  constexpr int kChunk = kChunkOf<T, NT>;
  // FP64 near field of the kernels smooth at r = 0: squared distances from
  // the matrix cores (kGramNear), 16 sources per distance MFMA
  constexpr bool kGram = sizeof(T) == 8 && !MULTI && kGramNear<KT> && D <= 3;
  extern __shared__ unsigned char smem_raw[];
  const int P1 = ip.P1, q = ip.q;
  const int qpad = (q + 3) & ~3;
  T *F = reinterpret_cast<T *>(smem_raw);              // [D][P1][kRowsP] factors of the rows
  int *aidx = reinterpret_cast<int *>(F + D * P1 * kRowsP);  // [qpad] packed indices
  T *U = reinterpret_cast<T *>(aidx + ((qpad + 3) & ~3));    // union region
  T *Ms = U;                                            // [kChunk][NCP] coefficients
  T *Ys = U;  // [D][kChunk] source points, or kGram: [kChunk][4] distance rows
  T *Xs = U + kYsRows<T, D> * kChunk;                   // [kChunk][NCP] source rhs
  const int lane = threadIdx.x & 63;
  // wave index as a scalar: the row-tile guards below are then uniform branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int task = pat.near_order[blockIdx.x];
  if (task < 0) return;
  const int li = pat.near_task[4 * task], n0 = pat.near_task[4 * task + 1],
            n1 = pat.near_task[4 * task + 2], flags = pat.near_task[4 * task + 3];
  const bool split = (flags & kTaskSplit) != 0, do_far = (flags & kTaskFar) != 0;
  const int c = rows.leaves[li];
  const int b0 = rows.begin[c], sz = rows.size[c];
  for (int e = threadIdx.x; e < qpad; e += blockDim.x)
    aidx[e] = basis_index<T, D>(e < q ? e : 0, ip);
  // kGram: coordinates centred on the leaf's bounding box and scaled by sigma
  T gctr[D] = {}, gsig = T(1);
  if constexpr (kGram) {
    gsig = gram_sigma<KT>(kp);
#pragma unroll
    for (int d = 0; d < D; ++d)
      gctr[d] = rows.bbmin[c * D + d] + T(0.5) * rows.bbsize[c * D + d];
  }
  for (int pbase = 0; pbase < sz; pbase += kRows) {
    // 1-D basis factors of the rows of this pass (thread <-> row)
    __syncthreads();
    {
      const int i = pbase + threadIdx.x;
      if (i < sz) {
#pragma unroll
        for (int d = 0; d < D; ++d)
          basis_1d<T>(rows.unit[(int64_t)d * rows.npts + b0 + i], ip,
                      F + (d * P1) * kRowsP + threadIdx.x, kRowsP);
      } else {
        for (int d = 0; d < D; ++d)
          for (int a = 0; a < P1; ++a) F[(d * P1 + a) * kRowsP + threadIdx.x] = T(0);
      }
    }
    // this lane's A rows: tile (wave + 4 rt), row lr
    T x[kRT][D];
    bool tile_on[kRT];
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt) {
      const int tile_row0 = pbase + (wave + kWaves * rt) * 16;
      tile_on[rt] = tile_row0 < sz;
      const int i = tile_row0 + lr;
#pragma unroll
      for (int d = 0; d < D; ++d)
        x[rt][d] = i < sz ? rows.pts[(int64_t)d * rows.npts + b0 + i] : T(0);
    }
    // kGram: the distance MFMA's right operand, k-slot lk of target lr of
    // each row tile ([x', 1], x' = sigma (x - centre of the leaf)), and its
    // accumulator input |x'|^2
    T gx[kRT], gxx[kRT];
    if constexpr (kGram) {
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        T xx = 0, sel = lk == 3 ? T(1) : T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const T xd = gsig * (x[rt][d] - gctr[d]);
          xx = fmat(xd, xd, xx);
          sel = lk == d ? xd : sel;
        }
        gx[rt] = sel;
        gxx[rt] = xx;
      }
    }
    for (int rc = 0; rc < nrhs; rc += NC) {
      acc_t acc[kRT][NT];
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = acc_t{0, 0, 0, 0};
      int tv[NT];  // kernel variant of each column tile of this pass (MULTI)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) tv[nt] = kp.tv[min((rc >> 4) + nt, kMaxVariantGroups - 1)];
      // column tiles of this pass that hold rhs columns (the rest are skipped)
      const int ntl = min(NT, (nrhs - rc + 15) >> 4);
      // far field: A = V^T (rows x q), B = l_t (q x NC), K-steps of 4 moments
      // (applied by the first task of the leaf)
      for (int a0 = 0; do_far && a0 < qpad; a0 += kChunk) {
        const int na = min(kChunk, qpad - a0);
        __syncthreads();
        for (int e = threadIdx.x; e < kChunk * NC; e += blockDim.x) {
          const int al = e / NC, r = e % NC;
          const int alpha = a0 + al;
          Ms[al * NCP + r] = (alpha < q && rc + r < nrhs)
                                 ? tlhs[((int64_t)c * q + alpha) * nrhs + rc + r]
                                 : T(0);
        }
        __syncthreads();
        for (int ks = 0; ks < na; ks += 4) {
          const int al = ks + lk;
          const int pk = aidx[a0 + al];
          T b[NT];
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) b[nt] = Ms[al * NCP + nt * 16 + lr];
#pragma unroll
          for (int rt = 0; rt < kRT; ++rt) {
            if (!tile_on[rt]) continue;
            const int row = (wave + kWaves * rt) * 16 + lr;
            T v = F[(0 * P1 + unpack(pk, 0)) * kRowsP + row];
#pragma unroll
            for (int d = 1; d < D; ++d) v *= F[(d * P1 + unpack(pk, d)) * kRowsP + row];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
              if (nt < ntl) acc[rt][nt] = M::mma(v, b[nt], acc[rt][nt]);
          }
        }
      }
      // near field: A = K(rows, sources) generated per lane, B = source rhs.
      // The chunks of the near list are staged through registers one chunk
      // ahead: the global loads of chunk c+1 overlap the MFMA work of chunk c.
      constexpr int KX = kChunk * NC / 256;          // staged rhs values per thread
      // staged coordinates per thread (kGram: thread j < kChunk holds the D
      // coordinates of source j)
      constexpr int KY = kGram ? D : (D * kChunk + 255) / 256;
      static_assert(kChunk * NC % 256 == 0, "staging layout");
      T py[KY], px[KX];
      int nb = n0, j0 = 0, pnj = 0;
      // a run of sources with adjacent point ranges is staged as one source
      auto fetch = [&]() {
        const int run = pat.near_run[nb];
        const int s = pat.near_src[nb], sl = pat.near_src[nb + run - 1];
        const int sb = cols.begin[s], ss = cols.begin[sl] + cols.size[sl] - sb;
        pnj = min(kChunk, ss - j0);
        if constexpr (kGram) {
          const int jj = threadIdx.x;
#pragma unroll
          for (int d = 0; d < D; ++d)
            py[d] = jj < pnj ? cols.pts[(int64_t)d * cols.npts + sb + j0 + jj] : gctr[d];
        } else {
#pragma unroll
          for (int k = 0; k < KY; ++k) {
            const int e = threadIdx.x + 256 * k;
            const int d = e / kChunk, jj = e % kChunk;
            py[k] = (e < D * kChunk && jj < pnj)
                        ? cols.pts[(int64_t)d * cols.npts + sb + j0 + jj]
                        : T(0);
          }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
          const int e = threadIdx.x + 256 * k;
          const int jj = e / NC, r = e % NC;
          px[k] = (jj < pnj && rc + r < nrhs) ? xs[(int64_t)(sb + j0 + jj) * nrhs + rc + r]
                                              : T(0);
        }
        j0 += kChunk;
        if (j0 >= ss) {
          nb += run;
          j0 = 0;
        }
      };
      bool more = nb < n1;
      if (more) fetch();
      while (more) {
        const int nj = pnj;
        __syncthreads();
        if constexpr (kGram) {
          // distance row of source j: [-2 y', |y'|^2] in k-slots 0..D-1 and
          // 3 (slots D..2 zero), y' = sigma (y - centre)
          if (threadIdx.x < kChunk) {
            T g[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int d = 0; d < D; ++d) {
              const T yd = gsig * (py[d] - gctr[d]);
              g[d] = T(-2) * yd;
              g[3] = fmat(yd, yd, g[3]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) Ys[threadIdx.x * 4 + k] = g[k];
          }
        } else {
#pragma unroll
          for (int k = 0; k < KY; ++k) {
            const int e = threadIdx.x + 256 * k;
            if (e < D * kChunk) Ys[e] = py[k];
          }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
          const int e = threadIdx.x + 256 * k;
          Xs[(e / NC) * NCP + e % NC] = px[k];
        }
        __syncthreads();
        more = nb < n1;
        if (more) fetch();
        if constexpr (kGram) {
          // 16 sources per group: one distance MFMA per row tile gives lane
          // (lr, lk) the squared distances of target lr to sources
          // lk + 4 v (v = 0..3: the FP64 accumulator layout), i.e. exactly
          // its A value of K-step v.  Rows past nj carry zero rhs and the
          // finite distance row of the leaf's centre.
          auto group = [&](int s0, auto full) {
            const T ga = Ys[(s0 + lr) * 4 + lk];
            T bv[4][NT];
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
              for (int nt = 0; nt < NT; ++nt) bv[v][nt] = Xs[(s0 + 4 * v + lk) * NCP + nt * 16 + lr];
#pragma unroll
            for (int rt = 0; rt < kRT; ++rt) {
              if (!tile_on[rt]) continue;
              const acc_t r2 = M::mma(ga, gx[rt], acc_t{gxx[rt], gxx[rt], gxx[rt], gxx[rt]});
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                if (!decltype(full)::value && s0 + 4 * v >= nj) break;
                const T e = gram_kernel<KT>(__builtin_fabs(r2[v]), kp);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                  if (nt < ntl) acc[rt][nt] = M::mma(e, bv[v][nt], acc[rt][nt]);
              }
            }
          };
          int s0 = 0;
          for (; s0 + 16 <= nj; s0 += 16) group(s0, std::true_type{});
          if (s0 < nj) group(s0, std::false_type{});
        } else {
          const int nsteps = (nj + 3) & ~3;
#pragma unroll 2
          for (int ks = 0; ks < nsteps; ks += 4) {
            const int sj = ks + lk;
            T y[D];
#pragma unroll
            for (int d = 0; d < D; ++d) y[d] = Ys[d * kChunk + sj];
            T b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = Xs[sj * NCP + nt * 16 + lr];
#pragma unroll
            for (int rt = 0; rt < kRT; ++rt) {
              if (!tile_on[rt]) continue;
              T r2 = 0;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const T dd = x[rt][d] - y[d];
                r2 = fmat(dd, dd, r2);
              }
              if constexpr (MULTI) {
                T ev[3];
                kernel_r2_all<KT, T>(r2, kp, ev);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                  if (nt < ntl) acc[rt][nt] = M::mma(pick_variant(ev, tv[nt]), b[nt], acc[rt][nt]);
              } else {
                const T e = kernel_r2<KT, T>(r2, kp);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                  if (nt < ntl) acc[rt][nt] = M::mma(e, b[nt], acc[rt][nt]);
              }
            }
          }
        }
      }
      // store: C row crow(lk, v) of tile (wave + 4 rt), column lr of rhs tile nt
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        if (!tile_on[rt]) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int col = rc + nt * 16 + lr;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = pbase + (wave + kWaves * rt) * 16 + M::crow(lk, v);
            if (i < sz && col < nrhs) {
              const int64_t row = natural ? (int64_t)rows.perm[b0 + i] : (int64_t)(b0 + i);
              double *p = out + row + (int64_t)col * ldo;
              if (split)
                atomic_accumulate(p, (double)acc[rt][nt][v]);
              else
                *p = (double)acc[rt][nt][v];
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Coupling on MFMA: one workgroup per target cluster; rows = the q target
// interpolation nodes (tiles of 16, kRT per wave, so q <= 256), K = the q
// source nodes of each far cluster, N = 16*NT rhs columns.  For q <= 64 (at
// most kRT row tiles) every wave takes all row tiles and a quarter of the
// K-steps, and the four partial sums are added in LDS (in wave order, so the
// result does not depend on scheduling): all waves stay busy.
template <typename T, int D, int KT, int NT, bool MULTI>
__global__ __launch_bounds__(256, (kMinWavesGram<T, NT, KT, MULTI>)) void k_coupling_mfma(DevTreeView<T> rows, DevTreeView<T> cols,
                                                       DevPatternView pat, InterpParams<T> ip,
                                                       KernelParams<T> kp, int nrhs,
                                                       const T *__restrict__ mom,
                                                       T *__restrict__ tlhs) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int NC = 16 * NT;
  constexpr int NCP = NC + 16;  // padded LDS row stride
  constexpr int kChunk = kChunkOf<T, NT>;
  // FP64 entries of the kernels smooth at r = 0: squared distances from the
  // matrix cores, as in the leaf output (16 source nodes per distance MFMA)
  constexpr bool kGram = sizeof(T) == 8 && !MULTI && kGramNear<KT> && D <= 3;
  extern __shared__ unsigned char smem_raw[];
  const int q = ip.q;
  const int qpad = (q + 3) & ~3;
  T *U = reinterpret_cast<T *>(smem_raw);  // [D][qpad] unit coordinates of the nodes
  // [D][kChunk] coordinates of the chunk's rows, or kGram: [kChunk][4] distance rows
  T *Yc = U + D * qpad;
  T *Ms = Yc + kYsRows<T, D> * kChunk;     // [kChunk][NCP] moments of the chunk's rows
  T *Red = Ms + kChunk * NCP;  // [2 kRT * 16][NC] partial sums of the K-split (q <= 128)
  const int lane = threadIdx.x & 63;
  // wave index as a scalar: the row-tile guards below are then uniform branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int task = pat.far_order[blockIdx.x];
  if (task < 0) return;
  const int t = pat.far_task[4 * task], f0 = pat.far_task[4 * task + 1],
            f1 = pat.far_task[4 * task + 2];
  const bool split = (pat.far_task[4 * task + 3] & kTaskSplit) != 0;
  // K-split: with R row tiles, kspl groups of waves (4 for R <= kRT, 2 for
  // R <= 2 kRT, else 1) share the K-steps, so that every wave holds up to kRT
  // row tiles; wave = rg + wrows * sg
  const int R = (q + 15) / 16;
  const int kspl = R <= kRT ? 4 : (R <= 2 * kRT ? 2 : 1);
  const int wrows = kWaves / kspl, rg = wave % wrows, sg = wave / wrows;
  const int ks0 = 4 * sg, kstep = 4 * kspl;
  // The K dimension runs over the far sources' nodes.  A chunk of kChunk
  // rows holds G = kChunk / qpad whole sources when that is 2 or more (small
  // q: fewer stages and barriers per source), else kChunk nodes of one
  // source.  Row bl of a chunk is node bl % qpad of the chunk's source
  // bl / qpad (G >= 2) or node be + bl (G = 1); rows past a source's q nodes
  // or past the list carry zero moments and the coordinates of a real node.
  const int G = max(1, kChunk / qpad);
  const float rqpad = 1.0f / (float)qpad;  // bl / qpad = floor((bl + 0.5) / qpad), bl < 256
  // bl / qpad = (bl * qmagic) >> 16 for bl < 128 and qpad <= 64 (the G >= 2 case)
  const int qmagic = (65536 + qpad - 1) / qpad;
  for (int e = threadIdx.x; e < D * qpad; e += blockDim.x) {
    const int d = e / qpad, b = e - d * qpad, bb = b < q ? b : 0;
    U[e] = ip.td ? ip.unodes[d * q + bb] : ip.xi[unpack(basis_index<T, D>(bb, ip), d)];
  }
  // this lane's target node per row tile
  T x[kRT][D];
  bool tile_on[kRT];
#pragma unroll
  for (int rt = 0; rt < kRT; ++rt) {
    const int tile_row0 = (rg + wrows * rt) * 16;
    tile_on[rt] = tile_row0 < q;
    const int row = tile_row0 + lr < q ? tile_row0 + lr : 0;
    const int pk = basis_index<T, D>(row, ip);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T u = ip.td ? ip.unodes[d * q + row] : ip.xi[unpack(pk, d)];
      x[rt][d] = u * rows.bbsize[t * D + d] + rows.bbmin[t * D + d];
    }
  }
  // kGram: coordinates centred on the target cluster's box and scaled by
  // sigma; the distance MFMA's right operand (k-slot lk of target node lr:
  // [x', 1]) and accumulator input |x'|^2 per row tile
  T gctr[D] = {}, gsig = T(1), gx[kRT], gxx[kRT];
  if constexpr (kGram) {
    gsig = gram_sigma<KT>(kp);
#pragma unroll
    for (int d = 0; d < D; ++d)
      gctr[d] = rows.bbmin[t * D + d] + T(0.5) * rows.bbsize[t * D + d];
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt) {
      T xx = 0, sel = lk == 3 ? T(1) : T(0);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const T xd = gsig * (x[rt][d] - gctr[d]);
        xx = fmat(xd, xd, xx);
        sel = lk == d ? xd : sel;
      }
      gx[rt] = sel;
      gxx[rt] = xx;
    }
  }
  __syncthreads();  // U
  // source slot and node of row bl of a chunk of nsrc whole sources (G >= 2);
  // node -1: a padding row (its coordinates are then those of slot 0, node 0)
  auto chunk_row = [&](int bl, int nsrc, int &slot, int &node) {
    slot = (int)(((float)bl + 0.5f) * rqpad);
    node = bl - slot * qpad;
    if (slot >= nsrc || node >= q) node = -1;
    if (slot >= nsrc) slot = 0;
  };
  for (int rc = 0; rc < nrhs; rc += NC) {
    acc_t acc[kRT][NT];
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = acc_t{0, 0, 0, 0};
    int tv[NT];  // kernel variant of each column tile of this pass (MULTI)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) tv[nt] = kp.tv[min((rc >> 4) + nt, kMaxVariantGroups - 1)];
    // column tiles of this pass that hold rhs columns (the rest are skipped)
    const int ntl = min(NT, (nrhs - rc + 15) >> 4);
    // K-steps of the staged chunk, rows [0, nrows)
    auto compute_chunk = [&](int nrows) {
      if constexpr (kGram) {
        // row groups of 16 (wave group sg takes every kspl-th): lane (lr, lk)
        // gets the squared distances of target lr to rows lk + 4 v of the
        // group, its A values of the group's K-steps v = 0..3.  Rows past
        // nrows (at most kChunk) carry zero moments and a real node.
        for (int g0 = 16 * sg; g0 < nrows; g0 += 16 * kspl) {
          const T ga = Yc[(g0 + lr) * 4 + lk];
          T bv[4][NT];
#pragma unroll
          for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bv[v][nt] = Ms[(g0 + 4 * v + lk) * NCP + nt * 16 + lr];
#pragma unroll
          for (int rt = 0; rt < kRT; ++rt) {
            if (!tile_on[rt]) continue;
            const acc_t r2 = M::mma(ga, gx[rt], acc_t{gxx[rt], gxx[rt], gxx[rt], gxx[rt]});
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const T e = gram_kernel<KT>(__builtin_fabs(r2[v]), kp);
#pragma unroll
              for (int nt = 0; nt < NT; ++nt)
                if (nt < ntl) acc[rt][nt] = M::mma(e, bv[v][nt], acc[rt][nt]);
            }
          }
        }
        return;
      }
#pragma unroll 2
      for (int ks = ks0; ks < nrows; ks += kstep) {
        const int bl = ks + lk;
        T y[D];
#pragma unroll
        for (int d = 0; d < D; ++d) y[d] = Yc[d * kChunk + bl];
        T b[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) b[nt] = Ms[bl * NCP + nt * 16 + lr];
#pragma unroll
        for (int rt = 0; rt < kRT; ++rt) {
          if (!tile_on[rt]) continue;
          T r2 = 0;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const T dd = x[rt][d] - y[d];
            r2 = fmat(dd, dd, r2);
          }
          if constexpr (MULTI) {
            T ev[3];
            kernel_r2_all<KT, T>(r2, kp, ev);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
              if (nt < ntl) acc[rt][nt] = M::mma(pick_variant(ev, tv[nt]), b[nt], acc[rt][nt]);
          } else {
            const T e = kernel_r2<KT, T>(r2, kp);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
              if (nt < ntl) acc[rt][nt] = M::mma(e, b[nt], acc[rt][nt]);
          }
        }
      }
    };
    // Chunks are fetched into registers (coordinates py, moments pm) and then
    // written to LDS between two barriers.  FP32 fetches the next chunk right
    // after the barriers, so its global loads overlap the MFMA work of the
    // current one; FP64 (one wave per SIMD) fetches just before staging, as a
    // lookahead measured slower there.
    constexpr int KM = kChunk * NC / 256;          // staged moments per thread
    // staged coordinates per thread (kGram: thread bl < kChunk holds the D
    // coordinates of row bl)
    constexpr int KY = kGram ? D : (D * kChunk + 255) / 256;
    static_assert(kChunk * NC % 256 == 0, "staging layout");
    T py[KY], pm[KM];
    int f = f0, nbe = 0;  // next chunk: sources from f, nodes from nbe (G = 1)
    int pn = 0;           // rows of the fetched chunk
    auto fetch = [&]() {
      const int nsrc = G > 1 ? min(G, f1 - f) : 1;
      const int be = nbe;
      pn = G > 1 ? nsrc * qpad : min(kChunk, qpad - be);
      if (G > 1) {
        // several sources: per-row source for the coordinates; the rows of
        // one wave-wide moment load (64 / NC consecutive rows, starting at a
        // multiple of 64 / NC) never straddle a source since qpad is a
        // multiple of 4, so that load's source is wave-uniform
        if constexpr (kGram) {
          const int bl = threadIdx.x;
          if (bl < kChunk) {
            int slot, node;
            chunk_row(bl, nsrc, slot, node);
            const int s = pat.far_src[f + slot];
#pragma unroll
            for (int d = 0; d < D; ++d)
              py[d] = U[d * qpad + (node < 0 ? 0 : node)] * cols.bbsize[s * D + d] +
                      cols.bbmin[s * D + d];
          }
        }
#pragma unroll
        for (int k = 0; k < (kGram ? 0 : KY); ++k) {
          const int e = threadIdx.x + 256 * k;
          const int d = e / kChunk, bl = e % kChunk;
          if (e < D * kChunk) {
            int slot, node;
            chunk_row(bl, nsrc, slot, node);
            const int s = pat.far_src[f + slot];
            py[k] = U[d * qpad + (node < 0 ? 0 : node)] * cols.bbsize[s * D + d] +
                    cols.bbmin[s * D + d];
          }
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const int e = threadIdx.x + 256 * k;
          const int bl = e / NC, r = e % NC;
          const int slot = (__builtin_amdgcn_readfirstlane(bl) * qmagic) >> 16;
          const int node = bl - slot * qpad;
          const T *ms = mom + (int64_t)pat.far_src[f + min(slot, nsrc - 1)] * q * nrhs;
          pm[k] = (slot < nsrc && node < q && rc + r < nrhs) ? ms[node * nrhs + rc + r] : T(0);
        }
      } else {
        // one source, nodes [be, be + kChunk)
        const int s = pat.far_src[f];
        const T *ms = mom + (int64_t)s * q * nrhs;
        if constexpr (kGram) {
          const int bl = threadIdx.x;
          if (bl < kChunk) {
            const int node = be + bl < q ? be + bl : 0;
#pragma unroll
            for (int d = 0; d < D; ++d)
              py[d] = U[d * qpad + node] * cols.bbsize[s * D + d] + cols.bbmin[s * D + d];
          }
        }
#pragma unroll
        for (int k = 0; k < (kGram ? 0 : KY); ++k) {
          const int e = threadIdx.x + 256 * k;
          const int d = e / kChunk, bl = e % kChunk;
          if (e < D * kChunk) {
            const int node = be + bl < q ? be + bl : 0;
            py[k] = U[d * qpad + node] * cols.bbsize[s * D + d] + cols.bbmin[s * D + d];
          }
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const int e = threadIdx.x + 256 * k;
          const int bl = e / NC, r = e % NC;
          const int node = be + bl;
          pm[k] = (node < q && rc + r < nrhs) ? ms[node * nrhs + rc + r] : T(0);
        }
      }
      if (G > 1) {
        f += nsrc;
      } else {
        nbe += kChunk;
        if (nbe >= qpad) {
          ++f;
          nbe = 0;
        }
      }
    };
    auto stage = [&]() {
      if constexpr (kGram) {
        // distance row of node row bl: [-2 y', |y'|^2] in k-slots 0..D-1 and
        // 3 (slots D..2 zero), y' = sigma (y - centre)
        if (threadIdx.x < kChunk) {
          T g[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const T yd = gsig * (py[d] - gctr[d]);
            g[d] = T(-2) * yd;
            g[3] = fmat(yd, yd, g[3]);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) Yc[threadIdx.x * 4 + k] = g[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < KY; ++k) {
          const int e = threadIdx.x + 256 * k;
          if (e < D * kChunk) Yc[e] = py[k];
        }
      }
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const int e = threadIdx.x + 256 * k;
        Ms[(e / NC) * NCP + e % NC] = pm[k];
      }
    };
    if constexpr (sizeof(T) == 4) {
      bool more = f < f1;
      if (more) fetch();
      while (more) {
        const int nrows = pn;
        __syncthreads();
        stage();
        __syncthreads();
        more = f < f1;
        if (more) fetch();
        compute_chunk(nrows);
      }
    } else {
      while (f < f1) {
        fetch();
        __syncthreads();
        stage();
        __syncthreads();
        compute_chunk(pn);
      }
    }
    if (kspl > 1) {
      // sum the wave groups' partials in LDS in group order, then store
      // row-contiguous
      for (int g = 0; g < kspl; ++g) {
        if (sg == g) {
#pragma unroll
          for (int rt = 0; rt < kRT; ++rt) {
            if (!tile_on[rt]) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
              if (nt >= ntl) continue;
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int alpha = (rg + wrows * rt) * 16 + M::crow(lk, v);
                T *dst = Red + alpha * NC + nt * 16 + lr;
                *dst = (g == 0 ? T(0) : *dst) + acc[rt][nt][v];
              }
            }
          }
        }
        __syncthreads();
      }
      for (int e = threadIdx.x; e < q * NC; e += blockDim.x) {
        const int alpha = e / NC, c = e - alpha * NC;
        if (c < 16 * ntl && rc + c < nrhs) {
          T *p = tlhs + ((int64_t)t * q + alpha) * nrhs + rc + c;
          if (split)
            atomic_accumulate(p, Red[e]);
          else
            *p = Red[e];
        }
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        if (!tile_on[rt]) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int col = rc + nt * 16 + lr;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int alpha = (rg + wrows * rt) * 16 + M::crow(lk, v);
            if (alpha < q && col < nrhs) {
              T *p = tlhs + ((int64_t)t * q + alpha) * nrhs + col;
              if (split)
                atomic_accumulate(p, acc[rt][nt][v]);
              else
                *p = acc[rt][nt][v];
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Leaf moments on FP32 MFMA, q <= 128: mom[t] (q x NC) = V_t (q x |t|) X_t
// (|t| x NC) per leaf, one workgroup per leaf.  Row tiles are the basis
// functions (16 each), K-steps 4 points; lane (row lr, k-slot lk) forms its
// A value V_alpha(x_p) as the product of the D 1-D factors of point p,
// staged in LDS 64 points at a time with the points' rhs rows.  Wave w owns
// the row tiles w and w + 4 (q <= 128: at most 8 tiles) and all NT column
// tiles.
template <int D, int NT>
__global__ __launch_bounds__(256) void k_upward_leaf_mfma(DevTreeView<float> tr,
                                                          InterpParams<float> ip,
                                                          const float *__restrict__ xs, int nrhs,
                                                          float *__restrict__ mom) {
  using M = Mfma<float>;
  using acc_t = typename M::acc_t;
  constexpr int NC = 16 * NT;
  constexpr int kPts = 64;          // points staged per chunk
  constexpr int kPtsP = kPts + 4;   // padded stride of a factor row
  constexpr int NCP = NC + 4;       // padded stride of a staged rhs row
  extern __shared__ unsigned char smem_raw[];
  float *F = reinterpret_cast<float *>(smem_raw);  // [D][P1][kPtsP] factors
  float *X = F + D * ip.P1 * kPtsP;                 // [kPts][NCP] rhs rows
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int c = tr.leaves[blockIdx.x];
  const int b0 = tr.begin[c], sz = tr.size[c];
  const int P1 = ip.P1, q = ip.q;
  // the two row tiles of this wave and the multi-index of this lane's row
  bool on[2];
  int pk[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int row0 = (wave + 4 * t) * 16;
    on[t] = row0 < q;
    pk[t] = basis_index<float, D>(row0 + lr < q ? row0 + lr : 0, ip);
  }
  for (int rc = 0; rc < nrhs; rc += NC) {
    acc_t acc[2][NT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[t][nt] = acc_t{0, 0, 0, 0};
    for (int p0 = 0; p0 < sz; p0 += kPts) {
      const int np = min(kPts, sz - p0);
      __syncthreads();
      for (int e = threadIdx.x; e < D * kPts; e += 256) {
        const int d = e / kPts, jj = e % kPts;
        if (jj < np) {
          basis_1d<float>(tr.unit[(int64_t)d * tr.npts + b0 + p0 + jj], ip,
                          F + (d * P1) * kPtsP + jj, kPtsP);
        } else {
          for (int a = 0; a < P1; ++a) F[(d * P1 + a) * kPtsP + jj] = 0.f;
        }
      }
      for (int e = threadIdx.x; e < kPts * NC; e += 256) {
        const int jj = e / NC, r = e % NC;
        X[jj * NCP + r] = (jj < np && rc + r < nrhs) ? xs[(int64_t)(b0 + p0 + jj) * nrhs + rc + r]
                                                     : 0.f;
      }
      __syncthreads();
      const int nsteps = (np + 3) & ~3;
      for (int ks = 0; ks < nsteps; ks += 4) {
        const int jj = ks + lk;
        float b[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) b[nt] = X[jj * NCP + nt * 16 + lr];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          if (!on[t]) continue;
          float v = F[(0 * P1 + unpack(pk[t], 0)) * kPtsP + jj];
#pragma unroll
          for (int d = 1; d < D; ++d) v *= F[(d * P1 + unpack(pk[t], d)) * kPtsP + jj];
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) acc[t][nt] = M::mma(v, b[nt], acc[t][nt]);
        }
      }
    }
    // C row crow(lk, v) of tile t is basis function (wave + 4 t) 16 + crow
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if (!on[t]) continue;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int col = rc + nt * 16 + lr;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int alpha = (wave + 4 * t) * 16 + M::crow(lk, v);
          if (alpha < q && col < nrhs) mom[((int64_t)c * q + alpha) * nrhs + col] = acc[t][nt][v];
        }
      }
    }
  }
}

template <typename T, int NT>
size_t leaf_output_mfma_lds(int D, int P1, int q) {
  constexpr int kChunk = kChunkOf<T, NT>;
  const int qpad = (q + 3) & ~3;
  const size_t u1 = (size_t)kChunk * (16 * NT + 16);
  const int ys = (sizeof(T) == 8 && D < 4) ? 4 : D;  // kYsRows
  const size_t u2 = (size_t)ys * kChunk + (size_t)kChunk * (16 * NT + 16);
  return (size_t)D * P1 * kRowsP * sizeof(T) + (size_t)((qpad + 3) & ~3) * sizeof(int) +
         (u1 > u2 ? u1 : u2) * sizeof(T);
}

template <typename T, int NT>
size_t coupling_mfma_lds(int D, int q, bool td) {
  (void)td;  // both schemes stage per-row coordinates
  constexpr int kChunk = kChunkOf<T, NT>;
  const int qpad = (q + 3) & ~3;
  const size_t red = q <= 2 * kRT * 16 ? (size_t)2 * kRT * 16 * 16 * NT : 0;  // K-split partials
  const int ys = (sizeof(T) == 8 && D < 4) ? 4 : D;  // kYsRows
  return ((size_t)D * qpad + (size_t)ys * kChunk + (size_t)kChunk * (16 * NT + 16) + red) *
         sizeof(T);
}

// Specialised entry code for the common kernels, a run-time switch otherwise.
inline bool kt_specialised(int base) {
  return base == KT_EXPONENTIAL || base == KT_MATERN32 || base == KT_MATERN52 ||
         base == KT_GAUSSIAN;
}

// True when the launch needs the MULTI kernels: several variants, or one
// derivative variant of a specialised kernel (evaluated with the shared code
// instead of the generic switch).
inline bool mfma_multi(int kt, int vmask) {
  const int base = kt % KT_COUNT;
  return (vmask & (vmask - 1)) != 0 || (vmask != (1 << KV_VALUE) && kt_specialised(base));
}

template <typename T, int D, int NT, bool MULTI>
hipError_t launch_mfma_nt(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                          const DevPatternView &pat, const InterpParams<T> &ip,
                          const KernelParams<T> &kp, int nrhs, const T *xs, const T *mom,
                          T *tlhs, double *out, int64_t ldo, bool natural, hipStream_t stream) {
  const size_t lds_out = leaf_output_mfma_lds<T, NT>(D, ip.P1, ip.q);
  const size_t lds_cpl = coupling_mfma_lds<T, NT>(D, ip.q, ip.td != 0);
  const int base = kp.kt % KT_COUNT;
  // MULTI kernels take the base kernel; single-variant kernels take kp.kt
  // (a derivative variant there goes to the run-time switch, K = -1)
  const int key = MULTI ? base : kp.kt;
#define FMCA_MFMA_CASE(K)                                                                  \
  case K:                                                                                  \
    if (output)                                                                            \
      hipLaunchKernelGGL((k_leaf_output_mfma<T, D, K, NT, MULTI>), dim3(pat.nnear_order),  \
                         dim3(256), lds_out, stream, rows, cols, pat, ip, kp, nrhs, xs,    \
                         tlhs, out, ldo, natural ? 1 : 0);                                 \
    else                                                                                   \
      hipLaunchKernelGGL((k_coupling_mfma<T, D, K, NT, MULTI>), dim3(pat.nfar_order),      \
                         dim3(256), lds_cpl, stream, rows, cols, pat, ip, kp, nrhs, mom,   \
                         tlhs);                                                            \
    break;
  switch (key) {
    FMCA_MFMA_CASE(KT_EXPONENTIAL)
    FMCA_MFMA_CASE(KT_MATERN32)
    FMCA_MFMA_CASE(KT_MATERN52)
    FMCA_MFMA_CASE(KT_GAUSSIAN)
    default:
      if (output)
        hipLaunchKernelGGL((k_leaf_output_mfma<T, D, -1, NT, MULTI>), dim3(pat.nnear_order),
                           dim3(256), lds_out, stream, rows, cols, pat, ip, kp, nrhs, xs, tlhs,
                           out, ldo, natural ? 1 : 0);
      else
        hipLaunchKernelGGL((k_coupling_mfma<T, D, -1, NT, MULTI>), dim3(pat.nfar_order),
                           dim3(256), lds_cpl, stream, rows, cols, pat, ip, kp, nrhs, mom, tlhs);
  }
#undef FMCA_MFMA_CASE
  return hipGetLastError();
}

// rhs columns per pass: 16 or 32.  Wider passes (64 columns) take fewer
// entry evaluations but drop to one wave per SIMD, and the staging latency per
// source cluster, hidden by occupancy, then costs more (measured 7-8 % slower
// at 64 rhs on uniform and on clustered points).
template <typename T>
int mfma_nt(int nrhs, bool multi) {
  (void)multi;
  return nrhs <= 16 ? 1 : 2;
}

template <typename T, int D, bool MULTI>
hipError_t launch_mfma_dm(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                          const DevPatternView &pat, const InterpParams<T> &ip,
                          const KernelParams<T> &kp, int nrhs, const T *xs, const T *mom,
                          T *tlhs, double *out, int64_t ldo, bool natural, hipStream_t stream) {
  if (mfma_nt<T>(nrhs, MULTI) == 1)
    return launch_mfma_nt<T, D, 1, MULTI>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs,
                                          out, ldo, natural, stream);
  return launch_mfma_nt<T, D, 2, MULTI>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs,
                                        out, ldo, natural, stream);
}

template <typename T, int D>
hipError_t launch_mfma_d(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                         const DevPatternView &pat, const InterpParams<T> &ip,
                         const KernelParams<T> &kp, int nrhs, const T *xs, const T *mom, T *tlhs,
                         double *out, int64_t ldo, bool natural, hipStream_t stream) {
  if (mfma_multi(kp.kt, kp.vmask))
    return launch_mfma_dm<T, D, true>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out,
                                      ldo, natural, stream);
  return launch_mfma_dm<T, D, false>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out,
                                     ldo, natural, stream);
}

}  // namespace

// Leaf moments of the FP32 forms on the matrix cores (q <= 128, >= 8 rhs);
// false: the caller takes the VALU kernel.
bool upward_leaf_mfma_supported(int D, int P1, int q, int nrhs) {
  return q <= 128 && nrhs >= kMfmaMinRhs && D >= 1 && D <= 3 &&
         ((size_t)D * P1 * 68 + 64 * 36) * sizeof(float) <= 64 * 1024;
}

hipError_t launch_upward_leaf_mfma(const DevTreeView<float> &tr, const InterpParams<float> &ip,
                                   const float *xs, int nrhs, float *mom, hipStream_t stream) {
  const int D = tr.dim, P1 = ip.P1;
  const int NT = nrhs > 16 ? 2 : 1;
  const size_t lds = ((size_t)D * P1 * 68 + (size_t)64 * (16 * NT + 4)) * sizeof(float);
#define FMCA_UPL(DV, NTV)                                                                   \
  hipLaunchKernelGGL((k_upward_leaf_mfma<DV, NTV>), dim3(tr.nleaves), dim3(256), lds, stream, \
                     tr, ip, xs, nrhs, mom)
  if (D == 1) {
    if (NT == 2) FMCA_UPL(1, 2); else FMCA_UPL(1, 1);
  } else if (D == 2) {
    if (NT == 2) FMCA_UPL(2, 2); else FMCA_UPL(2, 1);
  } else {
    if (NT == 2) FMCA_UPL(3, 2); else FMCA_UPL(3, 1);
  }
#undef FMCA_UPL
  return hipGetLastError();
}

template <typename T>
bool mfma_supported(bool output, int D, int P1, int q, int nrhs, int kt, int vmask, int td) {
  if (nrhs < kMfmaMinRhs) return false;  // the VALU forms are faster there
  const size_t limit = 64 * 1024;
  const int nt = mfma_nt<T>(nrhs, mfma_multi(kt, vmask));
  if (output)
    return (nt == 1 ? leaf_output_mfma_lds<T, 1>(D, P1, q)
                    : leaf_output_mfma_lds<T, 2>(D, P1, q)) <= limit;
  if (q > kWaves * kRT * 16) return false;
  return (nt == 1 ? coupling_mfma_lds<T, 1>(D, q, td != 0) : coupling_mfma_lds<T, 2>(D, q, td != 0)) <=
         limit;
}

template <typename T>
hipError_t launch_mfma(bool output, const DevTreeView<T> &rows, const DevTreeView<T> &cols,
                       const DevPatternView &pat, const InterpParams<T> &ip,
                       const KernelParams<T> &kp, int nrhs, const T *xs, const T *mom, T *tlhs,
                       double *out, int64_t ldo, bool natural, hipStream_t stream) {
  switch (rows.dim) {
    case 1:
      return launch_mfma_d<T, 1>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                                 natural, stream);
    case 2:
      return launch_mfma_d<T, 2>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                                 natural, stream);
    case 3:
      return launch_mfma_d<T, 3>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                                 natural, stream);
  }
  return hipErrorInvalidValue;
}

template bool mfma_supported<float>(bool, int, int, int, int, int, int, int);
template bool mfma_supported<double>(bool, int, int, int, int, int, int, int);
template hipError_t launch_mfma<float>(bool, const DevTreeView<float> &,
                                       const DevTreeView<float> &, const DevPatternView &,
                                       const InterpParams<float> &, const KernelParams<float> &,
                                       int, const float *, const float *, float *, double *,
                                       int64_t, bool, hipStream_t);
template hipError_t launch_mfma<double>(bool, const DevTreeView<double> &,
                                        const DevTreeView<double> &, const DevPatternView &,
                                        const InterpParams<double> &, const KernelParams<double> &,
                                        int, const double *, const double *, double *, double *,
                                        int64_t, bool, hipStream_t);

}  // namespace fmca_amd
