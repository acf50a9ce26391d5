// Matrix-core contractions of the FP32 precision form on BF16 inputs.
//
// The kernel entries are evaluated in FP32 exactly as in the FP32 MFMA forms
// (kernels_mfma.hip), and so are the right operands (rhs rows, moments).  Each
// FP32 operand v is split into two BF16 terms, hi = bf16(v) (round to nearest
// even) and lo = bf16(v - hi), and a product a.b is formed as
//     a_hi.b_hi + a_lo.b_hi + a_hi.b_lo
// with v_mfma_f32_16x16x32_bf16 and FP32 accumulation.  The dropped terms are
// of relative size 2^-16 (a_lo.b_lo and the remainders of the splits), so a
// product carries about the error of an FP32 evaluation of the entry; the sum
// accumulates in FP32 as in the FP32 forms.  One 16x16x32 BF16 instruction
// does 8 times the multiply-adds of a 16x16x4 FP32 one in half the cycles, so
// the three of them cost 3/16 of the FP32 matrix-core time per entry and the
// kernels become bound by the VALU work of generating the entries.
//
// Operand layout (gfx950): lane l supplies A[l % 16][k] and B[k][l % 16] for
// the eight k of K-slot group l / 16; the two operands use the same k <-> (lane
// group, element) map, so the contraction pairs them whatever that map is.
// Lane l's eight A values are this lane's row and eight consecutive sources
// (or source nodes) of the K-step; B is read as 16 bytes from LDS, where the
// staged right operand is held transposed (column-major over sources).  The
// accumulator layout is that of the 16x16 FP32 forms: item v of lane l is
// C[4 (l / 16) + v][l % 16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "device.hpp"
#include "device_math.hpp"

namespace fmca_amd {

namespace {

using namespace dev;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int kWaves = 4;                // waves per workgroup
constexpr int kRT = 4;                   // 16-row tiles per wave -> 256 rows per pass
constexpr int kRows = kWaves * kRT * 16;
constexpr int kRowsP = kRows + 16;       // padded LDS row stride of the basis factors
constexpr int kKB = 32;                  // sources per BF16 K-step
constexpr int kChunk = 128;              // sources / nodes staged per chunk
constexpr int kChunkP = kChunk + 8;      // LDS stride of a transposed column (bf16)
constexpr int kFarChunk = 64;            // moments staged per chunk of the far output

__device__ __forceinline__ f32x4 mma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma_f32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 splat4(float v) { return f32x4{v, v, v, v}; }

__device__ __forceinline__ int unpack(int p, int d) { return (p >> (8 * d)) & 0xff; }

// hi = bf16(v) rounded to nearest even, lo = bf16(v - hi)
__device__ __forceinline__ void split_bf16(float v, u16 &hi, u16 &lo) {
  const __bf16 h = (__bf16)v;
  const float hf = (float)h;
  const __bf16 l = (__bf16)(v - hf);
  hi = __builtin_bit_cast(u16, h);
  lo = __builtin_bit_cast(u16, l);
}

// Pairs: one packed conversion gives both high parts, whose FP32 values are
// the two 16-bit halves shifted into place (no second conversion).
__device__ __forceinline__ void split8(const float (&v)[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
  for (int i = 0; i < 8; i += 2) {
    const __bf16 h0 = (__bf16)v[i], h1 = (__bf16)v[i + 1];
    hi[i] = h0;
    hi[i + 1] = h1;
    const uint32_t u = (uint32_t)__builtin_bit_cast(u16, h0) |
                       ((uint32_t)__builtin_bit_cast(u16, h1) << 16);
    const float f0 = __builtin_bit_cast(float, u << 16);
    const float f1 = __builtin_bit_cast(float, u & 0xffff0000u);
    lo[i] = (__bf16)(v[i] - f0);
    lo[i + 1] = (__bf16)(v[i + 1] - f1);
  }
}

// Lane l of the distance MFMA's output holds sources 4 (l / 16) + v of each
// 16-source half of a K-step; the BF16 contraction reads k-slots
// 8 (l / 16) .. + 7, so source s of a 32-block is staged at this position.
__device__ __forceinline__ int gram_pos(int jj) {
  return (jj & ~31) | (((jj & 15) >> 2) << 3) | ((jj & 16) >> 2) | (jj & 3);
}

// the 8 bf16 values at p (16-byte aligned) as one operand
__device__ __forceinline__ bf16x8 load8(const u16 *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// In-place BF16 split of the right operands of one action (the permuted rhs
// before the leaf output, the source moments before the coupling): x[i]
// becomes hi | lo << 16, so the staging loops copy the two halves instead
// of splitting every value again for every target that reads it.
__global__ __launch_bounds__(256) void k_split_inplace(float *__restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    u16 h, l;
    split_bf16(x[i], h, l);
    x[i] = __builtin_bit_cast(float, (uint32_t)h | ((uint32_t)l << 16));
  }
}

// ---------------------------------------------------------------------------
// Leaf output: one workgroup per target-leaf task, 4 waves, each wave owns row
// tiles {wave, wave + 4, wave + 8, wave + 12} of the current 256-row pass and
// 16 NT rhs columns.  Far field: A = V_t^T (rows x q) on FP32 MFMA as in the
// FP32 form (a small share of the work); near field on BF16 MFMA.
template <int D, int KT, int NT>
__global__ __launch_bounds__(256, 2) void k_leaf_output_bf16(
    DevTreeView<float> rows, DevTreeView<float> cols, DevPatternView pat,
    InterpParams<float> ip, KernelParams<float> kp, int nrhs, const float *__restrict__ xs,
    const float *__restrict__ tlhs, double *__restrict__ out, int64_t ldo, int natural) {
  constexpr int NC = 16 * NT;
  constexpr int NCP = NC + 16;  // padded stride of the staged far moments
  constexpr bool kGram = kGramNear<KT>;
  extern __shared__ unsigned char smem_raw[];
  const int P1 = ip.P1, q = ip.q;
  const int qpad = (q + 3) & ~3;
  float *F = reinterpret_cast<float *>(smem_raw);               // [D][P1][kRowsP]
  int *aidx = reinterpret_cast<int *>(F + D * P1 * kRowsP);     // [qpad] packed indices
  float *Un = reinterpret_cast<float *>(aidx + ((qpad + 3) & ~3));  // union region
  float *Ms = Un;                                  // far: [kFarChunk][NCP] moments
  // near: [D][kChunk] source points, or with kGram [4][kChunk] rows of the
  // distance MFMA's source operand
  float *Ys = Un;
  u16 *Xh = reinterpret_cast<u16 *>(Ys + 4 * kChunk);  // near: [NC][kChunkP] rhs, hi
  u16 *Xl = Xh + NC * kChunkP;                         // near: [NC][kChunkP] rhs, lo
  const int lane = threadIdx.x & 63;
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
    aidx[e] = basis_index<float, D>(e < q ? e : 0, ip);
  for (int pbase = 0; pbase < sz; pbase += kRows) {
    __syncthreads();
    {
      const int i = pbase + threadIdx.x;
      if (i < sz) {
#pragma unroll
        for (int d = 0; d < D; ++d)
          basis_1d<float>(rows.unit[(int64_t)d * rows.npts + b0 + i], ip,
                          F + (d * P1) * kRowsP + threadIdx.x, kRowsP);
      } else {
        for (int d = 0; d < D; ++d)
          for (int a = 0; a < P1; ++a) F[(d * P1 + a) * kRowsP + threadIdx.x] = 0.f;
      }
    }
    float x[kRT][D];
    bool tile_on[kRT];
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt) {
      const int tile_row0 = pbase + (wave + kWaves * rt) * 16;
      tile_on[rt] = tile_row0 < sz;
      const int i = tile_row0 + lr;
#pragma unroll
      for (int d = 0; d < D; ++d)
        x[rt][d] = i < sz ? rows.pts[(int64_t)d * rows.npts + b0 + i] : 0.f;
    }
    // distance MFMA operands of the targets: xb = component lk of [x', 1]
    // (x' = x - centre of the leaf), xx = |x'|^2
    float ctr[D], xb[kRT], xx[kRT];
    const float sig = gram_sigma<KT>(kp);
#pragma unroll
    for (int d = 0; d < D; ++d)
      ctr[d] = rows.bbmin[(int64_t)c * D + d] + 0.5f * rows.bbsize[(int64_t)c * D + d];
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt) {
      float s2 = 0.f, b = lk == 3 ? 1.f : 0.f;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float v = (x[rt][d] - ctr[d]) * sig;
        s2 = fmat(v, v, s2);
        if (d == lk) b = v;
      }
      xx[rt] = s2;
      xb[rt] = b;
    }
    for (int rc = 0; rc < nrhs; rc += NC) {
      f32x4 acc[kRT][NT];
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = f32x4{0, 0, 0, 0};
      const int ntl = min(NT, (nrhs - rc + 15) >> 4);
      // far field (first task of the leaf): FP32 MFMA, K-steps of 4 moments
      for (int a0 = 0; do_far && a0 < qpad; a0 += kFarChunk) {
        const int na = min(kFarChunk, qpad - a0);
        __syncthreads();
        for (int e = threadIdx.x; e < kFarChunk * NC; e += blockDim.x) {
          const int al = e / NC, r = e % NC;
          const int alpha = a0 + al;
          Ms[al * NCP + r] = (alpha < q && rc + r < nrhs)
                                 ? tlhs[((int64_t)c * q + alpha) * nrhs + rc + r]
                                 : 0.f;
        }
        __syncthreads();
        for (int ks = 0; ks < na; ks += 4) {
          const int al = ks + lk;
          const int pk = aidx[a0 + al];
          float b[NT];
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) b[nt] = Ms[al * NCP + nt * 16 + lr];
#pragma unroll
          for (int rt = 0; rt < kRT; ++rt) {
            if (!tile_on[rt]) continue;
            const int row = (wave + kWaves * rt) * 16 + lr;
            float v = F[(0 * P1 + unpack(pk, 0)) * kRowsP + row];
#pragma unroll
            for (int d = 1; d < D; ++d) v *= F[(d * P1 + unpack(pk, d)) * kRowsP + row];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
              if (nt < ntl) acc[rt][nt] = mma_f32(v, b[nt], acc[rt][nt]);
          }
        }
      }
      // near field: chunks of kChunk sources fetched one chunk ahead into
      // registers, staged as coordinates Ys and the BF16 split of the rhs
      constexpr int KX = kChunk * NC / 256;          // rhs values per thread
      // coordinates per thread: with kGram thread jj < kChunk loads all D of
      // source jj (its row of the distance operand needs |y'|^2)
      constexpr int KY = kGram ? D : (D * kChunk + 255) / 256;
      static_assert(kChunk * NC % 256 == 0, "staging layout");
      static_assert(kChunk <= 256, "one source per thread in the distance staging");
      float py[KY], px[KX];
      int nb = n0, j0 = 0, pnj = 0;
      auto fetch = [&]() {
        const int run = pat.near_run[nb];
        const int s = pat.near_src[nb], sl = pat.near_src[nb + run - 1];
        const int sb = cols.begin[s], ss = cols.begin[sl] + cols.size[sl] - sb;
        pnj = min(kChunk, ss - j0);
        if constexpr (kGram) {
          const int jj = threadIdx.x;
#pragma unroll
          for (int d = 0; d < D; ++d)  // padding sources sit at the centre
            py[d] = (jj < kChunk && jj < pnj) ? cols.pts[(int64_t)d * cols.npts + sb + j0 + jj]
                                              : ctr[d];
        } else {
#pragma unroll
          for (int k = 0; k < KY; ++k) {
            const int e = threadIdx.x + 256 * k;
            const int d = e / kChunk, jj = e % kChunk;
            py[k] = (e < D * kChunk && jj < pnj)
                        ? cols.pts[(int64_t)d * cols.npts + sb + j0 + jj]
                        : 0.f;
          }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
          const int e = threadIdx.x + 256 * k;
          const int jj = e / NC, r = e % NC;
          px[k] = (jj < pnj && rc + r < nrhs) ? xs[(int64_t)(sb + j0 + jj) * nrhs + rc + r] : 0.f;
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
          const int jj = threadIdx.x;
          if (jj < kChunk) {
            float s2 = 0.f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
              const float v = (py[d] - ctr[d]) * sig;
              Ys[d * kChunk + jj] = -2.f * v;
              s2 = fmat(v, v, s2);
            }
#pragma unroll
            for (int d = D; d < 3; ++d) Ys[d * kChunk + jj] = 0.f;
            Ys[3 * kChunk + jj] = s2;
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
          const int jj = e / NC, r = e % NC;
          const int pos = kGram ? gram_pos(jj) : jj;
          const uint32_t w = __builtin_bit_cast(uint32_t, px[k]);  // hi | lo << 16
          Xh[r * kChunkP + pos] = (u16)(w & 0xffffu);
          Xl[r * kChunkP + pos] = (u16)(w >> 16);
        }
        __syncthreads();
        more = nb < n1;
        if (more) fetch();
        const int nsteps = (nj + kKB - 1) & ~(kKB - 1);
        // all four row tiles valid and both column tiles used (the common
        // case): no branches in a K-step, one basic block for the scheduler
        bool all_on = ntl == NT;
#pragma unroll
        for (int rt = 0; rt < kRT; ++rt) all_on = all_on && tile_on[rt];
        auto gram_steps = [&](auto all_t) {
          constexpr bool ALL = decltype(all_t)::value;
          for (int ks = 0; ks < nsteps; ks += kKB) {
            const int j8 = ks + 8 * lk;  // this lane group's eight k-slots
            // the distances of row tile rt + 1 are issued before the entries
            // of row tile rt are formed, so the matrix core works under the VALU
            const float ya = Ys[lk * kChunk + ks + lr], yb = Ys[lk * kChunk + ks + 16 + lr];
            f32x4 g0 = f32x4{0, 0, 0, 0}, g1 = f32x4{0, 0, 0, 0};
            if (ALL || tile_on[0]) {
              g0 = mma_f32(ya, xb[0], splat4(xx[0]));
              g1 = mma_f32(yb, xb[0], splat4(xx[0]));
            }
            // right operands of the K-step, shared by the row tiles
            bf16x8 bh[NT], bl[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
              bh[nt] = load8(Xh + (nt * 16 + lr) * kChunkP + j8);
              bl[nt] = load8(Xl + (nt * 16 + lr) * kChunkP + j8);
            }
#pragma unroll
            for (int rt = 0; rt < kRT; ++rt) {
              f32x4 n0 = f32x4{0, 0, 0, 0}, n1 = f32x4{0, 0, 0, 0};
              if (rt + 1 < kRT && (ALL || tile_on[rt + 1])) {
                n0 = mma_f32(ya, xb[rt + 1], splat4(xx[rt + 1]));
                n1 = mma_f32(yb, xb[rt + 1], splat4(xx[rt + 1]));
              }
              const f32x4 c0 = g0, c1 = g1;
              g0 = n0;
              g1 = n1;
              if (!ALL && !tile_on[rt]) continue;
              float ev[8];
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                ev[v] = gram_kernel<KT>(fabsf(c0[v]), kp);
                ev[4 + v] = gram_kernel<KT>(fabsf(c1[v]), kp);
              }
              bf16x8 ah, al;
              split8(ev, ah, al);
#pragma unroll
              for (int nt = 0; nt < NT; ++nt) {
                if (!ALL && nt >= ntl) continue;
                acc[rt][nt] = mma_bf16(ah, bh[nt], acc[rt][nt]);
                acc[rt][nt] = mma_bf16(al, bh[nt], acc[rt][nt]);
                acc[rt][nt] = mma_bf16(ah, bl[nt], acc[rt][nt]);
              }
            }
          }
        };
        if constexpr (kGram) {
          if (all_on)
            gram_steps(std::true_type{});
          else
            gram_steps(std::false_type{});
        }
        // kernels with a cusp at r = 0: differences on the VALU
        auto direct_steps = [&](auto all_t) {
          constexpr bool ALL = decltype(all_t)::value;
          for (int ks = 0; ks < nsteps; ks += kKB) {
            const int j8 = ks + 8 * lk;  // this lane group's eight sources
            float y[D][8];
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
              for (int i = 0; i < 8; ++i) y[d][i] = Ys[d * kChunk + j8 + i];
            bf16x8 bh[NT], bl[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
              bh[nt] = load8(Xh + (nt * 16 + lr) * kChunkP + j8);
              bl[nt] = load8(Xl + (nt * 16 + lr) * kChunkP + j8);
            }
#pragma unroll
            for (int rt = 0; rt < kRT; ++rt) {
              if (!ALL && !tile_on[rt]) continue;
              float ev[8];
#pragma unroll
              for (int i = 0; i < 8; ++i) {
                float r2 = 0.f;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                  const float dd = x[rt][d] - y[d][i];
                  r2 = fmat(dd, dd, r2);
                }
                ev[i] = kernel_r2<KT, float>(r2, kp);
              }
              bf16x8 ah, al;
              split8(ev, ah, al);
#pragma unroll
              for (int nt = 0; nt < NT; ++nt) {
                if (!ALL && nt >= ntl) continue;
                acc[rt][nt] = mma_bf16(ah, bh[nt], acc[rt][nt]);
                acc[rt][nt] = mma_bf16(al, bh[nt], acc[rt][nt]);
                acc[rt][nt] = mma_bf16(ah, bl[nt], acc[rt][nt]);
              }
            }
          }
        };
        // 32 columns: right operands loaded once per K-step for the four row
        // tiles, and the branch-free K-step when all tiles are valid (13 %
        // faster for EXPONENTIAL at 32 rhs); 16 columns and the run-time
        // kernel switch (which would spill) keep the per-tile loads (5 %
        // faster at 16 columns)
        constexpr bool kHoist = !kGram && NT == 2 && KT >= 0;
        if constexpr (kHoist) {
          if (all_on)
            direct_steps(std::true_type{});
          else
            direct_steps(std::false_type{});
        }
        for (int ks = 0; !kGram && !kHoist && ks < nsteps; ks += kKB) {
          const int j8 = ks + 8 * lk;  // this lane group's eight sources
          float y[D][8];
#pragma unroll
          for (int d = 0; d < D; ++d)
#pragma unroll
            for (int i = 0; i < 8; ++i) y[d][i] = Ys[d * kChunk + j8 + i];
#pragma unroll
          for (int rt = 0; rt < kRT; ++rt) {
            if (!tile_on[rt]) continue;
            float ev[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
              float r2 = 0.f;
#pragma unroll
              for (int d = 0; d < D; ++d) {
                const float dd = x[rt][d] - y[d][i];
                r2 = fmat(dd, dd, r2);
              }
              ev[i] = kernel_r2<KT, float>(r2, kp);
            }
            bf16x8 ah, al;
            split8(ev, ah, al);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
              if (nt >= ntl) continue;
              const bf16x8 bh = load8(Xh + (nt * 16 + lr) * kChunkP + j8);
              const bf16x8 bl = load8(Xl + (nt * 16 + lr) * kChunkP + j8);
              acc[rt][nt] = mma_bf16(ah, bh, acc[rt][nt]);
              acc[rt][nt] = mma_bf16(al, bh, acc[rt][nt]);
              acc[rt][nt] = mma_bf16(ah, bl, acc[rt][nt]);
            }
          }
        }
      }
      // store: C row 4 lk + v of tile (wave + 4 rt), column lr of rhs tile nt
#pragma unroll
      for (int rt = 0; rt < kRT; ++rt) {
        if (!tile_on[rt]) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int col = rc + nt * 16 + lr;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = pbase + (wave + kWaves * rt) * 16 + 4 * lk + v;
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
// Coupling: one workgroup per target-cluster task; rows = the q target nodes
// (tiles of 16, kRT per wave, q <= 256), K = the nodes of the far sources in
// K-steps of 32, N = 16 NT rhs columns.  A chunk of kChunk rows holds
// G = kChunk / qpad whole sources when that is 2 or more, else kChunk nodes of
// one source; rows past a source's q nodes or past the list carry zero
// moments.  For q <= 64 every wave takes all row tiles and every fourth
// K-step, and the partial sums are added in LDS in wave order.
template <int D, int KT, int NT>
__global__ __launch_bounds__(256, 2) void k_coupling_bf16(
    DevTreeView<float> rows, DevTreeView<float> cols, DevPatternView pat,
    InterpParams<float> ip, KernelParams<float> kp, int nrhs, const float *__restrict__ mom,
    float *__restrict__ tlhs) {
  constexpr int NC = 16 * NT;
  extern __shared__ unsigned char smem_raw[];
  const int q = ip.q;
  const int qpad = (q + 3) & ~3;
  float *U = reinterpret_cast<float *>(smem_raw);  // [D][qpad] unit node coordinates
  // [4][kChunk] rows of the distance MFMA's source operand: -2 y' (k-slots
  // 0..D-1), |y'|^2 (k-slot 3), y' = node - centre of the target cluster.
  // Far blocks are separated (dist >= radius / eta), so the cancellation in
  // |y'|^2 - 2 x'.y' + |x'|^2 stays at the rounding of the direct form.
  float *Yc = U + D * qpad;
  u16 *Mh = reinterpret_cast<u16 *>(Yc + 4 * kChunk);  // [NC][kChunkP] moments, hi
  u16 *Ml = Mh + NC * kChunkP;                         // [NC][kChunkP] moments, lo
  float *Red = reinterpret_cast<float *>(Ml + NC * kChunkP);  // [2 kRT 16][NC] K-split sums
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int task = pat.far_order[blockIdx.x];
  if (task < 0) return;
  const int t = pat.far_task[4 * task], f0 = pat.far_task[4 * task + 1],
            f1 = pat.far_task[4 * task + 2];
  const bool split = (pat.far_task[4 * task + 3] & kTaskSplit) != 0;
  const int R = (q + 15) / 16;
  const int kspl = R <= kRT ? 4 : (R <= 2 * kRT ? 2 : 1);
  const int wrows = kWaves / kspl, rg = wave % wrows, sg = wave / wrows;
  const int ks0 = kKB * sg, kstep = kKB * kspl;
  const int G = max(1, kChunk / qpad);
  const float rqpad = 1.0f / (float)qpad;  // bl / qpad = floor((bl + 0.5) / qpad), bl < 256
  for (int e = threadIdx.x; e < D * qpad; e += blockDim.x) {
    const int d = e / qpad, b = e - d * qpad, bb = b < q ? b : 0;
    U[e] = ip.td ? ip.unodes[d * q + bb] : ip.xi[unpack(basis_index<float, D>(bb, ip), d)];
  }
  float ctr[D], xb[kRT], xx[kRT];
  const float sig = gram_sigma<KT>(kp);
  bool tile_on[kRT];
#pragma unroll
  for (int d = 0; d < D; ++d) ctr[d] = rows.bbmin[t * D + d] + 0.5f * rows.bbsize[t * D + d];
#pragma unroll
  for (int rt = 0; rt < kRT; ++rt) {
    const int tile_row0 = (rg + wrows * rt) * 16;
    tile_on[rt] = tile_row0 < q;
    const int row = tile_row0 + lr < q ? tile_row0 + lr : 0;
    const int pk = basis_index<float, D>(row, ip);
    float s2 = 0.f, b = lk == 3 ? 1.f : 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float u = ip.td ? ip.unodes[d * q + row] : ip.xi[unpack(pk, d)];
      const float v = (u * rows.bbsize[t * D + d] + rows.bbmin[t * D + d] - ctr[d]) * sig;
      s2 = fmat(v, v, s2);
      if (d == lk) b = v;
    }
    xx[rt] = s2;  // distance MFMA operands of the target nodes
    xb[rt] = b;
  }
  __syncthreads();  // U
  // source slot and node of row bl of a chunk of nsrc whole sources (G >= 2);
  // node -1: a padding row
  auto chunk_row = [&](int bl, int nsrc, int &slot, int &node) {
    slot = (int)(((float)bl + 0.5f) * rqpad);
    node = bl - slot * qpad;
    if (slot >= nsrc || node >= q) node = -1;
    if (slot >= nsrc) slot = 0;
  };
  for (int rc = 0; rc < nrhs; rc += NC) {
    f32x4 acc[kRT][NT];
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = f32x4{0, 0, 0, 0};
    const int ntl = min(NT, (nrhs - rc + 15) >> 4);
    bool all_on = ntl == NT;
#pragma unroll
    for (int rt = 0; rt < kRT; ++rt) all_on = all_on && tile_on[rt];
    // all four row tiles valid and both column tiles used (the common case):
    // no branches in a K-step, one basic block for the scheduler
    auto chunk_steps = [&](int nrows, auto all_t) {
      constexpr bool ALL = decltype(all_t)::value;
      for (int ks = ks0; ks < nrows; ks += kstep) {
        const int j8 = ks + 8 * lk;  // this lane group's eight k-slots (gram_pos order)
        const float ya = Yc[lk * kChunk + ks + lr], yb = Yc[lk * kChunk + ks + 16 + lr];
        // distances of row tile rt + 1 issued before the entries of rt
        f32x4 g0 = f32x4{0, 0, 0, 0}, g1 = f32x4{0, 0, 0, 0};
        if (ALL || tile_on[0]) {
          g0 = mma_f32(ya, xb[0], splat4(xx[0]));
          g1 = mma_f32(yb, xb[0], splat4(xx[0]));
        }
        bf16x8 bh[NT], bl[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          bh[nt] = load8(Mh + (nt * 16 + lr) * kChunkP + j8);
          bl[nt] = load8(Ml + (nt * 16 + lr) * kChunkP + j8);
        }
#pragma unroll
        for (int rt = 0; rt < kRT; ++rt) {
          f32x4 n0 = f32x4{0, 0, 0, 0}, n1 = f32x4{0, 0, 0, 0};
          if (rt + 1 < kRT && (ALL || tile_on[rt + 1])) {
            n0 = mma_f32(ya, xb[rt + 1], splat4(xx[rt + 1]));
            n1 = mma_f32(yb, xb[rt + 1], splat4(xx[rt + 1]));
          }
          const f32x4 c0 = g0, c1 = g1;
          g0 = n0;
          g1 = n1;
          if (!ALL && !tile_on[rt]) continue;
          float ev[8];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            ev[v] = gram_kernel<KT>(fabsf(c0[v]), kp);
            ev[4 + v] = gram_kernel<KT>(fabsf(c1[v]), kp);
          }
          bf16x8 ah, al;
          split8(ev, ah, al);
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) {
            if (!ALL && nt >= ntl) continue;
            acc[rt][nt] = mma_bf16(ah, bh[nt], acc[rt][nt]);
            acc[rt][nt] = mma_bf16(al, bh[nt], acc[rt][nt]);
            acc[rt][nt] = mma_bf16(ah, bl[nt], acc[rt][nt]);
          }
        }
      }
    };
    auto compute_chunk = [&](int nrows) {
      if (all_on)
        chunk_steps(nrows, std::true_type{});
      else
        chunk_steps(nrows, std::false_type{});
    };
    constexpr int KM = kChunk * NC / 256;  // staged moments per thread
    static_assert(kChunk * NC % 256 == 0, "staging layout");
    static_assert(kChunk <= 256, "one node per thread in the distance staging");
    float py[D], pm[KM];  // py: the D coordinates of node threadIdx.x of the chunk
    int f = f0, nbe = 0;
    int pn = 0;
    auto fetch = [&]() {
      const int nsrc = G > 1 ? min(G, f1 - f) : 1;
      const int be = nbe;
      pn = G > 1 ? nsrc * qpad : min(kChunk, qpad - be);
      if (G > 1) {
        if (threadIdx.x < kChunk) {
          int slot, node;
          chunk_row(threadIdx.x, nsrc, slot, node);
          const int s = pat.far_src[f + slot];
#pragma unroll
          for (int d = 0; d < D; ++d)
            py[d] = U[d * qpad + (node < 0 ? 0 : node)] * cols.bbsize[s * D + d] +
                    cols.bbmin[s * D + d];
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const int e = threadIdx.x + 256 * k;
          const int bl = e / NC, r = e % NC;
          int slot, node;
          chunk_row(bl, nsrc, slot, node);
          const float *ms = mom + (int64_t)pat.far_src[f + slot] * q * nrhs;
          pm[k] = (node >= 0 && rc + r < nrhs) ? ms[node * nrhs + rc + r] : 0.f;
        }
      } else {
        const int s = pat.far_src[f];
        const float *ms = mom + (int64_t)s * q * nrhs;
        if (threadIdx.x < kChunk) {
          const int node = be + (int)threadIdx.x < q ? be + (int)threadIdx.x : 0;
#pragma unroll
          for (int d = 0; d < D; ++d)
            py[d] = U[d * qpad + node] * cols.bbsize[s * D + d] + cols.bbmin[s * D + d];
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const int e = threadIdx.x + 256 * k;
          const int bl = e / NC, r = e % NC;
          const int node = be + bl;
          pm[k] = (bl < pn && node < q && rc + r < nrhs) ? ms[node * nrhs + rc + r] : 0.f;
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
      if (threadIdx.x < kChunk) {
        const int bl = threadIdx.x;
        float s2 = 0.f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float v = (py[d] - ctr[d]) * sig;
          Yc[d * kChunk + bl] = -2.f * v;
          s2 = fmat(v, v, s2);
        }
#pragma unroll
        for (int d = D; d < 3; ++d) Yc[d * kChunk + bl] = 0.f;
        Yc[3 * kChunk + bl] = s2;
      }
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        const int e = threadIdx.x + 256 * k;
        const int bl = e / NC, r = e % NC;
        const uint32_t w = __builtin_bit_cast(uint32_t, pm[k]);  // hi | lo << 16
        Mh[r * kChunkP + gram_pos(bl)] = (u16)(w & 0xffffu);
        Ml[r * kChunkP + gram_pos(bl)] = (u16)(w >> 16);
      }
    };
    // no register lookahead of the next chunk here: its 18 values per thread
    // push the 32-column form past 256 VGPRs (spills)
    while (f < f1) {
      fetch();
      const int nrows = (pn + kKB - 1) & ~(kKB - 1);
      __syncthreads();
      stage();
      __syncthreads();
      compute_chunk(nrows);
    }
    if (kspl > 1) {
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
                const int alpha = (rg + wrows * rt) * 16 + 4 * lk + v;
                float *dst = Red + alpha * NC + nt * 16 + lr;
                *dst = (g == 0 ? 0.f : *dst) + acc[rt][nt][v];
              }
            }
          }
        }
        __syncthreads();
      }
      for (int e = threadIdx.x; e < q * NC; e += blockDim.x) {
        const int alpha = e / NC, cc = e - alpha * NC;
        if (cc < 16 * ntl && rc + cc < nrhs) {
          float *p = tlhs + ((int64_t)t * q + alpha) * nrhs + rc + cc;
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
            const int alpha = (rg + wrows * rt) * 16 + 4 * lk + v;
            if (alpha < q && col < nrhs) {
              float *p = tlhs + ((int64_t)t * q + alpha) * nrhs + col;
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

template <int NT>
size_t leaf_output_bf16_lds(int D, int P1, int q) {
  constexpr int NC = 16 * NT;
  const int qpad = (q + 3) & ~3;
  const size_t far = (size_t)kFarChunk * (NC + 16) * sizeof(float);
  const size_t near = (size_t)4 * kChunk * sizeof(float) + (size_t)2 * NC * kChunkP * sizeof(u16);
  return (size_t)D * P1 * kRowsP * sizeof(float) + (size_t)((qpad + 3) & ~3) * sizeof(int) +
         (far > near ? far : near);
}

template <int NT>
size_t coupling_bf16_lds(int D, int q) {
  constexpr int NC = 16 * NT;
  const int qpad = (q + 3) & ~3;
  const size_t red = q <= 2 * kRT * 16 ? (size_t)2 * kRT * 16 * NC * sizeof(float) : 0;
  return (size_t)D * qpad * sizeof(float) + (size_t)4 * kChunk * sizeof(float) +
         (size_t)2 * NC * kChunkP * sizeof(u16) + red;
}

template <int D, int NT>
hipError_t launch_bf16_nt(bool output, const DevTreeView<float> &rows,
                          const DevTreeView<float> &cols, const DevPatternView &pat,
                          const InterpParams<float> &ip, const KernelParams<float> &kp, int nrhs,
                          const float *xs, const float *mom, float *tlhs, double *out,
                          int64_t ldo, bool natural, hipStream_t stream) {
  const size_t lds_out = leaf_output_bf16_lds<NT>(D, ip.P1, ip.q);
  const size_t lds_cpl = coupling_bf16_lds<NT>(D, ip.q);
#define FMCA_BF16_CASE(K)                                                                   \
  case K:                                                                                   \
    if (output)                                                                             \
      hipLaunchKernelGGL((k_leaf_output_bf16<D, K, NT>), dim3(pat.nnear_order), dim3(256),  \
                         lds_out, stream, rows, cols, pat, ip, kp, nrhs, xs, tlhs, out, ldo, \
                         natural ? 1 : 0);                                                  \
    else                                                                                    \
      hipLaunchKernelGGL((k_coupling_bf16<D, K, NT>), dim3(pat.nfar_order), dim3(256),      \
                         lds_cpl, stream, rows, cols, pat, ip, kp, nrhs, mom, tlhs);        \
    break;
  switch (kp.kt) {
    FMCA_BF16_CASE(KT_EXPONENTIAL)
    FMCA_BF16_CASE(KT_MATERN32)
    FMCA_BF16_CASE(KT_MATERN52)
    FMCA_BF16_CASE(KT_GAUSSIAN)
    default:
      if (output)
        hipLaunchKernelGGL((k_leaf_output_bf16<D, -1, NT>), dim3(pat.nnear_order), dim3(256),
                           lds_out, stream, rows, cols, pat, ip, kp, nrhs, xs, tlhs, out, ldo,
                           natural ? 1 : 0);
      else
        hipLaunchKernelGGL((k_coupling_bf16<D, -1, NT>), dim3(pat.nfar_order), dim3(256),
                           lds_cpl, stream, rows, cols, pat, ip, kp, nrhs, mom, tlhs);
  }
#undef FMCA_BF16_CASE
  return hipGetLastError();
}

template <int D>
hipError_t launch_bf16_d(bool output, const DevTreeView<float> &rows,
                         const DevTreeView<float> &cols, const DevPatternView &pat,
                         const InterpParams<float> &ip, const KernelParams<float> &kp, int nrhs,
                         const float *xs, const float *mom, float *tlhs, double *out, int64_t ldo,
                         bool natural, hipStream_t stream) {
  if (nrhs <= 16)
    return launch_bf16_nt<D, 1>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                                natural, stream);
  return launch_bf16_nt<D, 2>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                              natural, stream);
}

}  // namespace

bool bf16_supported(bool output, int D, int P1, int q, int nrhs, int kt, int vmask) {
  if (nrhs < kMfmaMinRhs) return false;
  if (vmask != (1 << KV_VALUE) || kt >= KT_COUNT) return false;  // kernel values only
  const size_t limit = 64 * 1024;
  const int nt = nrhs <= 16 ? 1 : 2;
  if (output)
    return (nt == 1 ? leaf_output_bf16_lds<1>(D, P1, q) : leaf_output_bf16_lds<2>(D, P1, q)) <=
           limit;
  if (q > kWaves * kRT * 16) return false;
  return (nt == 1 ? coupling_bf16_lds<1>(D, q) : coupling_bf16_lds<2>(D, q)) <= limit;
}

hipError_t launch_bf16(bool output, const DevTreeView<float> &rows,
                       const DevTreeView<float> &cols, const DevPatternView &pat,
                       const InterpParams<float> &ip, const KernelParams<float> &kp, int nrhs,
                       float *xs, float *mom, float *tlhs, double *out, int64_t ldo,
                       bool natural, hipStream_t stream) {
  {
    // the right operand this launch stages, split once here (rewritten by
    // the next action's permutation / upward pass)
    float *x = output ? xs : mom;
    const int64_t n = (output ? cols.npts : (int64_t)cols.ncl * ip.q) * nrhs;
    if (n > 0) {
      const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
      hipLaunchKernelGGL(k_split_inplace, dim3(grid), dim3(256), 0, stream, x, n);
    }
  }
  switch (rows.dim) {
    case 1:
      return launch_bf16_d<1>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                              natural, stream);
    case 2:
      return launch_bf16_d<2>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                              natural, stream);
    case 3:
      return launch_bf16_d<3>(output, rows, cols, pat, ip, kp, nrhs, xs, mom, tlhs, out, ldo,
                              natural, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace fmca_amd
