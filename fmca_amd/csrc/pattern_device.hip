// The block tree of FMCA's H2MatrixBase::computePattern (H2MatrixBase.h:160-191)
// expanded on the GPU, level by level: each pair of clusters of a level is
// tested with CompareCluster::compare (Admissibility.h:95-115) and either
// refines into its four son pairs or becomes a leaf block.  The offsets of
// both come from one exclusive scan over the level, so children and blocks
// land in the order of the host's expansion (pair by pair, sons as FMCA
// enumerates them), which keeps the plan deterministic and equal to the
// host's.  The host then regroups the blocks (finish_pattern).
//
// The distance test runs in FP64 with the host's operation order and no
// contraction; where |max(r_t, r_s) - eta * dist| is within 1e-14 of the
// quantities compared, the device sqrt's rounding could decide the outcome,
// and the expansion is handed back to the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "plan.hpp"

namespace fmca_amd {
namespace {

constexpr int kTB = 256;

struct GeoD {
  double mn[kMaxDim], mx[kMaxDim], rad;
  int32_t son0, son1;
  uint32_t size;
  int32_t pad;
};

// 0 refine, 1 low-rank block, 2 dense block; +4 if the decision is within
// rounding of the admissibility threshold.
__device__ int classify(const GeoD &r, const GeoD &c, int d, double eta, uint32_t qq) {
#pragma clang fp contract(off)
  double s2 = 0;
  for (int i = 0; i < d; ++i) {
    const double v = 0.5 * (r.mn[i] - c.mn[i] + r.mx[i] - c.mx[i]);
    s2 += v * v;
  }
  const double sq = sqrt(s2);
  double dist = sq - r.rad - c.rad;
  dist = dist > 0 ? dist : 0;
  const double rad = r.rad > c.rad ? r.rad : c.rad;
  const double ed = eta * dist;
  const double margin = 1e-14 * (eta * (sq + r.rad + c.rad) + rad);
  const int amb = fabs(rad - ed) <= margin ? 4 : 0;
  if (rad > ed) return ((r.son0 < 0 || c.son0 < 0) ? 2 : 0) | amb;
  // low rank iff q^2 < |t||s| in 32-bit unsigned arithmetic, as FMCA compares
  return (qq < r.size * c.size ? 1 : 2) | amb;
}

// Per pair: its code; per workgroup: the sum of (children << 32 | blocks).
__global__ __launch_bounds__(kTB) void k_classify(const int2 *__restrict__ front, int64_t nf,
                                                  const GeoD *__restrict__ gr,
                                                  const GeoD *__restrict__ gc, int d, double eta,
                                                  uint32_t qq, int8_t *__restrict__ code,
                                                  unsigned long long *__restrict__ wsum,
                                                  int *__restrict__ ambiguous) {
  __shared__ unsigned long long sm[kTB];
  const int64_t i = blockIdx.x * (int64_t)kTB + threadIdx.x;
  unsigned long long v = 0;
  if (i < nf) {
    const int2 p = front[i];
    int c = classify(gr[p.x], gc[p.y], d, eta, qq);
    if (c & 4) *ambiguous = 1;
    c &= 3;
    code[i] = (int8_t)c;
    v = c == 0 ? (4ull << 32) : 1ull;
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) wsum[blockIdx.x] = sm[0];
}

// One workgroup: exclusive scan of the workgroup sums in place; the total
// goes to wsum[nw].
__global__ __launch_bounds__(kTB) void k_scan_sums(unsigned long long *__restrict__ wsum,
                                                   int64_t nw) {
  __shared__ unsigned long long sc[kTB];
  unsigned long long carry = 0;
  for (int64_t b = 0; b < nw; b += kTB) {
    const int64_t k = b + threadIdx.x;
    const unsigned long long v = k < nw ? wsum[k] : 0ull;
    sc[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kTB; o <<= 1) {
      const unsigned long long add = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : 0ull;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    if (k < nw) wsum[k] = carry + sc[threadIdx.x] - v;
    carry += sc[kTB - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) wsum[nw] = carry;
}

// Writes each pair's four son pairs (refine) or its block, at the scanned
// offsets.  Son order as FMCA's: row son i varies fastest (H2MatrixBase.h:181-186).
__global__ __launch_bounds__(kTB) void k_emit(const int2 *__restrict__ front, int64_t nf,
                                              const GeoD *__restrict__ gr,
                                              const GeoD *__restrict__ gc,
                                              const int8_t *__restrict__ code,
                                              const unsigned long long *__restrict__ wsum,
                                              int2 *__restrict__ next, int32_t *__restrict__ brow,
                                              int32_t *__restrict__ bcol,
                                              int8_t *__restrict__ blr) {
  __shared__ unsigned long long sc[kTB];
  const int64_t i = blockIdx.x * (int64_t)kTB + threadIdx.x;
  int c = -1;
  unsigned long long v = 0;
  if (i < nf) {
    c = code[i];
    v = c == 0 ? (4ull << 32) : 1ull;
  }
  sc[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < kTB; o <<= 1) {
    const unsigned long long add = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : 0ull;
    __syncthreads();
    sc[threadIdx.x] += add;
    __syncthreads();
  }
  if (c < 0) return;
  const unsigned long long off = wsum[blockIdx.x] + sc[threadIdx.x] - v;
  const int2 p = front[i];
  if (c == 0) {
    const int64_t o = (int64_t)(off >> 32);
    const GeoD &r = gr[p.x], &s = gc[p.y];
    next[o + 0] = make_int2(r.son0, s.son0);
    next[o + 1] = make_int2(r.son1, s.son0);
    next[o + 2] = make_int2(r.son0, s.son1);
    next[o + 3] = make_int2(r.son1, s.son1);
  } else {
    const int64_t o = (int64_t)(off & 0xffffffffull);
    brow[o] = p.x;
    bcol[o] = p.y;
    blr[o] = c == 1 ? 1 : 0;
  }
}

struct HipFail : std::runtime_error {
  hipError_t code;
  HipFail(const std::string &what, hipError_t c) : std::runtime_error(what), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipFail(std::string(what) + ": " + hipGetErrorString(e), e);
}

template <typename T>
struct DevArray {
  T *p = nullptr;
  size_t n = 0;
  ~DevArray() { device_free(p, n * sizeof(T)); }
  void reserve(size_t k) {
    if (k <= n) return;
    device_free(p, n * sizeof(T));
    p = nullptr;
    n = 0;
    k = std::max<size_t>(k + k / 4, 1024);
    hip_check((hipError_t)device_malloc((void **)&p, k * sizeof(T)), "device allocation");
    n = k;
  }
};

std::vector<GeoD> geometry(const Tree &T) {
  std::vector<GeoD> g(T.ncl);
  for (int c = 0; c < T.ncl; ++c) {
    GeoD &x = g[c];
    for (int i = 0; i < kMaxDim; ++i) x.mn[i] = x.mx[i] = 0;
    for (int i = 0; i < T.dim; ++i) {
      x.mn[i] = T.bbmin[(size_t)c * T.dim + i];
      x.mx[i] = T.bbmax[(size_t)c * T.dim + i];
    }
    x.rad = cluster_radius(T, c);
    x.son0 = T.son0[c];
    x.son1 = T.son1[c];
    x.size = (uint32_t)T.size[c];
    x.pad = 0;
  }
  return g;
}

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};
}  // namespace

int expand_blocks_device(Pattern &pat, const Tree &R, const Tree &C, double eta, int q,
                         std::string *err) {
  try {
    StreamGuard sg;
    hip_check(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking), "hipStreamCreate");
    hipStream_t st = sg.s;
    const std::vector<GeoD> hr = geometry(R);
    const std::vector<GeoD> hc = &C == &R ? std::vector<GeoD>() : geometry(C);
    DevArray<GeoD> dr, dc;
    dr.reserve(hr.size());
    hip_check(hipMemcpyAsync(dr.p, hr.data(), sizeof(GeoD) * hr.size(), hipMemcpyHostToDevice, st),
              "upload rows");
    const GeoD *gc = dr.p;
    if (&C != &R) {
      dc.reserve(hc.size());
      hip_check(hipMemcpyAsync(dc.p, hc.data(), sizeof(GeoD) * hc.size(),
                               hipMemcpyHostToDevice, st), "upload cols");
      gc = dc.p;
    }
    DevArray<int2> front[2];
    DevArray<int8_t> code, blr;
    DevArray<int32_t> brow, bcol;
    DevArray<unsigned long long> wsum;
    DevArray<int> amb;
    amb.reserve(1);
    front[0].reserve(1);
    const int2 root = make_int2(0, 0);
    const int zero = 0;
    hip_check(hipMemcpyAsync(front[0].p, &root, sizeof(int2), hipMemcpyHostToDevice, st), "root");
    hip_check(hipMemcpyAsync(amb.p, &zero, sizeof(int), hipMemcpyHostToDevice, st), "flag");
    const uint32_t qq = uint32_t(q) * uint32_t(q);
    int64_t nf = 1;
    int cur = 0;
    pat.block_row.clear();
    pat.block_col.clear();
    pat.block_lr.clear();
    while (nf > 0) {
      const int64_t nw = (nf + kTB - 1) / kTB;
      code.reserve(nf);
      wsum.reserve(nw + 1);
      k_classify<<<(unsigned)nw, kTB, 0, st>>>(front[cur].p, nf, dr.p, gc, R.dim, eta, qq, code.p,
                                               wsum.p, amb.p);
      k_scan_sums<<<1, kTB, 0, st>>>(wsum.p, nw);
      hip_check(hipGetLastError(), "classify");
      unsigned long long tot = 0;
      int a = 0;
      hip_check(hipMemcpyAsync(&tot, wsum.p + nw, sizeof(tot), hipMemcpyDeviceToHost, st), "total");
      hip_check(hipMemcpyAsync(&a, amb.p, sizeof(int), hipMemcpyDeviceToHost, st), "flag");
      hip_check(hipStreamSynchronize(st), "classify");
      if (a) return 0;
      const int64_t nchild = (int64_t)(tot >> 32), nblk = (int64_t)(tot & 0xffffffffull);
      front[cur ^ 1].reserve(nchild);
      brow.reserve(nblk);
      bcol.reserve(nblk);
      blr.reserve(nblk);
      k_emit<<<(unsigned)nw, kTB, 0, st>>>(front[cur].p, nf, dr.p, gc, code.p, wsum.p,
                                           front[cur ^ 1].p, brow.p, bcol.p, blr.p);
      hip_check(hipGetLastError(), "emit");
      const size_t b0 = pat.block_row.size();
      pat.block_row.resize(b0 + nblk);
      pat.block_col.resize(b0 + nblk);
      pat.block_lr.resize(b0 + nblk);
      if (nblk) {
        hip_check(hipMemcpyAsync(pat.block_row.data() + b0, brow.p, 4 * nblk,
                                 hipMemcpyDeviceToHost, st), "blocks");
        hip_check(hipMemcpyAsync(pat.block_col.data() + b0, bcol.p, 4 * nblk,
                                 hipMemcpyDeviceToHost, st), "blocks");
        hip_check(hipMemcpyAsync(pat.block_lr.data() + b0, blr.p, nblk, hipMemcpyDeviceToHost, st),
                  "blocks");
      }
      hip_check(hipStreamSynchronize(st), "emit");
      cur ^= 1;
      nf = nchild;
    }
    return 1;
  } catch (const HipFail &e) {
    if (err) *err = e.what();
    (void)hipGetLastError();
    return -(int)e.code;
  }
}

}  // namespace fmca_amd
