// Cluster tree construction on the GPU: FMCA's ClusterTree with
// CardinalityBisection (ClusterTreeInitializer_ClusterTree.h:24-132,
// ClusterTreeSplitter.h:60-81), built level by level.
//
// The level loop, the split boxes and the shrink-to-fit boxes stay on the
// host (build_tree_with, O(#clusters)); this file supplies the O(n) parts as
// a TreeSplitter:
//  - split: every node of a level is halved at its median along its axis by a
//    radix select (8 passes of 8 bits over order-preserving 64-bit keys,
//    per-segment histograms) and a stable partition (ballot prefix sums);
//  - finish_leaves: each leaf is sorted by its father's axis in LDS (bitonic)
//    and reduced to its point box.
// With pairwise distinct coordinates at every split point and inside every
// leaf the result is the host planner's (and FMCA's) permutation, bit for
// bit.  Equal coordinates there make the split report a tie and the caller
// builds on the host, as the host planner's own fast build does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "plan.hpp"

namespace fmca_amd {
namespace {

constexpr int kTB = 256;        // threads per workgroup
constexpr int kWaves = kTB / 64;
constexpr int kChunk = 4096;    // positions per workgroup in the split kernels
constexpr int kLeafMax = 2048;  // largest leaf the LDS sort takes

struct DSeg {
  int64_t begin, size;
  int32_t axis, pad;
};

// Order-preserving map of a finite double to uint64 (equal doubles, +0 and
// -0 included, give equal keys).
__device__ __forceinline__ uint64_t order_key(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x + 0.0);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

double key_value(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double x;
  std::memcpy(&x, &u, sizeof(x));
  return x;
}

__global__ void k_iota(uint32_t *perm, int64_t n) {
  for (int64_t j = blockIdx.x * (int64_t)kTB + threadIdx.x; j < n; j += (int64_t)gridDim.x * kTB)
    perm[j] = (uint32_t)j;
}

// Per-axis min / max keys over all points; flag = 1 if a coordinate is -0.0
// (the host's min / max of signed zeros depends on the order of the points).
__global__ __launch_bounds__(kTB) void k_point_box(const double *__restrict__ P, int dim,
                                                   int64_t n, unsigned long long *kmin,
                                                   unsigned long long *kmax, int *flag) {
  __shared__ unsigned long long smn[kTB], smx[kTB];
  for (int i = 0; i < dim; ++i) {
    unsigned long long a = ~0ull, b = 0ull;
    int negzero = 0;
    for (int64_t j = blockIdx.x * (int64_t)kTB + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * kTB) {
      const double x = P[i + (int64_t)dim * j];
      negzero |= (uint64_t)__double_as_longlong(x) == 0x8000000000000000ull;
      const unsigned long long k = order_key(x);
      a = k < a ? k : a;
      b = k > b ? k : b;
    }
    if (negzero) *flag = 1;
    smn[threadIdx.x] = a;
    smx[threadIdx.x] = b;
    __syncthreads();
    for (int s = kTB / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        smn[threadIdx.x] = smn[threadIdx.x + s] < smn[threadIdx.x] ? smn[threadIdx.x + s]
                                                                   : smn[threadIdx.x];
        smx[threadIdx.x] = smx[threadIdx.x + s] > smx[threadIdx.x] ? smx[threadIdx.x + s]
                                                                   : smx[threadIdx.x];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      atomicMin(kmin + i, smn[0]);
      atomicMax(kmax + i, smx[0]);
    }
    __syncthreads();
  }
}

// Workgroup w covers positions [begin + c * kChunk, ...) of segment s,
// w = s * nch + c.
__device__ __forceinline__ void chunk_range(const DSeg &sg, int c, int64_t &j0, int64_t &j1) {
  j0 = sg.begin + (int64_t)c * kChunk;
  j1 = sg.begin + sg.size;
  j1 = j1 < j0 + kChunk ? j1 : j0 + kChunk;
}

__global__ __launch_bounds__(kTB) void k_seg_keys(const double *__restrict__ P, int dim,
                                                  const uint32_t *__restrict__ perm,
                                                  const DSeg *__restrict__ segs, int nch,
                                                  uint64_t *__restrict__ key) {
  const DSeg sg = segs[blockIdx.x / nch];
  int64_t j0, j1;
  chunk_range(sg, blockIdx.x % nch, j0, j1);
  for (int64_t j = j0 + threadIdx.x; j < j1; j += kTB)
    key[j] = order_key(P[sg.axis + (int64_t)dim * perm[j]]);
}

// Histogram of digit `pass` (bits 56 - 8 pass .. 63 - 8 pass) over the keys
// whose higher digits equal the segment's prefix.  Runs of one digit per
// thread are added at once (early passes see few distinct digits).
__global__ __launch_bounds__(kTB) void k_hist(const uint64_t *__restrict__ key,
                                              const DSeg *__restrict__ segs, int nch,
                                              const uint64_t *__restrict__ prefix, int pass,
                                              unsigned *__restrict__ hist) {
  __shared__ unsigned h[256];
  const int s = blockIdx.x / nch;
  const DSeg sg = segs[s];
  int64_t j0, j1;
  chunk_range(sg, blockIdx.x % nch, j0, j1);
  h[threadIdx.x] = 0;
  __syncthreads();
  const int shift = 56 - 8 * pass;
  const uint64_t pre = prefix[s];
  int cur = -1;
  unsigned run = 0;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += kTB) {
    const uint64_t k = key[j];
    if (pass > 0 && ((k ^ pre) >> (shift + 8)) != 0) continue;
    const int d = (int)((k >> shift) & 255);
    if (d == cur) {
      ++run;
    } else {
      if (run) atomicAdd(&h[cur], run);
      cur = d;
      run = 1;
    }
  }
  if (run) atomicAdd(&h[cur], run);
  __syncthreads();
  const unsigned v = h[threadIdx.x];
  if (v) atomicAdd(&hist[(int64_t)s * 256 + threadIdx.x], v);
}

// One workgroup per segment: the digit bucket that holds rank[s]; extends the
// prefix by it, lowers rank[s] by the keys in lower buckets, and clears the
// histogram for the next pass.
__global__ __launch_bounds__(kTB) void k_pick(unsigned *__restrict__ hist, int pass,
                                              uint64_t *__restrict__ prefix,
                                              int64_t *__restrict__ rank) {
  __shared__ int64_t sc[kTB];
  const int s = blockIdx.x;
  const int64_t r = rank[s];
  const int64_t v = hist[(int64_t)s * 256 + threadIdx.x];
  hist[(int64_t)s * 256 + threadIdx.x] = 0;
  sc[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < kTB; o <<= 1) {
    const int64_t add = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
    __syncthreads();
    sc[threadIdx.x] += add;
    __syncthreads();
  }
  const int64_t incl = sc[threadIdx.x], excl = incl - v;
  if (excl <= r && r < incl) {
    prefix[s] |= (uint64_t)threadIdx.x << (56 - 8 * pass);
    rank[s] = r - excl;
  }
}

// Per chunk: the number of keys below the segment's median key, and the
// largest of them (atomicMax into lmax[s]).  Skipped for a segment whose
// median is tied (rank left over), whose build is abandoned.
__global__ __launch_bounds__(kTB) void k_count(const uint64_t *__restrict__ key,
                                               const DSeg *__restrict__ segs, int nch,
                                               const uint64_t *__restrict__ prefix,
                                               const int64_t *__restrict__ rank,
                                               int *__restrict__ cnt,
                                               unsigned long long *__restrict__ lmax) {
  __shared__ int sc[kTB];
  __shared__ unsigned long long sm[kTB];
  const int s = blockIdx.x / nch;
  if (rank[s] != 0) return;
  const DSeg sg = segs[s];
  int64_t j0, j1;
  chunk_range(sg, blockIdx.x % nch, j0, j1);
  const uint64_t med = prefix[s];
  int c = 0;
  unsigned long long m = 0;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += kTB) {
    const uint64_t k = key[j];
    if (k < med) {
      ++c;
      m = k > m ? k : m;
    }
  }
  sc[threadIdx.x] = c;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sc[threadIdx.x] += sc[threadIdx.x + o];
      sm[threadIdx.x] = sm[threadIdx.x + o] > sm[threadIdx.x] ? sm[threadIdx.x + o]
                                                              : sm[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[blockIdx.x] = sc[0];
    if (sm[0]) atomicMax(lmax + s, sm[0]);
  }
}

// One workgroup per segment: exclusive scan of its chunks' counts.
__global__ __launch_bounds__(kTB) void k_scan(const int64_t *__restrict__ rank, int nch,
                                              int *__restrict__ cnt) {
  __shared__ int sc[kTB];
  const int s = blockIdx.x;
  if (rank[s] != 0) return;
  int carry = 0;
  for (int b = 0; b < nch; b += kTB) {
    const int c = b + (int)threadIdx.x;
    const int v = c < nch ? cnt[(int64_t)s * nch + c] : 0;
    sc[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kTB; o <<= 1) {
      const int add = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : 0;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    if (c < nch) cnt[(int64_t)s * nch + c] = carry + sc[threadIdx.x] - v;
    carry += sc[kTB - 1];
    __syncthreads();
  }
}

// Stable partition of each segment: keys below the median to the first half
// (half = size / 2 of them, as no tie was found), the others after it, each
// in their present order.
__global__ __launch_bounds__(kTB) void k_partition(const uint64_t *__restrict__ key,
                                                   const DSeg *__restrict__ segs, int nch,
                                                   const uint64_t *__restrict__ prefix,
                                                   const int64_t *__restrict__ rank,
                                                   const int *__restrict__ cnt,
                                                   const uint32_t *__restrict__ pin,
                                                   uint32_t *__restrict__ pout) {
  __shared__ int wtot[kWaves];
  const int s = blockIdx.x / nch;
  if (rank[s] != 0) return;
  const DSeg sg = segs[s];
  const int c = blockIdx.x % nch;
  int64_t j0, j1;
  chunk_range(sg, c, j0, j1);
  const uint64_t med = prefix[s];
  const int64_t half = sg.size / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t left = cnt[blockIdx.x];  // left keys before this tile in the segment
  for (int64_t t0 = j0; t0 < j1; t0 += kTB) {
    const int64_t j = t0 + threadIdx.x;
    const bool valid = j < j1;
    const bool isl = valid && key[j] < med;
    const uint64_t mask = __ballot(isl);
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
    if (lane == 0) wtot[wave] = __popcll(mask);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < kWaves; ++w) {
      woff += w < wave ? wtot[w] : 0;
      tot += wtot[w];
    }
    if (valid) {
      const int64_t nl = left + woff + below;  // left keys before position j
      const int64_t i = j - sg.begin;
      const int64_t dst = isl ? sg.begin + nl : sg.begin + half + (i - nl);
      pout[dst] = pin[j];
    }
    left += tot;
    __syncthreads();
  }
}

// One workgroup per leaf: sort it by its father's axis (bitonic, in LDS) if it
// has one, flag equal keys, and reduce its point box.
__global__ __launch_bounds__(kTB) void k_leaves(const double *__restrict__ P, int dim,
                                                const DSeg *__restrict__ leaves,
                                                uint32_t *__restrict__ perm,
                                                double *__restrict__ mn, double *__restrict__ mx,
                                                int *__restrict__ tie) {
  __shared__ uint64_t K[kLeafMax];
  __shared__ uint32_t I[kLeafMax];
  __shared__ double rmn[kTB], rmx[kTB];
  const DSeg lf = leaves[blockIdx.x];
  uint32_t *id = perm + lf.begin;
  const int n = (int)lf.size;
  const bool sorted = lf.axis >= 0 && n > 1;  // the box then reads the sorted ids from LDS
  if (sorted) {
    int np = 1;
    while (np < n) np <<= 1;
    for (int j = threadIdx.x; j < np; j += kTB) {
      if (j < n) {
        const uint32_t p = id[j];
        I[j] = p;
        K[j] = order_key(P[lf.axis + (int64_t)dim * p]);
      } else {
        I[j] = 0;
        K[j] = ~0ull;
      }
    }
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1)
      for (int h = k >> 1; h > 0; h >>= 1) {
        for (int i = threadIdx.x; i < np; i += kTB) {
          const int o = i ^ h;
          if (o > i) {
            const bool up = (i & k) == 0;
            const uint64_t a = K[i], b = K[o];
            if ((a > b) == up) {
              K[i] = b;
              K[o] = a;
              const uint32_t t = I[i];
              I[i] = I[o];
              I[o] = t;
            }
          }
        }
        __syncthreads();
      }
    for (int j = threadIdx.x; j < n; j += kTB) {
      id[j] = I[j];
      if (j && K[j - 1] == K[j]) *tie = 1;
    }
  }
  for (int i = 0; i < dim; ++i) {
    double a = P[i + (int64_t)dim * (sorted ? I[0] : id[0])], b = a;
    for (int j = threadIdx.x; j < n; j += kTB) {
      const double x = P[i + (int64_t)dim * (sorted ? I[j] : id[j])];
      a = x < a ? x : a;
      b = x > b ? x : b;
    }
    rmn[threadIdx.x] = a;
    rmx[threadIdx.x] = b;
    __syncthreads();
    for (int o = kTB / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        rmn[threadIdx.x] = rmn[threadIdx.x + o] < rmn[threadIdx.x] ? rmn[threadIdx.x + o]
                                                                   : rmn[threadIdx.x];
        rmx[threadIdx.x] = rmx[threadIdx.x + o] > rmx[threadIdx.x] ? rmx[threadIdx.x + o]
                                                                   : rmx[threadIdx.x];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      mn[(int64_t)blockIdx.x * dim + i] = rmn[0];
      mx[(int64_t)blockIdx.x * dim + i] = rmx[0];
    }
    __syncthreads();
  }
}

__global__ void k_fill_u32(unsigned *p, int64_t n, unsigned v) {
  for (int64_t j = blockIdx.x * (int64_t)kTB + threadIdx.x; j < n; j += (int64_t)gridDim.x * kTB)
    p[j] = v;
}

struct HipFail : std::runtime_error {
  hipError_t code;
  HipFail(const std::string &what, hipError_t c) : std::runtime_error(what), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipFail(std::string(what) + ": " + hipGetErrorString(e), e);
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kTB - 1) / kTB, 4096)); }

template <typename T>
struct DevArray {
  T *p = nullptr;
  size_t n = 0;
  ~DevArray() { device_free(p, n * sizeof(T)); }
  // capacity for k elements; returns true if (re)allocated
  bool reserve(size_t k) {
    if (k <= n) return false;
    device_free(p, n * sizeof(T));
    p = nullptr;
    n = 0;
    hip_check((hipError_t)device_malloc((void **)&p, std::max<size_t>(k, 1) * sizeof(T)),
              "device allocation");
    n = k;
    return true;
  }
};

class DeviceSplitter final : public TreeSplitter {
 public:
  DeviceSplitter(int dim, int64_t n, const double *P, hipStream_t st) : dim_(dim), n_(n), st_(st) {
    P_.reserve((size_t)dim * n);
    perm_[0].reserve(n);
    perm_[1].reserve(n);
    key_.reserve(n);
    hip_check(hipMemcpyAsync(P_.p, P, sizeof(double) * dim * n, hipMemcpyHostToDevice, st_),
              "upload points");
    k_iota<<<grid_for(n), kTB, 0, st_>>>(perm_[0].p, n);
    hip_check(hipGetLastError(), "k_iota");
  }

  // Raw coordinate box of the points; false if a coordinate is -0.0.
  bool box(double *mn, double *mx) {
    std::vector<unsigned long long> init(2 * dim_);
    for (int i = 0; i < dim_; ++i) {
      init[i] = ~0ull;
      init[dim_ + i] = 0ull;
    }
    DevArray<unsigned long long> kb;
    DevArray<int> flag;
    kb.reserve(2 * dim_);
    flag.reserve(1);
    const int zero = 0;
    hip_check(hipMemcpyAsync(kb.p, init.data(), sizeof(unsigned long long) * 2 * dim_,
                             hipMemcpyHostToDevice, st_), "upload box init");
    hip_check(hipMemcpyAsync(flag.p, &zero, sizeof(int), hipMemcpyHostToDevice, st_), "flag");
    k_point_box<<<std::min(grid_for(n_), 1024), kTB, 0, st_>>>(P_.p, dim_, n_, kb.p, kb.p + dim_,
                                                               flag.p);
    hip_check(hipGetLastError(), "k_point_box");
    int f = 0;
    hip_check(hipMemcpyAsync(init.data(), kb.p, sizeof(unsigned long long) * 2 * dim_,
                             hipMemcpyDeviceToHost, st_), "download box");
    hip_check(hipMemcpyAsync(&f, flag.p, sizeof(int), hipMemcpyDeviceToHost, st_), "flag");
    hip_check(hipStreamSynchronize(st_), "point box");
    for (int i = 0; i < dim_; ++i) {
      mn[i] = key_value(init[i]);
      mx[i] = key_value(init[dim_ + i]);
    }
    return f == 0;
  }

  bool split(const std::vector<SplitSeg> &segs, double *lmax, double *rmin) override {
    const int ns = (int)segs.size();
    int64_t maxsz = 0;
    std::vector<DSeg> hs(ns);
    std::vector<uint64_t> pre(ns, 0);
    std::vector<int64_t> rank(ns);
    std::vector<unsigned long long> lm(ns, 0);
    for (int k = 0; k < ns; ++k) {
      hs[k] = {segs[k].begin, segs[k].size, segs[k].axis, 0};
      rank[k] = segs[k].size / 2;
      maxsz = std::max(maxsz, segs[k].size);
    }
    const int nch = (int)((maxsz + kChunk - 1) / kChunk);
    const int64_t nwg = (int64_t)ns * nch;
    if (nwg > INT32_MAX) throw HipFail("too many workgroups", hipErrorInvalidValue);
    seg_.reserve(ns);
    prefix_.reserve(ns);
    rank_.reserve(ns);
    lmax_.reserve(ns);
    cnt_.reserve(nwg);
    if (hist_.reserve((size_t)ns * 256)) {
      k_fill_u32<<<grid_for(hist_.n), kTB, 0, st_>>>(hist_.p, hist_.n, 0u);
      hip_check(hipGetLastError(), "k_fill_u32");
    }
    hip_check(hipMemcpyAsync(seg_.p, hs.data(), sizeof(DSeg) * ns, hipMemcpyHostToDevice, st_), "segs");
    hip_check(hipMemcpyAsync(prefix_.p, pre.data(), 8 * ns, hipMemcpyHostToDevice, st_), "prefix");
    hip_check(hipMemcpyAsync(rank_.p, rank.data(), 8 * ns, hipMemcpyHostToDevice, st_), "rank");
    hip_check(hipMemcpyAsync(lmax_.p, lm.data(), 8 * ns, hipMemcpyHostToDevice, st_), "lmax");
    uint32_t *pin = perm_[cur_].p, *pout = perm_[cur_ ^ 1].p;
    k_seg_keys<<<(int)nwg, kTB, 0, st_>>>(P_.p, dim_, pin, seg_.p, nch, key_.p);
    for (int pass = 0; pass < 8; ++pass) {
      k_hist<<<(int)nwg, kTB, 0, st_>>>(key_.p, seg_.p, nch, prefix_.p, pass, hist_.p);
      k_pick<<<ns, kTB, 0, st_>>>(hist_.p, pass, prefix_.p, rank_.p);
    }
    k_count<<<(int)nwg, kTB, 0, st_>>>(key_.p, seg_.p, nch, prefix_.p, rank_.p, cnt_.p, lmax_.p);
    k_scan<<<ns, kTB, 0, st_>>>(rank_.p, nch, cnt_.p);
    // positions outside the segments (finished leaves) keep their entries
    hip_check(hipMemcpyAsync(pout, pin, sizeof(uint32_t) * n_, hipMemcpyDeviceToDevice, st_),
              "copy permutation");
    k_partition<<<(int)nwg, kTB, 0, st_>>>(key_.p, seg_.p, nch, prefix_.p, rank_.p, cnt_.p, pin,
                                           pout);
    hip_check(hipGetLastError(), "split kernels");
    hip_check(hipMemcpyAsync(pre.data(), prefix_.p, 8 * ns, hipMemcpyDeviceToHost, st_), "prefix");
    hip_check(hipMemcpyAsync(rank.data(), rank_.p, 8 * ns, hipMemcpyDeviceToHost, st_), "rank");
    hip_check(hipMemcpyAsync(lm.data(), lmax_.p, 8 * ns, hipMemcpyDeviceToHost, st_), "lmax");
    hip_check(hipStreamSynchronize(st_), "split");
    cur_ ^= 1;
    for (int k = 0; k < ns; ++k) {
      if (rank[k] != 0) return false;  // the median's key also lies left of it
      lmax[k] = key_value(lm[k]);
      rmin[k] = key_value(pre[k]);
    }
    return true;
  }

  bool finish_leaves(const std::vector<SplitSeg> &leaves, double *mn, double *mx) override {
    const int nl = (int)leaves.size();
    std::vector<DSeg> hl(nl);
    for (int k = 0; k < nl; ++k) {
      hl[k] = {leaves[k].begin, leaves[k].size, leaves[k].axis, 0};
      if (leaves[k].axis >= 0 && leaves[k].size > kLeafMax) throw HipFail("leaf too large", hipErrorInvalidValue);
    }
    DevArray<DSeg> dl;
    DevArray<double> dmn, dmx;
    DevArray<int> tie;
    dl.reserve(nl);
    dmn.reserve((size_t)nl * dim_);
    dmx.reserve((size_t)nl * dim_);
    tie.reserve(1);
    int t = 0;
    hip_check(hipMemcpyAsync(dl.p, hl.data(), sizeof(DSeg) * nl, hipMemcpyHostToDevice, st_), "leaves");
    hip_check(hipMemcpyAsync(tie.p, &t, sizeof(int), hipMemcpyHostToDevice, st_), "tie");
    k_leaves<<<nl, kTB, 0, st_>>>(P_.p, dim_, dl.p, perm_[cur_].p, dmn.p, dmx.p, tie.p);
    hip_check(hipGetLastError(), "k_leaves");
    hip_check(hipMemcpyAsync(mn, dmn.p, sizeof(double) * nl * dim_, hipMemcpyDeviceToHost, st_), "mn");
    hip_check(hipMemcpyAsync(mx, dmx.p, sizeof(double) * nl * dim_, hipMemcpyDeviceToHost, st_), "mx");
    hip_check(hipMemcpyAsync(&t, tie.p, sizeof(int), hipMemcpyDeviceToHost, st_), "tie");
    hip_check(hipStreamSynchronize(st_), "leaves");
    return t == 0;
  }

  void permutation(std::vector<uint32_t> &perm) override {
    perm.resize(n_);
    hip_check(hipMemcpyAsync(perm.data(), perm_[cur_].p, sizeof(uint32_t) * n_,
                             hipMemcpyDeviceToHost, st_), "download permutation");
    hip_check(hipStreamSynchronize(st_), "permutation");
  }

 private:
  int dim_;
  int64_t n_;
  hipStream_t st_;
  int cur_ = 0;
  DevArray<double> P_;
  DevArray<uint32_t> perm_[2];
  DevArray<uint64_t> key_, prefix_;
  DevArray<int64_t> rank_;
  DevArray<unsigned long long> lmax_;
  DevArray<DSeg> seg_;
  DevArray<int> cnt_;
  DevArray<unsigned> hist_;
};

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};
}  // namespace

int build_tree_device(Tree &T, int dim, int64_t n, const double *P, int64_t min_csize,
                      std::string *err) {
  const int64_t thr = min_csize >= 1 ? (2 * min_csize - 1) : 1;
  if (thr > kLeafMax) {  // leaves beyond the LDS sort
    build_tree(T, dim, n, P, min_csize);
    return 0;
  }
  try {
    StreamGuard sg;
    hip_check(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking), "hipStreamCreate");
    bool built = false;
    {
      DeviceSplitter sp(dim, n, P, sg.s);
      double mn[kMaxDim], mx[kMaxDim];
      if (sp.box(mn, mx)) built = build_tree_with(T, dim, n, P, mn, mx, min_csize, sp);
    }
    if (built) return 1;
  } catch (const HipFail &e) {
    if (err) *err = e.what();
    (void)hipGetLastError();
    return -(int)e.code;
  }
  // equal coordinates at a split point or inside a leaf, or a -0.0
  build_tree(T, dim, n, P, min_csize);
  return 0;
}

}  // namespace fmca_amd
