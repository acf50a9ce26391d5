// C-ABI implementation: the actor owns the host plan and its HBM copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fmca_amd.h"
#include "device.hpp"
#include "plan.hpp"
#include "projector.hpp"
#include "samplets.hpp"

using namespace fmca_amd;

namespace {

thread_local std::string g_last_error;

fmca_status fail(fmca_status st, const std::string &msg) {
  g_last_error = msg;
  return st;
}

fmca_status from_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return FMCA_OK;
  const fmca_status st = e == hipErrorOutOfMemory ? FMCA_ERR_OUT_OF_MEMORY
                         : (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
                             ? FMCA_ERR_NO_DEVICE
                             : FMCA_ERR_DEVICE;
  return fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

void fmca_amd::set_last_error(const std::string &msg) { g_last_error = msg; }

namespace {

#define HIP_TRY(expr)                                            \
  do {                                                           \
    hipError_t _e = (expr);                                      \
    if (_e != hipSuccess) return from_hip(_e, #expr);            \
  } while (0)

// Device bytes held by all DevBufs of the process, and their high-water mark.
std::atomic<int64_t> g_dev_live{0}, g_dev_peak{0};
// Cap on g_dev_live set through fmca_set_device_memory_limit (0: none).  An
// allocation that would exceed it fails as hipMalloc does when HBM is full.
std::atomic<int64_t> g_dev_limit{0};

void note_device_bytes(int64_t delta) {
  const int64_t now = g_dev_live.fetch_add(delta) + delta;
  int64_t pk = g_dev_peak.load();
  while (now > pk && !g_dev_peak.compare_exchange_weak(pk, now)) {
  }
}

// A device allocation owned by the actor.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;  // requested size
  size_t held = 0;   // allocated size
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      note_device_bytes(-(int64_t)held);
    }
    p = nullptr;
    bytes = held = 0;
  }
  hipError_t alloc(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    const size_t h = std::max<size_t>(b, 256);
    const int64_t cap = g_dev_limit.load();
    if (cap > 0 && g_dev_live.load() + (int64_t)h > cap) {
      p = nullptr;
      return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc(&p, h);
    if (e == hipSuccess) {
      bytes = b;
      held = h;
      note_device_bytes((int64_t)h);
    } else {
      p = nullptr;
    }
    return e;
  }
  template <typename T>
  hipError_t upload(const std::vector<T> &v) {
    hipError_t e = alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// Integer arrays of one tree on the device (shared by both precisions).
struct DevTreeInts {
  DevBuf perm, pinv, begin, size, son0, son1, parent, leaves;
};

template <typename T>
struct DevTreeReals {
  DevBuf pts, unit, bbmin, bbsize, transfer;
};

struct DevPattern {
  DevBuf far_ptr, far_src, far_order, far_task, far_split;
  DevBuf near_ptr, near_src, near_run, near_order, near_task, near_split;
  int far_top_target = 0, far_top_source = 0;
  DevPatternView view(const Pattern &p) const {
    DevPatternView v;
    v.far_top_target = far_top_target;
    v.far_top_source = far_top_source;
    v.far_ptr = far_ptr.as<int32_t>();
    v.far_src = far_src.as<int32_t>();
    v.nfar_order = (int)p.far_order.size();
    v.far_order = far_order.as<int32_t>();
    v.far_task = far_task.as<int32_t>();
    v.nfar_split = (int)p.far_split.size();
    v.far_split = far_split.as<int32_t>();
    v.nnear_leaves = (int)p.near_ptr.size() - 1;
    v.near_ptr = near_ptr.as<int32_t>();
    v.near_src = near_src.as<int32_t>();
    v.near_run = near_run.as<int32_t>();
    v.nnear_order = (int)p.near_order.size();
    v.near_order = near_order.as<int32_t>();
    v.near_task = near_task.as<int32_t>();
    v.nnear_split = (int)p.near_split.size();
    v.near_split = near_split.as<int32_t>();
    return v;
  }
  hipError_t upload(const Pattern &p) {
    hipError_t e;
    // first far target and source (INT_MAX: no far blocks at all)
    far_top_target = far_top_source = std::numeric_limits<int>::max();
    for (size_t t = 0; t + 1 < p.far_ptr.size(); ++t)
      if (p.far_ptr[t + 1] > p.far_ptr[t]) {
        far_top_target = (int)t;
        break;
      }
    for (int32_t src : p.far_src) far_top_source = std::min(far_top_source, (int)src);
    if ((e = far_ptr.upload(p.far_ptr)) != hipSuccess) return e;
    if ((e = far_src.upload(p.far_src)) != hipSuccess) return e;
    if ((e = far_order.upload(p.far_order)) != hipSuccess) return e;
    if ((e = far_task.upload(p.far_task)) != hipSuccess) return e;
    if ((e = far_split.upload(p.far_split)) != hipSuccess) return e;
    if ((e = near_ptr.upload(p.near_ptr)) != hipSuccess) return e;
    if ((e = near_src.upload(p.near_src)) != hipSuccess) return e;
    if ((e = near_run.upload(p.near_run)) != hipSuccess) return e;
    if ((e = near_order.upload(p.near_order)) != hipSuccess) return e;
    if ((e = near_task.upload(p.near_task)) != hipSuccess) return e;
    return near_split.upload(p.near_split);
  }
};

template <typename T>
struct Workspace {
  DevBuf xs, mom, tlhs;
};

}  // namespace

struct fmca_h2_actor {
  int dim = 0, deg = 0;
  int64_t m = 0, n = 0;
  double eta = 0;
  int kt = 0;
  double l = 1, c = 1, nu = 1;
  int geodesic = 0;  // distance type: 0 Euclidean, 1 geodesic (fmca_h2_set_distance_type)
  Interpolation ip;
  Tree rows_tree, cols_tree;
  bool shared = false;  // Peval and P are the same point set: one tree
  // planning report: seconds in the tree builds and the pattern; where the
  // trees and the block expansion ran (0 host, 1 device, 2 device asked for,
  // host used)
  double tree_s = 0, pattern_s = 0;
  int rows_builder = 0, cols_builder = 0, pattern_builder = 0;
  Pattern pat, pat_t;
  bool have_pat_t = false;
  std::vector<double> trans_rows, trans_cols;
  std::vector<double> pts_rows, pts_cols;  // dim-major, cluster order
  // device state
  int device = -1;
  bool uploaded = false;
  DevTreeInts d_rows_i, d_cols_i;
  DevTreeReals<double> d_rows_64, d_cols_64;
  DevTreeReals<float> d_rows_32, d_cols_32;
  DevPattern d_pat, d_pat_t;
  // total-degree interpolation data: multi-indices, unit nodes (dim-major), invV
  DevBuf d_midx, d_unodes64, d_unodes32, d_invV64, d_invV32;
  Workspace<double> ws64;
  Workspace<float> ws32;
  DevBuf stage_in, stage_out;
  DevBuf cg_ws, cg_scal;  // CG vectors and per-column scalars
  // GCR basis: chunks of kBasisChunk slot pairs, slot pointer tables (held
  // during one solve)
  std::vector<std::unique_ptr<DevBuf>> basis;
  DevBuf basis_tab;
  DevBuf grad_x, grad_part;  // gradient: [alpha | G | alpha] rhs, partial sums of dL/dl
  // assembled blocks (fmca_h2_assemble) of precision st_prec (-1: none)
  DevBuf st_far, st_near, st_near_off, st_v_off, st_far_tgt, st_near_leaf, st_far_order,
      st_near_order;
  int st_prec = -1, st_Rq = 0, st_Kq = 0;
  int64_t st_bytes = 0;
  std::vector<int64_t> st_near_off_host;  // element offsets of the stored near blocks
  void release_stored() {
    for (DevBuf *b : {&st_far, &st_near, &st_near_off, &st_v_off, &st_far_tgt, &st_near_leaf,
                      &st_far_order, &st_near_order})
      b->release();
    st_prec = -1;
    st_bytes = 0;
  }
  hipStream_t stream = nullptr;
  std::mutex mtx;
  // Cross-stream order of the actor's workspaces: every action records
  // order_ev on its stream, and an action on another stream first waits for
  // it.  Work captured into a graph is left out (its order is the graph's).
  hipEvent_t order_ev = nullptr;
  hipStream_t order_stream = nullptr;
  bool order_valid = false;
  // per-phase profiling (events recorded on the launch stream)
  bool profiling = false;
  // events of the profiled actions not yet read, with each action's
  // precision; read events go back to ev_pool, so that a profiled action
  // records events but creates none once the pool is warm
  std::vector<std::pair<int, PhaseEvents>> pending;
  std::vector<hipEvent_t> ev_pool;
  double phase_ms[3][PH_COUNT] = {{0}};
  int64_t phase_count[3][PH_COUNT] = {{0}};

  const Tree &rows() const { return rows_tree; }
  const Tree &cols() const { return shared ? rows_tree : cols_tree; }

  ~fmca_h2_actor() {
    for (auto &pe : pending)
      for (auto &e : pe.second.ev) (void)hipEventDestroy(e);
    for (auto &e : ev_pool) (void)hipEventDestroy(e);
    if (order_ev) (void)hipEventDestroy(order_ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

std::vector<double> cluster_order_points(const Tree &T, const double *P) {
  const int d = T.dim;
  std::vector<double> out((size_t)d * T.npts);
  parallel_ranges(T.npts, [&](int64_t lo, int64_t hi) {
    for (int i = 0; i < d; ++i)
      for (int64_t j = lo; j < hi; ++j) out[(size_t)i * T.npts + j] = P[i + (size_t)d * T.perm[j]];
  });
  return out;
}

template <typename T>
std::vector<T> convert(const std::vector<double> &v) {
  return std::vector<T>(v.begin(), v.end());
}

hipError_t upload_ints(DevTreeInts &d, const Tree &T) {
  hipError_t e;
  std::vector<int32_t> b(T.begin.begin(), T.begin.end()), s(T.size.begin(), T.size.end());
  std::vector<uint32_t> pinv(T.perm.size());
  parallel_ranges((int64_t)T.perm.size(), [&](int64_t lo, int64_t hi) {
    for (int64_t j = lo; j < hi; ++j) pinv[T.perm[j]] = (uint32_t)j;
  });
  if ((e = d.perm.upload(T.perm)) != hipSuccess) return e;
  if ((e = d.pinv.upload(pinv)) != hipSuccess) return e;
  if ((e = d.begin.upload(b)) != hipSuccess) return e;
  if ((e = d.size.upload(s)) != hipSuccess) return e;
  if ((e = d.son0.upload(T.son0)) != hipSuccess) return e;
  if ((e = d.son1.upload(T.son1)) != hipSuccess) return e;
  if ((e = d.parent.upload(T.parent)) != hipSuccess) return e;
  return d.leaves.upload(T.leaves);
}

// Coordinates of every point in the unit box of its leaf, (x - bbmin) / bbsize:
// the argument of the leaf basis (NystromMoments::moment_matrix,
// NystromMoments.h:24-34).  Formed once in FP64, so that the FP32 kernels see
// it rounded once instead of amplifying the rounding of absolute coordinates
// by 1 / bbsize.
std::vector<double> leaf_unit_coords(const Tree &tr, const std::vector<double> &pts) {
  const int d = tr.dim;
  const size_t n = (size_t)tr.npts;
  std::vector<double> u(pts.size());
  parallel_ranges((int64_t)tr.leaves.size(), [&](int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
      const int c = tr.leaves[k];
      for (int i = 0; i < d; ++i) {
        const double mn = tr.bbmin[(size_t)c * d + i], sz = tr.bbsize[(size_t)c * d + i];
        for (int64_t j = tr.begin[c]; j < tr.begin[c] + tr.size[c]; ++j)
          u[i * n + j] = (pts[i * n + j] - mn) / sz;
      }
    }
  });
  return u;
}

template <typename T>
hipError_t upload_reals(DevTreeReals<T> &d, const Tree &tr, const std::vector<double> &pts,
                        const std::vector<double> &unit, const std::vector<double> &transfer) {
  hipError_t e;
  if ((e = d.pts.upload(convert<T>(pts))) != hipSuccess) return e;
  if ((e = d.unit.upload(convert<T>(unit))) != hipSuccess) return e;
  if ((e = d.bbmin.upload(convert<T>(tr.bbmin))) != hipSuccess) return e;
  if ((e = d.bbsize.upload(convert<T>(tr.bbsize))) != hipSuccess) return e;
  return d.transfer.upload(convert<T>(transfer));
}

template <typename T>
DevTreeView<T> make_view(const Tree &tr, const DevTreeInts &di, const DevTreeReals<T> &dr) {
  DevTreeView<T> v;
  v.dim = tr.dim;
  v.ncl = tr.ncl;
  v.nleaves = (int)tr.leaves.size();
  v.depth = tr.depth;
  v.npts = tr.npts;
  v.perm = di.perm.as<uint32_t>();
  v.pinv = di.pinv.as<uint32_t>();
  v.pts = dr.pts.template as<T>();
  v.unit = dr.unit.template as<T>();
  v.bbmin = dr.bbmin.template as<T>();
  v.bbsize = dr.bbsize.template as<T>();
  v.begin = di.begin.as<int32_t>();
  v.size = di.size.as<int32_t>();
  v.son0 = di.son0.as<int32_t>();
  v.son1 = di.son1.as<int32_t>();
  v.parent = di.parent.as<int32_t>();
  v.leaves = di.leaves.as<int32_t>();
  v.transfer = dr.transfer.template as<T>();
  return v;
}

fmca_status ensure_device(fmca_h2_actor *a) {
  if (a->uploaded) {
    HIP_TRY(hipSetDevice(a->device));
    return FMCA_OK;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(FMCA_ERR_NO_DEVICE, "no HIP device available");
  HIP_TRY(hipGetDevice(&a->device));
  HIP_TRY(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&a->order_ev, hipEventDisableTiming));
  HIP_TRY(upload_ints(a->d_rows_i, a->rows()));
  {
    const std::vector<double> unit = leaf_unit_coords(a->rows(), a->pts_rows);
    HIP_TRY(upload_reals<double>(a->d_rows_64, a->rows(), a->pts_rows, unit, a->trans_rows));
    HIP_TRY(upload_reals<float>(a->d_rows_32, a->rows(), a->pts_rows, unit, a->trans_rows));
  }
  if (!a->shared) {
    HIP_TRY(upload_ints(a->d_cols_i, a->cols()));
    const std::vector<double> unit = leaf_unit_coords(a->cols(), a->pts_cols);
    HIP_TRY(upload_reals<double>(a->d_cols_64, a->cols(), a->pts_cols, unit, a->trans_cols));
    HIP_TRY(upload_reals<float>(a->d_cols_32, a->cols(), a->pts_cols, unit, a->trans_cols));
  }
  HIP_TRY(a->d_pat.upload(a->pat));
  if (a->ip.scheme == IS_TOTAL_DEGREE) {
    const int q = a->ip.q, d = a->dim;
    std::vector<double> un((size_t)d * q);
    for (int k = 0; k < q; ++k)
      for (int i = 0; i < d; ++i) un[(size_t)i * q + k] = a->ip.nodes[(size_t)d * k + i];
    HIP_TRY(a->d_midx.upload(a->ip.midx));
    HIP_TRY(a->d_unodes64.upload(un));
    HIP_TRY(a->d_unodes32.upload(std::vector<float>(un.begin(), un.end())));
    HIP_TRY(a->d_invV64.upload(a->ip.invV));
    HIP_TRY(a->d_invV32.upload(std::vector<float>(a->ip.invV.begin(), a->ip.invV.end())));
  }
  a->uploaded = true;
  return FMCA_OK;
}

template <typename T>
InterpParams<T> interp_params(const fmca_h2_actor *a) {
  const Interpolation &ip = a->ip;
  InterpParams<T> p;
  p.P1 = ip.P1;
  p.q = ip.q;
  for (int k = 0; k < kMaxP1; ++k) {
    p.xi[k] = k < (int)ip.xi.size() ? (T)ip.xi[k] : T(0);
    p.w[k] = k < (int)ip.w.size() ? (T)ip.w[k] : T(0);
  }
  if (ip.scheme == IS_TOTAL_DEGREE) {
    const bool f64 = sizeof(T) == sizeof(double);
    p.td = 1;
    p.midx = a->d_midx.as<int32_t>();
    p.unodes = (f64 ? a->d_unodes64 : a->d_unodes32).as<T>();
    p.invV = (f64 ? a->d_invV64 : a->d_invV32).as<T>();
  }
  return p;
}

// Kernel variant of every 16-column group of the rhs of one action.
struct VariantGroups {
  int n = 1;
  int v[kMaxVariantGroups] = {KV_VALUE};
  static VariantGroups single(int variant) {
    VariantGroups g;
    g.v[0] = variant;
    return g;
  }
  bool uniform() const {
    for (int i = 1; i < n; ++i)
      if (v[i] != v[0]) return false;
    return true;
  }
};

template <typename T>
KernelParams<T> kernel_params(const fmca_h2_actor *a, const VariantGroups &g) {
  KernelParams<T> k;
  k.kt = a->kt + KT_COUNT * g.v[0];
  k.vmask = 0;
  for (int i = 0; i < kMaxVariantGroups; ++i) {
    k.tv[i] = g.v[std::min(i, g.n - 1)];
    k.vmask |= 1 << k.tv[i];
  }
  k.l = (T)a->l;
  k.c = (T)a->c;
  k.nu = (T)a->nu;
  const double sc = kernel_radial_scale(a->kt, a->l);
  k.s = (T)sc;
  k.sl2 = (T)(sc * 1.4426950408889634074);
  k.inv_l = (T)(1.0 / a->l);
  k.c2 = (T)(a->c * a->c);
  k.geodesic = a->geodesic;
  return k;
}

// A workspace region the caller provides (the device scratch of
// fmca_h2_action_device_noalloc): xs, mom and tlhs are carved from it.
struct Region {
  void *base;
  size_t bytes;
};

constexpr size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

// Enqueue one action on `stream`.  transpose swaps the roles of the trees.
template <typename T>
fmca_status enqueue(fmca_h2_actor *a, Workspace<T> &ws, const DevTreeReals<T> &r_rows,
                    const DevTreeReals<T> &r_cols, const double *d_rhs, int64_t ld_rhs,
                    int64_t nrhs, double *d_out, int64_t ld_out, bool natural, bool transpose,
                    hipStream_t stream, const VariantGroups &groups, const Region *fixed,
                    bool split_bf16 = false, const GradJob *grad = nullptr) {
  const Tree &tr_rows = a->rows();
  const Tree &tr_cols = a->cols();
  const DevTreeInts &i_cols = a->shared ? a->d_rows_i : a->d_cols_i;
  DevTreeView<T> v_rows = make_view<T>(tr_rows, a->d_rows_i, r_rows);
  DevTreeView<T> v_cols = make_view<T>(tr_cols, i_cols, r_cols);
  const Pattern *pat = &a->pat;
  const DevPattern *dpat = &a->d_pat;
  if (transpose) {
    std::swap(v_rows, v_cols);
    if (!a->shared) {
      pat = &a->pat_t;
      dpat = &a->d_pat_t;
    }
  }
  const Tree &t_tgt = transpose ? tr_cols : tr_rows;
  const Tree &t_src = transpose ? tr_rows : tr_cols;
  HostLevels l_tgt{t_tgt.depth, t_tgt.level_ptr.data()};
  HostLevels l_src{t_src.depth, t_src.level_ptr.data()};
  const int q = a->ip.q;
  const size_t b_xs = (size_t)t_src.npts * nrhs * sizeof(T);
  const size_t b_mom = (size_t)t_src.ncl * q * nrhs * sizeof(T);
  const size_t b_tlhs = (size_t)t_tgt.ncl * q * nrhs * sizeof(T);
  ActionBuffers<T> buf;
  if (fixed) {
    // the caller's region: xs | mom | tlhs, each start aligned to 256 bytes
    const size_t o_mom = align_up(b_xs), o_tlhs = o_mom + align_up(b_mom);
    if (o_tlhs + b_tlhs > fixed->bytes)
      return fail(FMCA_ERR_NOT_SUPPORTED,
                  "the action needs " + std::to_string(o_tlhs + b_tlhs) +
                      " bytes of workspace, the device scratch holds " +
                      std::to_string(fixed->bytes));
    unsigned char *base = static_cast<unsigned char *>(fixed->base);
    buf = ActionBuffers<T>{reinterpret_cast<T *>(base), reinterpret_cast<T *>(base + o_mom),
                           reinterpret_cast<T *>(base + o_tlhs)};
  } else {
    HIP_TRY(ws.xs.alloc(b_xs));
    HIP_TRY(ws.mom.alloc(b_mom));
    HIP_TRY(ws.tlhs.alloc(b_tlhs));
    buf = ActionBuffers<T>{ws.xs.template as<T>(), ws.mom.template as<T>(),
                           ws.tlhs.template as<T>()};
  }
  PhaseEvents *pe = nullptr;
  if (a->profiling && !fixed) {
    PhaseEvents ev;
    for (auto &x : ev.ev) {
      if (a->ev_pool.empty()) {
        HIP_TRY(hipEventCreate(&x));
      } else {
        x = a->ev_pool.back();
        a->ev_pool.pop_back();
      }
    }
    a->pending.emplace_back(sizeof(T) == 8   ? FMCA_PRECISION_FP64
                            : split_bf16 ? FMCA_PRECISION_FP32_BF16
                                         : FMCA_PRECISION_FP32,
                            ev);
    pe = &a->pending.back().second;
  }
  // the assembled blocks serve plain forward actions of their precision
  StoredBlocks<T> st;
  const bool use_stored = a->st_prec == (sizeof(T) == 8 ? FMCA_PRECISION_FP64 : FMCA_PRECISION_FP32) &&
                          !transpose && groups.uniform() && groups.v[0] == KV_VALUE;
  if (use_stored) {
    st.far = a->st_far.template as<T>();
    st.near = a->st_near.template as<T>();
    st.near_off = a->st_near_off.template as<int64_t>();
    st.v_off = a->st_v_off.template as<int64_t>();
    st.far_order = a->st_far_order.template as<int32_t>();
    st.near_order = a->st_near_order.template as<int32_t>();
    st.Rq = a->st_Rq;
    st.Kq = a->st_Kq;
  }
  KernelParams<T> kp = kernel_params<T>(a, groups);
  kp.split_bf16 = split_bf16 ? 1 : 0;
  hipError_t e = launch_action<T>(v_rows, v_cols, l_tgt, l_src, dpat->view(*pat),
                                  interp_params<T>(a), kp, d_rhs, ld_rhs,
                                  (int)nrhs, d_out, ld_out, natural, buf, stream, pe,
                                  use_stored ? &st : nullptr, grad);
  return from_hip(e, "launch_action");
}

// Writes the assembled blocks of precision T (fmca_h2_assemble).
template <typename T>
fmca_status assemble(fmca_h2_actor *a, const DevTreeReals<T> &r_rows,
                     const DevTreeReals<T> &r_cols) {
  const Tree &tr = a->rows();
  const Pattern &pat = a->pat;
  const int q = a->ip.q;
  const int Rq = (q + 15) & ~15, Kq = stored_kpad<T>(q);
  const int nl = (int)tr.leaves.size();
  // block -> target tables and the offsets of the leaf blocks
  std::vector<int32_t> far_tgt(pat.far_src.size()), near_leaf(pat.near_src.size());
  for (int t = 0; t < tr.ncl; ++t)
    for (int b = pat.far_ptr[t]; b < pat.far_ptr[t + 1]; ++b) far_tgt[b] = t;
  std::vector<int64_t> near_off(pat.near_src.size()), v_off(nl);
  int64_t off = 0;
  const Tree &tc = a->cols();
  for (int k = 0; k < nl; ++k) {
    const int64_t Rt = (tr.size[tr.leaves[k]] + 15) & ~15;
    v_off[k] = off;
    off += Rt * Kq;
    for (int e = pat.near_ptr[k]; e < pat.near_ptr[k + 1]; ++e) {
      near_leaf[e] = k;
      near_off[e] = off;
      off += Rt * stored_kpad<T>((int)tc.size[pat.near_src[e]]);
    }
  }
  // launch orders of the streaming kernels: longest list first (stable)
  std::vector<int32_t> far_order(tr.ncl), near_order(nl);
  std::iota(far_order.begin(), far_order.end(), 0);
  std::stable_sort(far_order.begin(), far_order.end(), [&](int x, int y) {
    return pat.far_ptr[x + 1] - pat.far_ptr[x] > pat.far_ptr[y + 1] - pat.far_ptr[y];
  });
  std::vector<int64_t> leaf_work(nl);
  for (int k = 0; k < nl; ++k)
    leaf_work[k] = (k + 1 < nl ? v_off[k + 1] : off) - v_off[k];
  std::iota(near_order.begin(), near_order.end(), 0);
  std::stable_sort(near_order.begin(), near_order.end(),
                   [&](int x, int y) { return leaf_work[x] > leaf_work[y]; });
  const int64_t nfar = (int64_t)pat.far_src.size(), nnear = (int64_t)pat.near_src.size();
  HIP_TRY(a->st_far.alloc((size_t)nfar * Rq * Kq * sizeof(T)));
  HIP_TRY(a->st_near.alloc((size_t)off * sizeof(T)));
  HIP_TRY(a->st_near_off.upload(near_off));
  a->st_near_off_host = near_off;
  HIP_TRY(a->st_v_off.upload(v_off));
  HIP_TRY(a->st_far_tgt.upload(far_tgt));
  HIP_TRY(a->st_near_leaf.upload(near_leaf));
  HIP_TRY(a->st_far_order.upload(far_order));
  HIP_TRY(a->st_near_order.upload(near_order));
  const DevTreeInts &i_cols = a->shared ? a->d_rows_i : a->d_cols_i;
  DevTreeView<T> v_rows = make_view<T>(tr, a->d_rows_i, r_rows);
  DevTreeView<T> v_cols = make_view<T>(tc, i_cols, r_cols);
  HIP_TRY(launch_assemble<T>(v_rows, v_cols, a->d_pat.view(pat), interp_params<T>(a),
                             kernel_params<T>(a, VariantGroups::single(KV_VALUE)), nfar,
                             a->st_far_tgt.template as<int32_t>(), nnear,
                             a->st_near_leaf.template as<int32_t>(), a->st_far.template as<T>(),
                             a->st_near.template as<T>(), a->st_near_off.template as<int64_t>(),
                             a->st_v_off.template as<int64_t>(), Rq, Kq, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  a->st_Rq = Rq;
  a->st_Kq = Kq;
  a->st_bytes = (int64_t)((size_t)nfar * Rq * Kq + off) * (int64_t)sizeof(T) +
                (int64_t)(near_off.size() + v_off.size()) * 8 +
                (int64_t)(far_tgt.size() + near_leaf.size() + far_order.size() +
                          near_order.size()) * 4;
  return FMCA_OK;
}

fmca_status enqueue_any_unordered(fmca_h2_actor *a, const double *d_rhs, int64_t ld_rhs,
                                  int64_t nrhs, double *d_out, int64_t ld_out, int precision,
                                  bool natural, bool transpose, hipStream_t stream,
                                  const VariantGroups &groups, const Region *fixed,
                                  const GradJob *grad);

// One action on `stream`, ordered after the actor's previous action when that
// ran on another stream: the workspaces (permuted rhs, moments, sums) are
// shared by every action of the actor.
fmca_status enqueue_any(fmca_h2_actor *a, const double *d_rhs, int64_t ld_rhs, int64_t nrhs,
                        double *d_out, int64_t ld_out, int precision, bool natural,
                        bool transpose, hipStream_t stream,
                        const VariantGroups &groups = VariantGroups::single(KV_VALUE),
                        const Region *fixed = nullptr, const GradJob *grad = nullptr) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(stream, &cs));
  const bool captured = cs != hipStreamCaptureStatusNone;
  if (!captured && a->order_valid && a->order_stream != stream)
    HIP_TRY(hipStreamWaitEvent(stream, a->order_ev, 0));
  fmca_status st = enqueue_any_unordered(a, d_rhs, ld_rhs, nrhs, d_out, ld_out, precision,
                                         natural, transpose, stream, groups, fixed, grad);
  if (st != FMCA_OK || captured) return st;
  HIP_TRY(hipEventRecord(a->order_ev, stream));
  a->order_stream = stream;
  a->order_valid = true;
  return FMCA_OK;
}

fmca_status enqueue_any_unordered(fmca_h2_actor *a, const double *d_rhs, int64_t ld_rhs,
                                  int64_t nrhs, double *d_out, int64_t ld_out, int precision,
                                  bool natural, bool transpose, hipStream_t stream,
                                  const VariantGroups &groups, const Region *fixed,
                                  const GradJob *grad) {
  if (a->geodesic) {
    // near the diagonal acos turns the rounding of an FP32 dot product into
    // an error of about 1e-7 / r in r, and the derivative in the points of a
    // geodesic kernel is not the Euclidean one of the gradient kernels
    if (precision != FMCA_PRECISION_FP64)
      return fail(FMCA_ERR_NOT_SUPPORTED, "the geodesic distance is evaluated in FP64 only");
    bool phi = grad != nullptr;
    for (int i = 0; i < groups.n; ++i) phi = phi || groups.v[i] == KV_PHI;
    if (phi)
      return fail(FMCA_ERR_NOT_SUPPORTED,
                  "derivatives in the points are Euclidean only; the geodesic distance "
                  "supports kernel values and d/dl");
  }
  if (transpose && !a->shared && !a->have_pat_t) {
    compute_pattern(a->pat_t, a->cols(), a->rows(), a->eta, a->ip.q);
    HIP_TRY(a->d_pat_t.upload(a->pat_t));
    a->have_pat_t = true;
  }
  if (precision == FMCA_PRECISION_FP64)
    return enqueue<double>(a, a->ws64, a->d_rows_64, a->shared ? a->d_rows_64 : a->d_cols_64,
                           d_rhs, ld_rhs, nrhs, d_out, ld_out, natural, transpose, stream,
                           groups, fixed, false, grad);
  if (precision == FMCA_PRECISION_FP32 || precision == FMCA_PRECISION_FP32_BF16)
    return enqueue<float>(a, a->ws32, a->d_rows_32, a->shared ? a->d_rows_32 : a->d_cols_32,
                          d_rhs, ld_rhs, nrhs, d_out, ld_out, natural, transpose, stream,
                          groups, fixed, precision == FMCA_PRECISION_FP32_BF16, grad);
  return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
}

// One action whose 16-column groups carry different variants, when both
// entry phases run on MFMA (the VALU forms evaluate one variant per launch).
template <typename T>
bool groups_in_one_action(const fmca_h2_actor *a, const VariantGroups &g, int64_t nrhs) {
  if (g.uniform()) return true;
  KernelParams<T> kp = kernel_params<T>(a, g);
  const int td = a->ip.scheme == IS_TOTAL_DEGREE;
  return mfma_supported<T>(true, a->dim, a->ip.P1, a->ip.q, (int)nrhs, kp.kt, kp.vmask, td) &&
         mfma_supported<T>(false, a->dim, a->ip.P1, a->ip.q, (int)nrhs, kp.kt, kp.vmask, td);
}

fmca_status host_action(fmca_h2_actor *a, const double *rhs, int64_t ld_rhs, int64_t nrhs,
                        double *out, int64_t ld_out, int precision, bool natural) {
  if (!a || (!rhs && nrhs) || (!out && nrhs) || nrhs < 0)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or negative nrhs");
  if (ld_rhs < a->n || ld_out < a->m) return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (nrhs == 0) return FMCA_OK;
  std::lock_guard<std::mutex> lock(a->mtx);
  fmca_status st = ensure_device(a);
  if (st != FMCA_OK) return st;
  const size_t in_bytes = (size_t)a->n * nrhs * sizeof(double);
  const size_t out_bytes = (size_t)a->m * nrhs * sizeof(double);
  HIP_TRY(a->stage_in.alloc(in_bytes));
  HIP_TRY(a->stage_out.alloc(out_bytes));
  HIP_TRY(hipMemcpy2DAsync(a->stage_in.p, a->n * sizeof(double), rhs, ld_rhs * sizeof(double),
                           a->n * sizeof(double), nrhs, hipMemcpyHostToDevice, a->stream));
  st = enqueue_any(a, a->stage_in.as<double>(), a->n, nrhs, a->stage_out.as<double>(), a->m,
                   precision, natural, false, a->stream);
  if (st != FMCA_OK) return st;
  HIP_TRY(hipMemcpy2DAsync(out, ld_out * sizeof(double), a->stage_out.p, a->m * sizeof(double),
                           a->m * sizeof(double), nrhs, hipMemcpyDeviceToHost, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return FMCA_OK;
}

// Device scratch of fmca_h2_action_device_noalloc: a fixed array in the code
// object, so that the entry point allocates nothing.  Each device holds its
// own copy until the process ends; the host learns its address per device
// with hipGetSymbolAddress (the array's name is not a device address in host
// code).  Every call on a device shares that device's copy.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mtx;
void *g_scratch_ptr[kMaxScratchDevices] = {nullptr};

// The device address of `scratch` on `device` (which is current).
fmca_status scratch_on_device(int device, void **ptr) {
  if (device < 0 || device >= kMaxScratchDevices)
    return fail(FMCA_ERR_NOT_SUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mtx);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FMCA_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return FMCA_OK;
}

}  // namespace

namespace fmca_amd {
int device_malloc(void **p, size_t bytes) {
  const size_t h = std::max<size_t>(bytes, 256);
  const int64_t cap = g_dev_limit.load();
  if (cap > 0 && g_dev_live.load() + (int64_t)h > cap) {
    *p = nullptr;
    return (int)hipErrorOutOfMemory;
  }
  const hipError_t e = hipMalloc(p, h);
  if (e == hipSuccess) note_device_bytes((int64_t)h);
  return (int)e;
}

void device_free(void *p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  note_device_bytes(-(int64_t)std::max<size_t>(bytes, 256));
}
}  // namespace fmca_amd

extern "C" {

const char *fmca_status_string(fmca_status s) {
  switch (s) {
    case FMCA_OK:
      return "ok";
    case FMCA_ERR_INVALID_ARGUMENT:
      return "invalid argument";
    case FMCA_ERR_UNKNOWN_KERNEL:
      return "unknown kernel type";
    case FMCA_ERR_NOT_SUPPORTED:
      return "not supported";
    case FMCA_ERR_NO_DEVICE:
      return "no HIP device";
    case FMCA_ERR_OUT_OF_MEMORY:
      return "out of device memory";
    case FMCA_ERR_DEVICE:
      return "device error";
    case FMCA_ERR_NOT_CONVERGED:
      return "not converged";
    case FMCA_ERR_NONFINITE:
      return "non-finite value";
  }
  return "unknown status";
}

const char *fmca_last_error(void) { return g_last_error.c_str(); }


fmca_status fmca_h2_create(const fmca_kernel_desc *kernel, int dim, int64_t m,
                           const double *Peval, int64_t n, const double *P, int deg, double eta,
                           fmca_h2_actor **actor) {
  return fmca_h2_create_interp(kernel, dim, m, Peval, n, P, deg, eta,
                               FMCA_INTERP_TENSOR_PRODUCT, actor);
}

fmca_status fmca_h2_create_interp(const fmca_kernel_desc *kernel, int dim, int64_t m,
                                  const double *Peval, int64_t n, const double *P, int deg,
                                  double eta, int interpolator, fmca_h2_actor **actor) {
  return fmca_h2_create_ex(kernel, dim, m, Peval, n, P, deg, eta, interpolator, FMCA_PLAN_HOST,
                           actor);
}

fmca_status fmca_h2_create_ex(const fmca_kernel_desc *kernel, int dim, int64_t m,
                              const double *Peval, int64_t n, const double *P, int deg,
                              double eta, int interpolator, int planner,
                              fmca_h2_actor **actor) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  if (planner != FMCA_PLAN_HOST && planner != FMCA_PLAN_DEVICE)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown planner");
  *actor = nullptr;
  if (!kernel || !kernel->ktype) return fail(FMCA_ERR_INVALID_ARGUMENT, "kernel is null");
  const int kt = parse_kernel_type(kernel->ktype);
  if (kt < 0) return fail(FMCA_ERR_UNKNOWN_KERNEL, std::string("kernel ") + kernel->ktype);
  if (dim < 1 || dim > kMaxDim)
    return fail(FMCA_ERR_NOT_SUPPORTED, "dimension must be 1, 2 or 3");
  if (deg < 0 || deg + 1 > kMaxP1) return fail(FMCA_ERR_NOT_SUPPORTED, "degree must be 0..9");
  if (interpolator != FMCA_INTERP_TENSOR_PRODUCT && interpolator != FMCA_INTERP_TOTAL_DEGREE)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown interpolator");
  if (m < 1 || n < 1 || !P || !Peval) return fail(FMCA_ERR_INVALID_ARGUMENT, "empty point set");
  if (n > (int64_t)INT32_MAX || m > (int64_t)INT32_MAX)
    return fail(FMCA_ERR_NOT_SUPPORTED, "more than 2^31-1 points");
  if (!(eta >= 0) || !std::isfinite(eta)) return fail(FMCA_ERR_INVALID_ARGUMENT, "eta");
  if (!(kernel->l > 0) || !std::isfinite(kernel->l))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "lengthscale must be positive");
  for (int64_t k = 0; k < (int64_t)dim * n; ++k)
    if (!std::isfinite(P[k])) return fail(FMCA_ERR_NONFINITE, "non-finite source coordinate");
  for (int64_t k = 0; k < (int64_t)dim * m; ++k)
    if (!std::isfinite(Peval[k])) return fail(FMCA_ERR_NONFINITE, "non-finite target coordinate");
  std::unique_ptr<fmca_h2_actor> a;
  try {
    a.reset(new fmca_h2_actor());
    a->dim = dim;
    a->deg = deg;
    a->m = m;
    a->n = n;
    a->eta = eta;
    a->kt = kt;
    a->l = kernel->l;
    a->c = kernel->c;
    a->nu = kernel->nu;
    a->ip.init(dim, deg,
               interpolator == FMCA_INTERP_TOTAL_DEGREE ? IS_TOTAL_DEGREE : IS_TENSOR);
    // H2ClusterTree::init (H2ClusterTree.h:65-74): minimum cluster size q
    a->shared = (P == Peval && m == n) ||
                (m == n && std::memcmp(P, Peval, sizeof(double) * dim * n) == 0);
    const auto t0 = std::chrono::steady_clock::now();
    // ClusterTreeInitializer<ClusterTree>::init for both trees, on the host or
    // on the current device
    auto tree = [&](Tree &T, int64_t cnt, const double *pts, int &where) -> fmca_status {
      if (planner == FMCA_PLAN_HOST) {
        build_tree(T, dim, cnt, pts, a->ip.q);
        where = 0;
        return FMCA_OK;
      }
      std::string err;
      const int r = build_tree_device(T, dim, cnt, pts, a->ip.q, &err);
      if (r == -(int)hipErrorOutOfMemory) {  // no room on the device: plan on the host
        build_tree(T, dim, cnt, pts, a->ip.q);
        where = 2;
        return FMCA_OK;
      }
      if (r < 0) return fail(from_hip((hipError_t)(-r), "tree"), "device tree build: " + err);
      where = r == 1 ? 1 : 2;
      return FMCA_OK;
    };
    fmca_status st = tree(a->rows_tree, m, Peval, a->rows_builder);
    if (st == FMCA_OK && !a->shared) st = tree(a->cols_tree, n, P, a->cols_builder);
    if (a->shared) a->cols_builder = a->rows_builder;
    if (st != FMCA_OK) return st;
    const auto t1 = std::chrono::steady_clock::now();
    if (planner == FMCA_PLAN_HOST) {
      compute_pattern(a->pat, a->rows(), a->cols(), eta, a->ip.q);
    } else {
      std::string err;
      const int r = expand_blocks_device(a->pat, a->rows(), a->cols(), eta, a->ip.q, &err);
      if (r < 0 && r != -(int)hipErrorOutOfMemory)
        return fail(from_hip((hipError_t)(-r), "pattern"), "device block expansion: " + err);
      if (r <= 0) {  // near the threshold, or no room on the device
        a->pat = Pattern();
        expand_blocks(a->pat, a->rows(), a->cols(), eta, a->ip.q);
      }
      a->pattern_builder = r == 1 ? 1 : 2;
      finish_pattern(a->pat, a->rows(), a->cols(), a->ip.q);
    }
    a->tree_s = std::chrono::duration<double>(t1 - t0).count();
    a->pattern_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    auto transfers = [&](std::vector<double> &T, const Tree &tr) {
      if (a->ip.scheme == IS_TOTAL_DEGREE)
        compute_transfers_dense(T, tr, a->ip);
      else
        compute_transfers(T, tr, a->ip);
    };
    transfers(a->trans_rows, a->rows());
    a->pts_rows = cluster_order_points(a->rows(), Peval);
    if (!a->shared) {
      transfers(a->trans_cols, a->cols());
      a->pts_cols = cluster_order_points(a->cols(), P);
    }
  } catch (const std::bad_alloc &) {
    return fail(FMCA_ERR_OUT_OF_MEMORY, "host allocation failed while planning");
  }
  *actor = a.release();
  return FMCA_OK;
}

fmca_status fmca_h2_plan_report(const fmca_h2_actor *actor, double *tree_seconds,
                                double *pattern_seconds, int *where) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  if (tree_seconds) *tree_seconds = actor->tree_s;
  if (pattern_seconds) *pattern_seconds = actor->pattern_s;
  if (where) {
    where[0] = actor->rows_builder;
    where[1] = actor->cols_builder;
    where[2] = actor->pattern_builder;
  }
  return FMCA_OK;
}

fmca_status fmca_h2_destroy(fmca_h2_actor *actor) {
  if (!actor) return FMCA_OK;
  if (actor->uploaded) (void)hipSetDevice(actor->device);
  delete actor;
  return FMCA_OK;
}

fmca_status fmca_h2_upload(fmca_h2_actor *actor) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  std::lock_guard<std::mutex> lock(actor->mtx);
  return ensure_device(actor);
}

fmca_status fmca_h2_action(fmca_h2_actor *actor, const double *rhs, int64_t ld_rhs, int64_t nrhs,
                           double *out, int64_t ld_out, int precision) {
  return host_action(actor, rhs, ld_rhs, nrhs, out, ld_out, precision, true);
}

fmca_status fmca_h2_action_cluster_order(fmca_h2_actor *actor, const double *rhs, int64_t ld_rhs,
                                         int64_t nrhs, double *out, int64_t ld_out,
                                         int precision) {
  return host_action(actor, rhs, ld_rhs, nrhs, out, ld_out, precision, false);
}

fmca_status fmca_h2_action_device(fmca_h2_actor *actor, const double *d_rhs, int64_t ld_rhs,
                                  int64_t nrhs, double *d_out, int64_t ld_out, int precision,
                                  int cluster_order, void *stream) {
  if (!actor || nrhs < 0 || (nrhs && (!d_rhs || !d_out)))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or negative nrhs");
  if (ld_rhs < actor->n || ld_out < actor->m)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (nrhs == 0) return FMCA_OK;
  std::lock_guard<std::mutex> lock(actor->mtx);
  fmca_status st = ensure_device(actor);
  if (st != FMCA_OK) return st;
  return enqueue_any(actor, d_rhs, ld_rhs, nrhs, d_out, ld_out, precision, cluster_order == 0,
                     false, (hipStream_t)stream);
}

fmca_status fmca_h2_action_device_noalloc(fmca_h2_actor *actor, const double *d_rhs,
                                          int64_t ld_rhs, int64_t nrhs, double *d_out,
                                          int64_t ld_out, int precision, int cluster_order,
                                          void *stream) {
  if (!actor || nrhs < 0 || (nrhs && (!d_rhs || !d_out)))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or negative nrhs");
  if (ld_rhs < actor->n || ld_out < actor->m)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (nrhs == 0) return FMCA_OK;
  std::lock_guard<std::mutex> lock(actor->mtx);
  if (!actor->uploaded)
    return fail(FMCA_ERR_INVALID_ARGUMENT,
                "fmca_h2_action_device_noalloc: the plan is not on the device yet "
                "(call fmca_h2_upload or one allocating action first)");
  HIP_TRY(hipSetDevice(actor->device));
  void *base = nullptr;
  fmca_status st = scratch_on_device(actor->device, &base);
  if (st != FMCA_OK) return st;
  const Region region{base, kScratchBytes};
  return enqueue_any(actor, d_rhs, ld_rhs, nrhs, d_out, ld_out, precision, cluster_order == 0,
                     false, (hipStream_t)stream, VariantGroups::single(KV_VALUE), &region);
}

fmca_status fmca_h2_action_transpose_device(fmca_h2_actor *actor, const double *d_rhs,
                                            int64_t ld_rhs, int64_t nrhs, double *d_out,
                                            int64_t ld_out, int precision, void *stream) {
  if (!actor || nrhs < 0 || (nrhs && (!d_rhs || !d_out)))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or negative nrhs");
  if (ld_rhs < actor->m || ld_out < actor->n)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (nrhs == 0) return FMCA_OK;
  std::lock_guard<std::mutex> lock(actor->mtx);
  fmca_status st = ensure_device(actor);
  if (st != FMCA_OK) return st;
  return enqueue_any(actor, d_rhs, ld_rhs, nrhs, d_out, ld_out, precision, true, true,
                     (hipStream_t)stream);
}

fmca_status fmca_h2_set_kernel(fmca_h2_actor *actor, const fmca_kernel_desc *kernel) {
  if (!actor || !kernel || !kernel->ktype) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  const int kt = parse_kernel_type(kernel->ktype);
  if (kt < 0) return fail(FMCA_ERR_UNKNOWN_KERNEL, std::string("kernel ") + kernel->ktype);
  if (!(kernel->l > 0) || !std::isfinite(kernel->l))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "lengthscale must be positive");
  std::lock_guard<std::mutex> lock(actor->mtx);
  // assembled blocks hold the old kernel's entries
  if (kt != actor->kt || kernel->l != actor->l || kernel->c != actor->c || kernel->nu != actor->nu)
    actor->release_stored();
  actor->kt = kt;
  actor->l = kernel->l;
  actor->c = kernel->c;
  actor->nu = kernel->nu;
  return FMCA_OK;
}

fmca_status fmca_h2_set_distance_type(fmca_h2_actor *actor, const char *dist_type) {
  if (!actor || !dist_type) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  std::string d(dist_type);
  for (auto &ch : d) ch = (char)std::toupper((unsigned char)ch);
  int geo;
  if (d == "EUCLIDEAN")
    geo = 0;
  else if (d == "GEODESIC")
    geo = 1;
  else
    return fail(FMCA_ERR_INVALID_ARGUMENT, "desired distance not implemented: " + d);
  std::lock_guard<std::mutex> lock(actor->mtx);
  if (geo != actor->geodesic) actor->release_stored();
  actor->geodesic = geo;
  return FMCA_OK;
}

fmca_status fmca_h2_distance_type(const fmca_h2_actor *actor, int *geodesic) {
  if (!actor || !geodesic) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  *geodesic = actor->geodesic;
  return FMCA_OK;
}

fmca_status fmca_h2_assemble(fmca_h2_actor *actor, int precision, int64_t *bytes) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  if (precision != FMCA_PRECISION_FP64 && precision != FMCA_PRECISION_FP32)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
  if (actor->ip.q > kStoredMaxRows / 2)
    return fail(FMCA_ERR_NOT_SUPPORTED, "assembled blocks need q <= 128");
  if (actor->geodesic && precision != FMCA_PRECISION_FP64)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the geodesic distance is evaluated in FP64 only");
  std::lock_guard<std::mutex> lock(actor->mtx);
  fmca_status st = ensure_device(actor);
  if (st != FMCA_OK) return st;
  actor->release_stored();
  st = precision == FMCA_PRECISION_FP64
           ? assemble<double>(actor, actor->d_rows_64,
                              actor->shared ? actor->d_rows_64 : actor->d_cols_64)
           : assemble<float>(actor, actor->d_rows_32,
                             actor->shared ? actor->d_rows_32 : actor->d_cols_32);
  if (st != FMCA_OK) {
    const std::string msg = fmca_last_error();
    actor->release_stored();
    return fail(st, msg);
  }
  actor->st_prec = precision;
  if (bytes) *bytes = actor->st_bytes;
  return FMCA_OK;
}

fmca_status fmca_h2_release_assembled(fmca_h2_actor *actor) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  std::lock_guard<std::mutex> lock(actor->mtx);
  if (actor->uploaded) HIP_TRY(hipSetDevice(actor->device));
  actor->release_stored();
  return FMCA_OK;
}

fmca_status fmca_h2_action_variant_device(fmca_h2_actor *actor, int variant, int transpose,
                                          const double *d_rhs, int64_t ld_rhs, int64_t nrhs,
                                          double *d_out, int64_t ld_out, int precision,
                                          void *stream) {
  if (!actor || nrhs < 0 || (nrhs && (!d_rhs || !d_out)))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or negative nrhs");
  if (variant < KV_VALUE || variant > KV_PHI) return fail(FMCA_ERR_INVALID_ARGUMENT, "variant");
  const int64_t rows_in = transpose ? actor->m : actor->n;
  const int64_t rows_out = transpose ? actor->n : actor->m;
  if (ld_rhs < rows_in || ld_out < rows_out)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (nrhs == 0) return FMCA_OK;
  std::lock_guard<std::mutex> lock(actor->mtx);
  fmca_status st = ensure_device(actor);
  if (st != FMCA_OK) return st;
  return enqueue_any(actor, d_rhs, ld_rhs, nrhs, d_out, ld_out, precision, true, transpose != 0,
                     (hipStream_t)stream, VariantGroups::single(variant));
}

fmca_status fmca_h2_action_variants_device(fmca_h2_actor *actor, const int *variants,
                                           int ngroups, int transpose, const double *d_rhs,
                                           int64_t ld_rhs, int64_t nrhs, double *d_out,
                                           int64_t ld_out, int precision, void *stream) {
  if (!actor || !variants || nrhs < 0 || (nrhs && (!d_rhs || !d_out)))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or negative nrhs");
  if (ngroups < 1 || ngroups > kMaxVariantGroups)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "ngroups must be in [1, 8]");
  if (nrhs > 16 * (int64_t)ngroups || nrhs <= 16 * (int64_t)(ngroups - 1))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "nrhs must fill the last 16-column group");
  VariantGroups g;
  g.n = ngroups;
  for (int i = 0; i < ngroups; ++i) {
    if (variants[i] < KV_VALUE || variants[i] > KV_PHI)
      return fail(FMCA_ERR_INVALID_ARGUMENT, "variant");
    g.v[i] = variants[i];
  }
  const int64_t rows_in = transpose ? actor->m : actor->n;
  const int64_t rows_out = transpose ? actor->n : actor->m;
  if (ld_rhs < rows_in || ld_out < rows_out)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (precision != FMCA_PRECISION_FP64 && precision != FMCA_PRECISION_FP32 &&
      precision != FMCA_PRECISION_FP32_BF16)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
  std::lock_guard<std::mutex> lock(actor->mtx);
  fmca_status st = ensure_device(actor);
  if (st != FMCA_OK) return st;
  const bool one = precision == FMCA_PRECISION_FP64 ? groups_in_one_action<double>(actor, g, nrhs)
                                                    : groups_in_one_action<float>(actor, g, nrhs);
  if (one)
    return enqueue_any(actor, d_rhs, ld_rhs, nrhs, d_out, ld_out, precision, true,
                       transpose != 0, (hipStream_t)stream, g);
  // one action per run of equal variants
  for (int g0 = 0; g0 < ngroups;) {
    int g1 = g0 + 1;
    while (g1 < ngroups && g.v[g1] == g.v[g0]) ++g1;
    const int64_t c0 = 16 * (int64_t)g0, c1 = std::min<int64_t>(16 * (int64_t)g1, nrhs);
    st = enqueue_any(actor, d_rhs + c0 * ld_rhs, ld_rhs, c1 - c0, d_out + c0 * ld_out, ld_out,
                     precision, true, transpose != 0, (hipStream_t)stream,
                     VariantGroups::single(g.v[g0]));
    if (st != FMCA_OK) return st;
    g0 = g1;
  }
  return FMCA_OK;
}

fmca_status fmca_h2_gradient_device(fmca_h2_actor *a, const double *d_alpha, int64_t ld_alpha,
                                   const double *d_g, int64_t ld_g, int64_t nrhs, int precision,
                                   double *d_dalpha, int64_t ld_dalpha, double *d_dpoints,
                                   int64_t ld_points, double *d_dl, void *stream_ptr) {
  if (!a || !d_alpha || !d_g || !d_dalpha || !d_dpoints || !d_dl)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  if (nrhs < 1 || nrhs > 16)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the gradient form takes 1 to 16 right-hand sides");
  if (!a->shared)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the gradient form needs one point set (targets = sources)");
  if (a->ip.scheme == IS_TOTAL_DEGREE)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the gradient form needs the tensor-product interpolator");
  if (a->geodesic)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the gradient form takes the Euclidean distance");
  if (ld_alpha < a->n || ld_g < a->n || ld_dalpha < a->n || ld_points < a->n)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (precision == FMCA_PRECISION_FP32_BF16) precision = FMCA_PRECISION_FP32;
  if (precision != FMCA_PRECISION_FP64 && precision != FMCA_PRECISION_FP32)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
  const int r = (int)nrhs, R = r <= 8 ? 8 : 16, ncol = 2 * R + 16;
  VariantGroups g;
  g.n = ncol / 16;
  for (int i = 0; i < g.n; ++i) g.v[i] = 16 * i < 2 * R ? KV_VALUE : KV_DL;
  const int vm = (1 << KV_VALUE) | (1 << KV_DL);
  const bool coupling_ok =
      precision == FMCA_PRECISION_FP64
          ? mfma_supported<double>(false, a->dim, a->ip.P1, a->ip.q, ncol, a->kt, vm, 0)
          : mfma_supported<float>(false, a->dim, a->ip.P1, a->ip.q, ncol, a->kt, vm, 0);
  if (!coupling_ok)
    return fail(FMCA_ERR_NOT_SUPPORTED,
                "the coupling of the gradient form does not fit the matrix cores");
  std::lock_guard<std::mutex> lock(a->mtx);
  fmca_status st = ensure_device(a);
  if (st != FMCA_OK) return st;
  hipStream_t sm = (hipStream_t)stream_ptr;
  const int64_t n = a->n;
  HIP_TRY(a->grad_x.alloc((size_t)n * ncol * sizeof(double)));
  const int nparts = (int)a->pat.near_order.size();
  HIP_TRY(a->grad_part.alloc((size_t)(nparts + grad_far_dot_blocks()) * sizeof(double)));
  double *X = a->grad_x.as<double>();
  // rhs [alpha | G | alpha] with zero padding, natural order
  HIP_TRY(launch_grad_zero(X, n * ncol, sm));
  const size_t w = (size_t)n * sizeof(double);
  HIP_TRY(hipMemcpy2DAsync(X, w, d_alpha, ld_alpha * sizeof(double), w, r,
                           hipMemcpyDeviceToDevice, sm));
  HIP_TRY(hipMemcpy2DAsync(X + (size_t)R * n, w, d_g, ld_g * sizeof(double), w, r,
                           hipMemcpyDeviceToDevice, sm));
  HIP_TRY(hipMemcpy2DAsync(X + (size_t)2 * R * n, w, d_alpha, ld_alpha * sizeof(double), w, r,
                           hipMemcpyDeviceToDevice, sm));
  for (int k = 0; k < r; ++k) HIP_TRY(launch_grad_zero(d_dalpha + k * ld_dalpha, n, sm));
  for (int d = 0; d < a->dim; ++d) HIP_TRY(launch_grad_zero(d_dpoints + d * ld_points, n, sm));
  GradJob job;
  job.r = r;
  job.R = R;
  job.ncol = ncol;
  job.dalpha = d_dalpha;
  job.ld_dalpha = ld_dalpha;
  job.dpoints = d_dpoints;
  job.ld_points = ld_points;
  job.dl_part = a->grad_part.as<double>();
  job.nnear_parts = nparts;
  job.dl_out = d_dl;
  return enqueue_any(a, X, n, ncol, nullptr, n, precision, true, false, sm, g, nullptr, &job);
}

fmca_status fmca_h2_indices(const fmca_h2_actor *actor, int which, uint32_t *out) {
  if (!actor || !out) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  const Tree &T = which == FMCA_ROWS ? actor->rows() : actor->cols();
  std::copy(T.perm.begin(), T.perm.end(), out);
  return FMCA_OK;
}

fmca_status fmca_h2_statistics(const fmca_h2_actor *actor, double out[6]) {
  if (!actor || !out) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  const Statistics s = statistics(actor->pat, actor->rows(), actor->cols(), actor->ip.q);
  out[0] = s.rows;
  out[1] = s.cols;
  out[2] = s.low_rank_blocks;
  out[3] = s.dense_blocks;
  out[4] = s.nz_per_row;
  out[5] = s.storage_gb;
  return FMCA_OK;
}

fmca_status fmca_h2_info(const fmca_h2_actor *actor, double out[13]) {
  if (!actor || !out) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  out[0] = actor->dim;
  out[1] = (double)actor->m;
  out[2] = (double)actor->n;
  out[3] = actor->ip.q;
  out[4] = actor->rows().ncl;
  out[5] = actor->cols().ncl;
  out[6] = (double)actor->pat.n_far;
  out[7] = (double)actor->pat.n_near;
  out[8] = actor->pat.near_entries;
  out[9] = actor->pat.far_entries;
  out[10] = actor->rows().depth;
  out[11] = actor->cols().depth;
  out[12] = actor->shared ? 1 : 0;
  return FMCA_OK;
}

}  // extern "C"

namespace {

// Position of element (i, j) of a stored R x K block in its matrix-core tile
// layout (the inverse of the device's tile_pos; layout in device.hpp).
template <typename T>
int64_t tile_index(int i, int j, int K) {
  constexpr int G = kStoredGroup<T>;
  const int ng = K / (4 * G);
  const int lane = (i & 15) + 16 * (j & 3);
  const int ks = j >> 2;
  const int64_t sg = (int64_t)(i >> 4) * ng + ks / G;
  return (sg * 64 + lane) * G + ks % G;
}

// Writes the projected blocks over the assembled ones of precision T.
template <typename T>
fmca_status store_projected(fmca_h2_actor *a, const ProjectedMatrix &pm) {
  const int q = a->ip.q, Rq = a->st_Rq, Kq = a->st_Kq;
  const size_t nfar = a->pat.far_src.size();
  std::vector<T> far((size_t)nfar * Rq * Kq, T(0));
  parallel_ranges((int64_t)nfar, [&](int64_t lo, int64_t hi) {
    for (int64_t b = lo; b < hi; ++b) {
      T *dst = far.data() + (size_t)b * Rq * Kq;
      const double *S = pm.far.data() + (size_t)b * q * q;
      for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) dst[tile_index<T>(i, j, Kq)] = (T)S[(size_t)i * q + j];
    }
  });
  if (nfar)
    HIP_TRY(hipMemcpy(a->st_far.p, far.data(), far.size() * sizeof(T), hipMemcpyHostToDevice));
  // near blocks: patch them into a host copy of the stored near region (which
  // also holds the leaf bases V_t^T)
  std::vector<T> near(a->st_near.bytes / sizeof(T));
  if (!near.empty())
    HIP_TRY(hipMemcpy(near.data(), a->st_near.p, near.size() * sizeof(T), hipMemcpyDeviceToHost));
  const Tree &tr = a->rows(), &tc = a->cols();
  const int nl = (int)tr.leaves.size();
  parallel_ranges(nl, [&](int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
      const int t = tr.leaves[k];
      const int nt = (int)tr.size[t];
      for (int e = a->pat.near_ptr[k]; e < a->pat.near_ptr[k + 1]; ++e) {
        const int ns = (int)tc.size[a->pat.near_src[e]];
        const int Kt = stored_kpad<T>(ns);
        T *dst = near.data() + a->st_near_off_host[e];
        const double *src = pm.near.data() + pm.near_off[e];
        for (int j = 0; j < ns; ++j)
          for (int i = 0; i < nt; ++i) dst[tile_index<T>(i, j, Kt)] = (T)src[(size_t)j * nt + i];
      }
    }
  });
  if (!near.empty())
    HIP_TRY(hipMemcpy(a->st_near.p, near.data(), near.size() * sizeof(T), hipMemcpyHostToDevice));
  return FMCA_OK;
}

}  // namespace

extern "C" {

fmca_status fmca_h2_project(fmca_h2_actor *a, const double *M, int64_t ld, int precision,
                            double dual_check[2]) {
  if (!a || !M) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  if (ld < a->m) return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (precision != FMCA_PRECISION_FP64 && precision != FMCA_PRECISION_FP32)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
  if (a->ip.scheme != IS_TENSOR)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the projector takes the tensor-product interpolator");
  if (a->ip.q > kStoredMaxRows / 2)
    return fail(FMCA_ERR_NOT_SUPPORTED, "projected blocks need q <= 128");
  if (a->geodesic && precision != FMCA_PRECISION_FP64)
    return fail(FMCA_ERR_NOT_SUPPORTED, "the geodesic distance is evaluated in FP64 only");
  std::lock_guard<std::mutex> lock(a->mtx);
  ProjectedMatrix pm;
  std::string err;
  if (!project_matrix(a->rows(), a->cols(), a->shared, a->ip, a->trans_rows,
                      a->shared ? a->trans_rows : a->trans_cols, a->pts_rows,
                      a->shared ? a->pts_rows : a->pts_cols, a->pat, M, ld, pm, &err))
    return fail(FMCA_ERR_NOT_SUPPORTED, err);
  if (dual_check) {
    dual_check[0] = pm.dual_check_rows;
    dual_check[1] = pm.dual_check_cols;
  }
  fmca_status st = ensure_device(a);
  if (st != FMCA_OK) return st;
  // the stored-block layout and the leaf bases, then the projected blocks over
  // the generated ones
  a->release_stored();
  st = precision == FMCA_PRECISION_FP64
           ? assemble<double>(a, a->d_rows_64, a->shared ? a->d_rows_64 : a->d_cols_64)
           : assemble<float>(a, a->d_rows_32, a->shared ? a->d_rows_32 : a->d_cols_32);
  if (st == FMCA_OK)
    st = precision == FMCA_PRECISION_FP64 ? store_projected<double>(a, pm)
                                          : store_projected<float>(a, pm);
  if (st != FMCA_OK) {
    a->release_stored();
    return st;
  }
  a->st_prec = precision;
  return FMCA_OK;
}

fmca_status fmca_h2_interpolation(const fmca_h2_actor *actor, int *interpolator, int *q,
                                  int32_t *multi_indices, double *nodes, double *invV) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  const Interpolation &ip = actor->ip;
  const int d = ip.dim, nq = ip.q;
  const bool td = ip.scheme == IS_TOTAL_DEGREE;
  if (interpolator) *interpolator = td ? FMCA_INTERP_TOTAL_DEGREE : FMCA_INTERP_TENSOR_PRODUCT;
  if (q) *q = nq;
  for (int k = 0; k < nq; ++k) {
    // FMCA position k -> packed multi-index
    int pk = 0;
    if (td) {
      pk = ip.midx[k];
    } else {
      int lin = ip.fmca_order[k];
      for (int i = 0; i < d; ++i) {
        pk |= (lin % ip.P1) << (8 * i);
        lin /= ip.P1;
      }
    }
    for (int i = 0; i < d; ++i) {
      const int ai = (pk >> (8 * i)) & 0xff;
      if (multi_indices) multi_indices[(size_t)d * k + i] = ai;
      if (nodes) nodes[(size_t)d * k + i] = td ? ip.nodes[(size_t)d * k + i] : ip.xi[ai];
    }
  }
  if (invV)
    for (int j = 0; j < nq; ++j)
      for (int i = 0; i < nq; ++i)
        invV[i + (size_t)nq * j] = td ? ip.invV[i + (size_t)nq * j] : (i == j ? 1. : 0.);
  return FMCA_OK;
}

fmca_status fmca_h2_clusters(const fmca_h2_actor *actor, int which, int64_t *begin, int64_t *size,
                             int32_t *level, int32_t *sons, double *bb) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  const Tree &T = which == FMCA_ROWS ? actor->rows() : actor->cols();
  const int d = T.dim;
  for (int c = 0; c < T.ncl; ++c) {
    if (begin) begin[c] = T.begin[c];
    if (size) size[c] = T.size[c];
    if (level) level[c] = T.level[c];
    if (sons) {
      sons[2 * c] = T.son0[c];
      sons[2 * c + 1] = T.son1[c];
    }
    if (bb)
      for (int i = 0; i < d; ++i) {
        bb[(size_t)3 * d * c + i] = T.bbmin[(size_t)c * d + i];
        bb[(size_t)3 * d * c + d + i] = T.bbmax[(size_t)c * d + i];
        bb[(size_t)3 * d * c + 2 * d + i] = T.bbsize[(size_t)c * d + i];
      }
  }
  return FMCA_OK;
}

fmca_status fmca_h2_launch_order(const fmca_h2_actor *actor, int which, int32_t *out,
                                 int64_t *len) {
  if (!actor || !len) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  if (which != 0 && which != 1) return fail(FMCA_ERR_INVALID_ARGUMENT, "which must be 0 or 1");
  const std::vector<int32_t> &o = which == 0 ? actor->pat.far_order : actor->pat.near_order;
  const std::vector<int32_t> &tk = which == 0 ? actor->pat.far_task : actor->pat.near_task;
  *len = (int64_t)o.size();
  if (out)
    for (size_t b = 0; b < o.size(); ++b)
      for (int k = 0; k < 4; ++k) out[4 * b + k] = o[b] < 0 ? -1 : tk[4 * (size_t)o[b] + k];
  return FMCA_OK;
}

fmca_status fmca_h2_blocks(const fmca_h2_actor *actor, int32_t *row, int32_t *col,
                           int8_t *low_rank) {
  if (!actor || !row || !col || !low_rank) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  const Pattern &p = actor->pat;
  std::copy(p.block_row.begin(), p.block_row.end(), row);
  std::copy(p.block_col.begin(), p.block_col.end(), col);
  std::copy(p.block_lr.begin(), p.block_lr.end(), low_rank);
  return FMCA_OK;
}

}  // extern "C"

namespace {

constexpr int kBasisChunk = 8;  // slot pairs per basis allocation

// Share of tol that the lower-precision GCR recurrence aims at before its
// FP64 check, so that the check meets tol despite the drift of the
// recurrence from the true residual (one FP64 action saved).  CG, with its
// replacements every 0.1 of residual drop, gains nothing from it (measured:
// 10 more iterations, the same FP64 actions).
constexpr double kInnerTarget = 0.5;

// GCR with FP64 residual replacement, the method FMCA_SOLVER_GCR of
// fmca_h2_cg_device (no FMCA counterpart).  GCR (Eisenstat, Elman & Schultz
// 1983): each iteration multiplies the residual, q0 = (K + ridge I) r,
// orthogonalises q0 against the stored images (one classical Gram-Schmidt
// pass) and applies the same combination to p0 = r, so a stored pair keeps
// q_j = (K + ridge I) p_j as that matvec computed it, whatever its rounding;
// the residual then drops by its projection on the new image, which
// minimises ||r|| over the window.  Lower-precision matvecs change the
// operator from one iteration to the next by their rounding; the pairs stay
// consistent, so the iteration count follows the FP64 one.  Replacement as in
// CG (every active column dropped by refine_factor, or one reached tol in
// FP64); a column whose FP64 residual meets tol stops, the others have the
// replaced residual projected on the window at once.  Pairs are stored as S
// (double for FP64 matvecs, float otherwise) in a circular window of at most
// m pairs; the oldest is dropped when it is full.
template <typename S>
fmca_status gcr_solve(fmca_h2_actor *a, const double *d_b, int64_t ld_b, int nr, double *d_x,
                      int64_t ld_x, const fmca_cg_options *opts, fmca_cg_result *res,
                      hipStream_t sm, std::chrono::steady_clock::time_point t0) {
  const int64_t n = a->n;
  const double ridge = opts->ridge;
  const bool cl = a->shared;
  const uint32_t *perm = a->d_rows_i.perm.as<uint32_t>();
  const size_t vec = (size_t)n * nr;
  const int prec = opts->precision;
  const int max_iter = opts->max_iter > 0 ? opts->max_iter : 100000;
  const int max_updates = opts->max_outer > 0 ? opts->max_outer : 1000;
  const double delta = (opts->refine_factor > 0 && opts->refine_factor < 1) ? opts->refine_factor
                                                                              : 0.1;
  // window size
  const size_t pair_bytes = 2 * vec * sizeof(S);
  int m = opts->max_basis;
  if (m < 0 || m > kBasisMax)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "max_basis must be in [0, 256]");
  if (m == 0) {
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    size_t budget = free_b / 2;
    const int64_t cap = g_dev_limit.load();
    if (cap > 0) {
      const int64_t left = cap - g_dev_live.load() - (int64_t)(4 * vec * sizeof(double));
      budget = std::min<size_t>(budget, left > 0 ? (size_t)left / 2 : 0);
    }
    m = (int)std::min<size_t>((size_t)kBasisMax, budget / pair_bytes);
    m = std::min(m, max_iter + 1);
  }
  if (m < 2) return fail(FMCA_ERR_OUT_OF_MEMORY, "GCR needs room for two basis pairs");
  HIP_TRY(a->cg_ws.alloc(4 * vec * sizeof(double)));
  // scalars: rr, qqrq, coef, qq table (doubles), then partial sums, then active flags
  const size_t nd = (size_t)nr * (1 + 2 + 2 * (size_t)kBasisMax);
  const size_t scratch_d =
      std::max(basis_scratch(nr, kBasisMax), blas1_dot_scratch(nr)) / sizeof(double);
  HIP_TRY(a->cg_scal.alloc((nd + scratch_d) * sizeof(double) + (size_t)nr * sizeof(int)));
  double *w = a->cg_ws.as<double>();
  double *B = w, *X = w + vec, *R = w + 2 * vec, *Y = w + 3 * vec;
  double *sc = a->cg_scal.as<double>();
  double *d_rr = sc, *d_qqrq = sc + nr, *d_coef = sc + 3 * nr,
         *d_qq = sc + 3 * nr + (size_t)nr * kBasisMax, *d_scratch = sc + nd;
  int *d_active = reinterpret_cast<int *>(sc + nd + scratch_d);
  auto chk = [&](hipError_t e, const char *what) { return from_hip(e, what); };
#define GCR_TRY(expr)                                 \
  do {                                                \
    fmca_status _s = chk((expr), #expr);              \
    if (_s != FMCA_OK) return _s;                     \
  } while (0)
  auto matvec = [&](const double *in, double *out, int p) -> fmca_status {
    if (p == FMCA_PRECISION_FP64) res->matvecs_fp64++; else res->matvecs_fp32++;
    return enqueue_any(a, in, n, nr, out, n, p, !cl, false, sm);
  };
  std::vector<double> rn2(nr), bn2(nr), target(nr), rmax(nr);
  std::vector<int> active(nr);
  auto read_rr = [&]() -> fmca_status {
    GCR_TRY(hipMemcpyAsync(rn2.data(), d_rr, nr * sizeof(double), hipMemcpyDeviceToHost, sm));
    GCR_TRY(hipStreamSynchronize(sm));
    return FMCA_OK;
  };
  auto push_active = [&]() -> hipError_t {
    return hipMemcpyAsync(d_active, active.data(), nr * sizeof(int), hipMemcpyHostToDevice, sm);
  };
  // basis slots: chunks allocated as the window grows, released at the end
  std::vector<const S *> qtab(m, nullptr), ptab(m, nullptr);
  a->basis.clear();
  HIP_TRY(a->basis_tab.alloc(2 * (size_t)m * sizeof(void *)));
  const S *const *d_qtab = reinterpret_cast<const S *const *>(a->basis_tab.p);
  const S *const *d_ptab = d_qtab + m;
  int allocated = 0;
  auto ensure_slot = [&](int slot) -> fmca_status {
    if (slot < allocated) return FMCA_OK;
    auto buf = std::make_unique<DevBuf>();
    const int k = std::min(kBasisChunk, m - allocated);
    hipError_t e = buf->alloc((size_t)k * pair_bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(FMCA_ERR_OUT_OF_MEMORY, "GCR basis allocation failed (lower max_basis)");
    }
    S *base = buf->as<S>();
    for (int j = 0; j < k; ++j) {
      ptab[allocated + j] = base + (size_t)j * vec;
      qtab[allocated + j] = base + (size_t)(k + j) * vec;
    }
    a->basis.push_back(std::move(buf));
    allocated += k;
    GCR_TRY(hipMemcpyAsync(a->basis_tab.p, qtab.data(), m * sizeof(void *),
                           hipMemcpyHostToDevice, sm));
    GCR_TRY(hipMemcpyAsync(reinterpret_cast<char *>(a->basis_tab.p) + m * sizeof(void *),
                           ptab.data(), m * sizeof(void *), hipMemcpyHostToDevice, sm));
    // the host tables change when the next chunk arrives
    GCR_TRY(hipStreamSynchronize(sm));
    return FMCA_OK;
  };
  fmca_status st;
  if (cl) {
    GCR_TRY(blas1_permute(n, nr, perm, true, d_b, ld_b, B, n, sm));
    GCR_TRY(blas1_permute(n, nr, perm, true, d_x, ld_x, X, n, sm));
  } else {
    GCR_TRY(hipMemcpy2DAsync(B, n * sizeof(double), d_b, ld_b * sizeof(double),
                             n * sizeof(double), nr, hipMemcpyDeviceToDevice, sm));
    GCR_TRY(hipMemcpy2DAsync(X, n * sizeof(double), d_x, ld_x * sizeof(double),
                             n * sizeof(double), nr, hipMemcpyDeviceToDevice, sm));
  }
  GCR_TRY(blas1_dot(n, nr, B, n, B, n, d_rr, d_scratch, sm));
  if ((st = read_rr()) != FMCA_OK) return st;
  bn2 = rn2;
  for (int j = 0; j < nr; ++j)
    if (bn2[j] == 0) bn2[j] = 1;  // zero rhs: absolute residual
  // r = b - (K + ridge I) x, FP64; exactly b when x = 0
  GCR_TRY(blas1_dot(n, nr, X, n, X, n, d_rr, d_scratch, sm));
  if ((st = read_rr()) != FMCA_OK) return st;
  if (std::all_of(rn2.begin(), rn2.end(), [](double v) { return v == 0; })) {
    GCR_TRY(hipMemcpyAsync(R, B, vec * sizeof(double), hipMemcpyDeviceToDevice, sm));
  } else {
    if ((st = matvec(X, Y, FMCA_PRECISION_FP64)) != FMCA_OK) return st;
    GCR_TRY(blas1_residual(n, nr, ridge, B, n, Y, n, X, n, R, n, sm));
  }
  GCR_TRY(blas1_dot(n, nr, R, n, R, n, d_rr, d_scratch, sm));
  if ((st = read_rr()) != FMCA_OK) return st;
  bool any = false;
  double worst = 0;
  for (int j = 0; j < nr; ++j) {
    rmax[j] = std::sqrt(rn2[j]);
    target[j] = opts->tol * std::sqrt(bn2[j]);
    active[j] = rmax[j] > target[j];
    any = any || active[j];
    worst = std::max(worst, rmax[j] / std::sqrt(bn2[j]));
  }
  res->rel_residual = worst;
  GCR_TRY(push_active());
  int first = 0, count = 0;
  while (any && res->iterations < max_iter && res->outer < max_updates) {
    // q0 = (K + ridge I) r
    if ((st = matvec(R, Y, prec)) != FMCA_OK) return st;
    GCR_TRY(blas1_add_ridge(n, nr, ridge, R, n, Y, n, sm));
    if (count == m) {  // full: drop the oldest pair, whose slot takes the new one
      first = (first + 1) % m;
      --count;
    }
    const int slot = (first + count) % m;
    if ((st = ensure_slot(slot)) != FMCA_OK) return st;
    GCR_TRY(basis_dots<S>(n, nr, Y, d_qtab, first, count, m, d_qq, d_active, d_coef, d_scratch,
                          sm));
    S *qk = const_cast<S *>(qtab[slot]), *pk = const_cast<S *>(ptab[slot]);
    GCR_TRY(basis_new_pair<S>(n, nr, Y, R, d_coef, d_qtab, d_ptab, first, count, m, qk, pk,
                              d_active, d_qqrq, d_scratch, sm));
    GCR_TRY(basis_update<S>(n, nr, pk, qk, d_qqrq, d_qq + (size_t)slot * nr, X, R, d_active,
                            d_rr, d_scratch, sm));
    ++count;
    if ((st = read_rr()) != FMCA_OK) return st;
    res->iterations++;
    bool any_done = false, all_ready = true;
    for (int j = 0; j < nr; ++j) {
      if (!active[j]) continue;
      const double rn = std::sqrt(rn2[j]);
      if (!std::isfinite(rn)) return fail(FMCA_ERR_NONFINITE, "GCR residual is not finite");
      // lower precision: the iterated residual runs to kInnerTarget tol
      const bool done = rn <= (prec == FMCA_PRECISION_FP64 ? 1.0 : kInnerTarget) * target[j];
      any_done = any_done || done;
      all_ready = all_ready && (done || rn < delta * rmax[j]);
    }
    const bool update = prec == FMCA_PRECISION_FP64 ? any_done : all_ready;
    if (!update) continue;
    // FP64 residual; columns that meet tol stop, the others are projected
    if ((st = matvec(X, Y, FMCA_PRECISION_FP64)) != FMCA_OK) return st;
    GCR_TRY(blas1_residual(n, nr, ridge, B, n, Y, n, X, n, R, n, sm));
    GCR_TRY(blas1_dot(n, nr, R, n, R, n, d_rr, d_scratch, sm));
    if ((st = read_rr()) != FMCA_OK) return st;
    res->outer++;
    worst = 0;
    any = false;
    for (int j = 0; j < nr; ++j) {
      const double rn = std::sqrt(rn2[j]);
      worst = std::max(worst, rn / std::sqrt(bn2[j]));
      if (active[j] && rn <= target[j]) active[j] = 0;
      any = any || active[j];
    }
    res->rel_residual = worst;
    GCR_TRY(push_active());
    if (!any) break;
    GCR_TRY(basis_dots<S>(n, nr, R, d_qtab, first, count, m, d_qq, d_active, d_coef, d_scratch,
                          sm));
    GCR_TRY(basis_project<S>(n, nr, d_coef, d_qtab, d_ptab, first, count, m, X, R, d_active,
                             d_rr, d_scratch, sm));
    if ((st = read_rr()) != FMCA_OK) return st;
    for (int j = 0; j < nr; ++j)
      if (active[j]) rmax[j] = std::sqrt(rn2[j]);
  }
  res->converged = !any;
  if (any) {  // stopped by a cap: report the FP64 residual of the returned x
    if ((st = matvec(X, Y, FMCA_PRECISION_FP64)) != FMCA_OK) return st;
    GCR_TRY(blas1_residual(n, nr, ridge, B, n, Y, n, X, n, R, n, sm));
    GCR_TRY(blas1_dot(n, nr, R, n, R, n, d_rr, d_scratch, sm));
    if ((st = read_rr()) != FMCA_OK) return st;
    worst = 0;
    for (int j = 0; j < nr; ++j) worst = std::max(worst, std::sqrt(rn2[j] / bn2[j]));
    res->rel_residual = worst;
  }
  if (cl)
    GCR_TRY(blas1_permute(n, nr, perm, false, X, n, d_x, ld_x, sm));
  else
    GCR_TRY(hipMemcpy2DAsync(d_x, ld_x * sizeof(double), X, n * sizeof(double),
                             n * sizeof(double), nr, hipMemcpyDeviceToDevice, sm));
  GCR_TRY(hipStreamSynchronize(sm));
#undef GCR_TRY
  a->basis.clear();
  a->basis_tab.release();
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res->converged ? FMCA_OK : fail(FMCA_ERR_NOT_CONVERGED, "GCR did not reach tol");
}

}  // namespace

extern "C" {

fmca_status fmca_h2_cg_device(fmca_h2_actor *a, const double *d_b, int64_t ld_b, int64_t nrhs,
                              double *d_x, int64_t ld_x, const fmca_cg_options *opts,
                              fmca_cg_result *res, void *stream_ptr) {
  if (!a || !opts || !res || nrhs < 1 || !d_b || !d_x)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer or nrhs < 1");
  if (a->m != a->n) return fail(FMCA_ERR_INVALID_ARGUMENT, "CG needs a square kernel matrix");
  if (ld_b < a->n || ld_x < a->n) return fail(FMCA_ERR_INVALID_ARGUMENT, "leading dimension");
  if (nrhs > 65535) return fail(FMCA_ERR_NOT_SUPPORTED, "more than 65535 right-hand sides");
  if (opts->precision != FMCA_PRECISION_FP64 && opts->precision != FMCA_PRECISION_FP32 &&
      opts->precision != FMCA_PRECISION_FP32_BF16)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
  if (!(opts->tol > 0) || !(opts->ridge >= 0))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "tol must be > 0 and ridge >= 0");
  std::lock_guard<std::mutex> lock(a->mtx);
  fmca_status st = ensure_device(a);
  if (st != FMCA_OK) return st;
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t sm = (hipStream_t)stream_ptr;
  if (opts->method != FMCA_SOLVER_CG && opts->method != FMCA_SOLVER_GCR)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown solver method");
  if (opts->method == FMCA_SOLVER_GCR) {
    *res = fmca_cg_result{};
    return opts->precision == FMCA_PRECISION_FP64
               ? gcr_solve<double>(a, d_b, ld_b, (int)nrhs, d_x, ld_x, opts, res, sm, t0)
               : gcr_solve<float>(a, d_b, ld_b, (int)nrhs, d_x, ld_x, opts, res, sm, t0);
  }
  const int64_t n = a->n;
  const int nr = (int)nrhs;
  const double ridge = opts->ridge;
  // Cluster order avoids a permutation per matvec; it needs one tree for both sides.
  const bool cl = a->shared;
  const uint32_t *perm = a->d_rows_i.perm.as<uint32_t>();
  const size_t vec = (size_t)n * nr;
  HIP_TRY(a->cg_ws.alloc(8 * vec * sizeof(double)));
  HIP_TRY(a->cg_scal.alloc((8 * (size_t)nr) * sizeof(double) + blas1_dot_scratch(nr)));
  double *w = a->cg_ws.as<double>();
  double *B = w, *X = w + vec, *R = w + 2 * vec, *Dv = w + 3 * vec, *RR = w + 4 * vec,
         *Pv = w + 5 * vec, *Q = w + 6 * vec, *Y = w + 7 * vec;
  double *sc = a->cg_scal.as<double>();
  double *d_dot = sc, *d_alpha = sc + nr, *d_beta = sc + 2 * nr, *d_one = sc + 3 * nr;
  double *d_scratch = sc + 8 * nr;
  std::vector<double> h(4 * (size_t)nr);
  double *h_dot = h.data(), *h_alpha = h.data() + nr, *h_beta = h.data() + 2 * nr,
         *h_one = h.data() + 3 * nr;
  for (int j = 0; j < nr; ++j) h_one[j] = 1.0;
  auto chk = [&](hipError_t e, const char *what) { return from_hip(e, what); };
#define CG_TRY(expr)                                  \
  do {                                                \
    fmca_status _s = chk((expr), #expr);              \
    if (_s != FMCA_OK) return _s;                     \
  } while (0)
  auto dot = [&](const double *u, const double *v, double *hout) -> fmca_status {
    CG_TRY(blas1_dot(n, nr, u, n, v, n, d_dot, d_scratch, sm));
    CG_TRY(hipMemcpyAsync(hout, d_dot, nr * sizeof(double), hipMemcpyDeviceToHost, sm));
    CG_TRY(hipStreamSynchronize(sm));
    return FMCA_OK;
  };
  auto upload = [&](double *dst, const double *src) -> hipError_t {
    return hipMemcpyAsync(dst, src, nr * sizeof(double), hipMemcpyHostToDevice, sm);
  };
  auto matvec = [&](const double *in, double *out, int prec) -> fmca_status {
    if (prec == FMCA_PRECISION_FP64) res->matvecs_fp64++; else res->matvecs_fp32++;
    return enqueue_any(a, in, n, nr, out, n, prec, !cl, false, sm);
  };
  *res = fmca_cg_result{};
  CG_TRY(upload(d_one, h_one));
  if (cl) {
    CG_TRY(blas1_permute(n, nr, perm, true, d_b, ld_b, B, n, sm));
    CG_TRY(blas1_permute(n, nr, perm, true, d_x, ld_x, X, n, sm));
  } else {
    CG_TRY(hipMemcpy2DAsync(B, n * sizeof(double), d_b, ld_b * sizeof(double), n * sizeof(double),
                            nr, hipMemcpyDeviceToDevice, sm));
    CG_TRY(hipMemcpy2DAsync(X, n * sizeof(double), d_x, ld_x * sizeof(double), n * sizeof(double),
                            nr, hipMemcpyDeviceToDevice, sm));
  }
  std::vector<double> bn2(nr), rn2(nr), rho(nr), target(nr);
  (void)Dv;
  (void)RR;
  if ((st = dot(B, B, bn2.data())) != FMCA_OK) return st;
  for (int j = 0; j < nr; ++j)
    if (bn2[j] == 0) bn2[j] = 1;  // zero rhs: absolute residual
  // r = b - (K + ridge I) x, FP64; exactly b when x = 0
  if ((st = dot(X, X, rn2.data())) != FMCA_OK) return st;
  if (std::all_of(rn2.begin(), rn2.end(), [](double v) { return v == 0; })) {
    CG_TRY(hipMemcpyAsync(R, B, vec * sizeof(double), hipMemcpyDeviceToDevice, sm));
  } else {
    if ((st = matvec(X, Y, FMCA_PRECISION_FP64)) != FMCA_OK) return st;
    CG_TRY(blas1_residual(n, nr, ridge, B, n, Y, n, X, n, R, n, sm));
  }
  // CG with reliable updates (Sleijpen & van der Vorst 1996; the form of
  // Clark et al. 2010): the recurrence runs with matvecs of the chosen
  // precision (FP64 vectors, inner products and updates throughout).  Once
  // every active column's iterated residual has dropped by refine_factor
  // since the last replacement or reached its target, the true residual
  // b - (K + ridge I) x is recomputed with the FP64 action and replaces it;
  // the search direction is kept and re-orthogonalised against the new
  // residual, p -= (r.p / r.r) r.  With FP64 matvecs only the final check
  // replaces the residual.
  const int prec = opts->precision;
  const int max_iter = opts->max_iter > 0 ? opts->max_iter : 100000;
  const int max_updates = opts->max_outer > 0 ? opts->max_outer : 1000;
  const double delta = (opts->refine_factor > 0 && opts->refine_factor < 1) ? opts->refine_factor
                                                                              : 0.1;
  if ((st = dot(R, R, rn2.data())) != FMCA_OK) return st;
  std::vector<char> active(nr);
  std::vector<double> rmax(nr);
  bool any = false;
  double worst = 0;
  for (int j = 0; j < nr; ++j) {
    rho[j] = rn2[j];
    rmax[j] = std::sqrt(rn2[j]);
    target[j] = opts->tol * std::sqrt(bn2[j]);
    active[j] = rmax[j] > target[j];
    any = any || active[j];
    worst = std::max(worst, rmax[j] / std::sqrt(bn2[j]));
  }
  res->rel_residual = worst;
  CG_TRY(hipMemcpyAsync(Pv, R, vec * sizeof(double), hipMemcpyDeviceToDevice, sm));
  while (any && res->iterations < max_iter && res->outer < max_updates) {
    if ((st = matvec(Pv, Q, prec)) != FMCA_OK) return st;
    CG_TRY(blas1_add_ridge(n, nr, ridge, Pv, n, Q, n, sm));
    if ((st = dot(Pv, Q, h_dot)) != FMCA_OK) return st;
    for (int j = 0; j < nr; ++j) {
      if (active[j] && !(h_dot[j] > 0)) {
        // negative or zero curvature: CG is undefined on an indefinite system
        res->indefinite = 1;
        res->converged = 0;
        res->seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (cl)
          CG_TRY(blas1_permute(n, nr, perm, false, X, n, d_x, ld_x, sm));
        else
          CG_TRY(hipMemcpy2DAsync(d_x, ld_x * sizeof(double), X, n * sizeof(double),
                                  n * sizeof(double), nr, hipMemcpyDeviceToDevice, sm));
        CG_TRY(hipStreamSynchronize(sm));
        return fail(FMCA_ERR_NOT_CONVERGED,
                    "CG met p.(K + ridge I)p <= 0: the H2 matrix plus ridge is indefinite");
      }
      h_alpha[j] = active[j] ? rho[j] / h_dot[j] : 0.0;
    }
    CG_TRY(upload(d_alpha, h_alpha));
    CG_TRY(blas1_axpy(n, nr, 1.0, d_alpha, Pv, n, X, n, sm));
    CG_TRY(blas1_axpy(n, nr, -1.0, d_alpha, Q, n, R, n, sm));
    if ((st = dot(R, R, rn2.data())) != FMCA_OK) return st;
    res->iterations++;
    // FP64: replace as soon as one column reaches its target.  Lower
    // precision: once all active columns have reached theirs or dropped by delta.
    bool any_done = false, all_ready = true;
    for (int j = 0; j < nr; ++j) {
      if (!active[j]) continue;
      const double rn = std::sqrt(rn2[j]);
      if (!std::isfinite(rn)) return fail(FMCA_ERR_NONFINITE, "CG residual is not finite");
      const bool done = rn <= target[j];
      any_done = any_done || done;
      all_ready = all_ready && (done || rn < delta * rmax[j]);
    }
    const bool update = prec == FMCA_PRECISION_FP64 ? any_done : all_ready;
    if (update) {
      if ((st = matvec(X, Y, FMCA_PRECISION_FP64)) != FMCA_OK) return st;
      CG_TRY(blas1_residual(n, nr, ridge, B, n, Y, n, X, n, R, n, sm));
      if ((st = dot(R, R, rn2.data())) != FMCA_OK) return st;
      res->outer++;
      worst = 0;
      for (int j = 0; j < nr; ++j) {
        const double rn = std::sqrt(rn2[j]);
        rmax[j] = rn;
        worst = std::max(worst, rn / std::sqrt(bn2[j]));
        if (active[j] && rn <= target[j]) active[j] = 0;
      }
      res->rel_residual = worst;
      if (prec != FMCA_PRECISION_FP64) {
        if ((st = dot(R, Pv, h_dot)) != FMCA_OK) return st;
        for (int j = 0; j < nr; ++j)
          h_alpha[j] = active[j] && rn2[j] != 0 ? h_dot[j] / rn2[j] : 0.0;
        CG_TRY(upload(d_alpha, h_alpha));
        CG_TRY(blas1_axpy(n, nr, -1.0, d_alpha, R, n, Pv, n, sm));
      }
    }
    any = false;
    for (int j = 0; j < nr; ++j) {
      h_beta[j] = active[j] && rho[j] != 0 ? rn2[j] / rho[j] : 0.0;
      if (active[j]) rho[j] = rn2[j];
      any = any || active[j];
    }
    if (!any) break;
    CG_TRY(upload(d_beta, h_beta));
    CG_TRY(blas1_xpby(n, nr, d_beta, R, n, Pv, n, sm));
  }
  res->converged = !any;
  if (any) {  // stopped by a cap: report the FP64 residual of the returned x
    if ((st = matvec(X, Y, FMCA_PRECISION_FP64)) != FMCA_OK) return st;
    CG_TRY(blas1_residual(n, nr, ridge, B, n, Y, n, X, n, R, n, sm));
    if ((st = dot(R, R, rn2.data())) != FMCA_OK) return st;
    worst = 0;
    for (int j = 0; j < nr; ++j) worst = std::max(worst, std::sqrt(rn2[j] / bn2[j]));
    res->rel_residual = worst;
  }
  if (cl)
    CG_TRY(blas1_permute(n, nr, perm, false, X, n, d_x, ld_x, sm));
  else
    CG_TRY(hipMemcpy2DAsync(d_x, ld_x * sizeof(double), X, n * sizeof(double), n * sizeof(double),
                            nr, hipMemcpyDeviceToDevice, sm));
  CG_TRY(hipStreamSynchronize(sm));
#undef CG_TRY
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res->converged ? FMCA_OK : fail(FMCA_ERR_NOT_CONVERGED, "CG did not reach tol");
}

fmca_status fmca_h2_set_profiling(fmca_h2_actor *actor, int enable) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  std::lock_guard<std::mutex> lock(actor->mtx);
  actor->profiling = enable != 0;
  return FMCA_OK;
}

fmca_status fmca_h2_phase_times_precision(fmca_h2_actor *actor, int precision, double ms[6],
                                          int64_t count[6], int reset) {
  if (!actor) return fail(FMCA_ERR_INVALID_ARGUMENT, "actor is null");
  if (precision < -1 || precision > FMCA_PRECISION_FP32_BF16)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown precision");
  std::lock_guard<std::mutex> lock(actor->mtx);
  for (auto &pp : actor->pending) {
    PhaseEvents &pe = pp.second;
    HIP_TRY(hipEventSynchronize(pe.ev[PH_COUNT]));
    for (int k = 0; k < PH_COUNT; ++k) {
      float t = 0;
      HIP_TRY(hipEventElapsedTime(&t, pe.ev[k], pe.ev[k + 1]));
      actor->phase_ms[pp.first][k] += t;
      actor->phase_count[pp.first][k] += 1;
    }
    for (auto &e : pe.ev) actor->ev_pool.push_back(e);
  }
  actor->pending.clear();
  for (int k = 0; k < PH_COUNT; ++k) {
    double t = 0;
    int64_t c = 0;
    for (int p = 0; p < 3; ++p) {
      if (precision != -1 && p != precision) continue;
      t += actor->phase_ms[p][k];
      c += actor->phase_count[p][k];
      if (reset) {
        actor->phase_ms[p][k] = 0;
        actor->phase_count[p][k] = 0;
      }
    }
    if (ms) ms[k] = t;
    if (count) count[k] = c;
  }
  return FMCA_OK;
}

fmca_status fmca_h2_phase_times(fmca_h2_actor *actor, double ms[6], int64_t count[6], int reset) {
  return fmca_h2_phase_times_precision(actor, -1, ms, count, reset);
}

fmca_status fmca_device_memory(int64_t *live_bytes, int64_t *peak_bytes, int reset_peak) {
  if (live_bytes) *live_bytes = g_dev_live.load();
  if (peak_bytes) *peak_bytes = g_dev_peak.load();
  if (reset_peak) g_dev_peak.store(g_dev_live.load());
  return FMCA_OK;
}

fmca_status fmca_set_device_memory_limit(int64_t limit_bytes) {
  if (limit_bytes < 0) return fail(FMCA_ERR_INVALID_ARGUMENT, "negative limit");
  g_dev_limit.store(limit_bytes);
  return FMCA_OK;
}

fmca_status fmca_diag_math(int fn, const double *d_x, double *d_y, int64_t n, void *stream) {
  if (n < 0 || (n && (!d_x || !d_y))) return fail(FMCA_ERR_INVALID_ARGUMENT, "null pointer");
  hipError_t e = launch_math_probe(fn, d_x, d_y, n, (hipStream_t)stream);
  if (e == hipErrorInvalidValue) return fail(FMCA_ERR_INVALID_ARGUMENT, "unknown function");
  return from_hip(e, "launch_math_probe");
}

}  // extern "C"
