// C ABI of the samplet tree, samplet transforms and samplet compression.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/fmca_amd.h"
#include "samplets.hpp"

using namespace fmca_amd;

struct fmca_samplet_tree {
  SampletTree S;
};

struct fmca_samplet_matrix {
  int64_t n = 0;
  HostArray<int64_t> rows, cols;
  HostArray<double> vals;
  double stats[FMCA_SAMPLET_STATS] = {0};
};

namespace {

fmca_status fail(fmca_status st, const std::string &msg) {
  set_last_error(msg);
  return st;
}

fmca_status from_hip(hipError_t e, const char *what) {
  const fmca_status st = e == hipErrorOutOfMemory ? FMCA_ERR_OUT_OF_MEMORY
                         : (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
                             ? FMCA_ERR_NO_DEVICE
                             : FMCA_ERR_DEVICE;
  return fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Device buffers of one compression, counted against the library's cap and
// released on every exit path.
struct DeviceArena {
  std::vector<std::pair<void *, size_t>> held;
  ~DeviceArena() {
    for (auto &h : held) device_free(h.first, h.second);
  }
  template <typename T>
  hipError_t upload(const std::vector<T> &v, const T **out) {
    void *p = nullptr;
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 256);
    const int e = device_malloc(&p, bytes);
    if (e != (int)hipSuccess) return (hipError_t)e;
    held.emplace_back(p, bytes);
    *out = reinterpret_cast<const T *>(p);
    if (v.empty()) return hipSuccess;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  hipError_t alloc(size_t bytes, void **out) {
    void *p = nullptr;
    bytes = std::max<size_t>(bytes, 256);
    const int e = device_malloc(&p, bytes);
    if (e != (int)hipSuccess) return (hipError_t)e;
    held.emplace_back(p, bytes);
    *out = p;
    return hipSuccess;
  }
};

#define HIP_TRY(expr)                                 \
  do {                                                \
    const hipError_t _e = (expr);                     \
    if (_e != hipSuccess) return from_hip(_e, #expr); \
  } while (0)

}  // namespace

extern "C" {

fmca_status fmca_samplet_tree_create(int dim, int64_t n, const double *P, int dtilde,
                                     fmca_samplet_tree **tree) {
  if (!tree) return fail(FMCA_ERR_INVALID_ARGUMENT, "tree is null");
  *tree = nullptr;
  if (dim < 1 || dim > 3) return fail(FMCA_ERR_NOT_SUPPORTED, "dimension must be 1, 2 or 3");
  if (n < 1 || !P) return fail(FMCA_ERR_INVALID_ARGUMENT, "empty point set");
  if (n > (int64_t)INT32_MAX) return fail(FMCA_ERR_NOT_SUPPORTED, "more than 2^31-1 points");
  for (int64_t k = 0; k < (int64_t)dim * n; ++k)
    if (!std::isfinite(P[k])) return fail(FMCA_ERR_NONFINITE, "non-finite point coordinate");
  if (dtilde > 5) return fail(FMCA_ERR_NOT_SUPPORTED, "dtilde must be at most 5");
  try {
    std::unique_ptr<fmca_samplet_tree> t(new fmca_samplet_tree());
    const std::string err = build_samplet_tree(t->S, dim, n, P, dtilde);
    if (!err.empty()) return fail(FMCA_ERR_INVALID_ARGUMENT, err);
    *tree = t.release();
  } catch (const std::bad_alloc &) {
    return fail(FMCA_ERR_OUT_OF_MEMORY, "host memory");
  }
  return FMCA_OK;
}

fmca_status fmca_samplet_tree_destroy(fmca_samplet_tree *tree) {
  delete tree;
  return FMCA_OK;
}

fmca_status fmca_samplet_tree_info(const fmca_samplet_tree *tree, int64_t *info) {
  if (!tree || !info) return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  const SampletTree &S = tree->S;
  info[0] = S.n;
  info[1] = S.dim;
  info[2] = S.dtilde;
  info[3] = S.tree.ncl;
  info[4] = S.ip.q;
  info[5] = S.m;
  info[6] = S.tree.depth;
  return FMCA_OK;
}

fmca_status fmca_samplet_tree_indices(const fmca_samplet_tree *tree, uint32_t *perm) {
  if (!tree || !perm) return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  std::memcpy(perm, tree->S.tree.perm.data(), sizeof(uint32_t) * tree->S.n);
  return FMCA_OK;
}

fmca_status fmca_samplet_tree_levels(const fmca_samplet_tree *tree, int32_t *levels) {
  if (!tree || !levels) return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  const std::vector<int32_t> l = samplet_levels(tree->S);
  std::memcpy(levels, l.data(), sizeof(int32_t) * l.size());
  return FMCA_OK;
}

fmca_status fmca_samplet_tree_clusters(const fmca_samplet_tree *tree, int64_t *begin,
                                       int64_t *size, int32_t *sons, int32_t *qdim,
                                       int32_t *nscal, int64_t *start) {
  if (!tree) return fail(FMCA_ERR_INVALID_ARGUMENT, "tree is null");
  const SampletTree &S = tree->S;
  const int ncl = S.tree.ncl;
  for (int c = 0; c < ncl; ++c) {
    if (begin) begin[c] = S.tree.begin[c];
    if (size) size[c] = S.tree.size[c];
    if (sons) {
      sons[2 * c] = S.tree.son0[c];
      sons[2 * c + 1] = S.tree.son1[c];
    }
    if (qdim) qdim[c] = S.qdim[c];
    if (nscal) nscal[c] = S.nscal[c];
    if (start) start[c] = S.start[c];
  }
  return FMCA_OK;
}

fmca_status fmca_samplet_tree_basis(const fmca_samplet_tree *tree, int cluster, double *Q,
                                    double *V) {
  if (!tree) return fail(FMCA_ERR_INVALID_ARGUMENT, "tree is null");
  const SampletTree &S = tree->S;
  if (cluster < 0 || cluster >= S.tree.ncl) return fail(FMCA_ERR_INVALID_ARGUMENT, "cluster");
  const int d = S.qdim[cluster];
  if (Q) std::memcpy(Q, S.Q.data() + S.q_off[cluster], sizeof(double) * d * d);
  if (V) std::memcpy(V, S.V.data() + S.v_off[cluster], sizeof(double) * S.ip.q * d);
  return FMCA_OK;
}

fmca_status fmca_samplet_transform(const fmca_samplet_tree *tree, int inverse, const double *in,
                                   int64_t ld_in, int64_t k, double *out, int64_t ld_out) {
  if (!tree || (!in && k > 0) || (!out && k > 0))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  const int64_t n = tree->S.n;
  if (k < 0 || ld_in < n || ld_out < n) return fail(FMCA_ERR_INVALID_ARGUMENT, "sizes");
  if (k == 0) return FMCA_OK;
  try {
    if (inverse)
      inverse_samplet_transform(tree->S, in, ld_in, k, out, ld_out);
    else
      samplet_transform(tree->S, in, ld_in, k, out, ld_out);
  } catch (const std::bad_alloc &) {
    return fail(FMCA_ERR_OUT_OF_MEMORY, "host memory");
  }
  return FMCA_OK;
}

fmca_status fmca_samplet_compress(const fmca_samplet_tree *tree, const fmca_kernel_desc *kernel,
                                  double eta, double threshold, fmca_samplet_matrix **out) {
  if (!tree || !kernel || !kernel->ktype || !out)
    return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const int kt = parse_kernel_type(kernel->ktype);
  if (kt < 0) return fail(FMCA_ERR_UNKNOWN_KERNEL, std::string("kernel ") + kernel->ktype);
  if (!(kernel->l > 0) || !std::isfinite(kernel->l))
    return fail(FMCA_ERR_INVALID_ARGUMENT, "lengthscale must be positive");
  if (!(eta >= 0) || !std::isfinite(eta)) return fail(FMCA_ERR_INVALID_ARGUMENT, "eta");
  if (!(threshold >= 0)) return fail(FMCA_ERR_INVALID_ARGUMENT, "threshold");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(FMCA_ERR_NO_DEVICE, "no HIP device");
  const SampletTree &S = tree->S;
  try {
    std::unique_ptr<fmca_samplet_matrix> M(new fmca_samplet_matrix());
    M->n = S.n;
    auto t0 = std::chrono::steady_clock::now();
    SampletProgram prog;
    plan_samplet_compression(prog, S, eta);
    M->stats[0] = (double)prog.pat_row.size();
    int64_t ndense = 0;
    for (const auto &j : prog.jobs) ndense += j.kind == 0;
    M->stats[1] = (double)ndense;
    M->stats[2] = (double)(prog.jobs.size() - ndense);
    M->stats[3] = (double)prog.job_elems;
    M->stats[4] = (double)prog.op.size();
    M->stats[5] = seconds_since(t0);

    // device inputs: points, cluster bases Q and W = invV^T V, boxes, nodes
    t0 = std::chrono::steady_clock::now();
    const Tree &T = S.tree;
    const int q = S.ip.q, ncl = T.ncl, dim = S.dim;
    std::vector<double> W(S.V.size());
    parallel_dynamic(ncl, 16, [&](int64_t c) {
      const double *Vc = S.V.data() + S.v_off[c];
      double *Wc = W.data() + S.v_off[c];
      for (int j = 0; j < S.qdim[c]; ++j)
        for (int a = 0; a < q; ++a) {
          double s = 0;
          for (int b = 0; b < q; ++b) s += S.ip.invV[b + (size_t)q * a] * Vc[b + (size_t)q * j];
          Wc[a + (size_t)q * j] = s;
        }
    });
    std::vector<double> unodes((size_t)dim * q);
    for (int k = 0; k < q; ++k)
      for (int i = 0; i < dim; ++i) unodes[(size_t)i * q + k] = S.ip.nodes[(size_t)dim * k + i];
    DeviceArena A;
    SampletDeviceData d;
    d.n = S.n;
    d.dim = dim;
    d.q = q;
    HIP_TRY(A.upload(S.pts, &d.pts));
    HIP_TRY(A.upload(S.Q, &d.Q));
    HIP_TRY(A.upload(W, &d.W));
    HIP_TRY(A.upload(S.q_off, &d.q_off));
    HIP_TRY(A.upload(S.v_off, &d.v_off));
    HIP_TRY(A.upload(T.begin, &d.begin));
    HIP_TRY(A.upload(T.size, &d.size));
    HIP_TRY(A.upload(S.qdim, &d.qdim));
    HIP_TRY(A.upload(T.bbmin, &d.bbmin));
    HIP_TRY(A.upload(T.bbsize, &d.bbsize));
    HIP_TRY(A.upload(unodes, &d.unodes));
    HIP_TRY(A.upload(prog.jobs, &d.jobs));
    std::vector<double>().swap(W);
    // the job blocks in batches of at most 2 GiB of output
    HostArray<double> blocks;  // every element is written by the copies below
    blocks.resize_uninit((size_t)prog.job_elems);
    const int64_t cap = std::min<int64_t>(prog.job_elems, (int64_t)1 << 28);
    void *dout = nullptr;
    HIP_TRY(A.alloc(sizeof(double) * std::max<int64_t>(cap, 1), &dout));
    hipStream_t stream;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // the blocks come back through a pinned staging buffer, copied on from
    // there by host threads (which also spreads the first touch of the pages
    // of `blocks`): several times the rate of a copy into pageable memory
    constexpr int64_t kStage = (int64_t)1 << 23;  // doubles (64 MiB)
    double *pinned = nullptr;
    if (hipHostMalloc((void **)&pinned, sizeof(double) * kStage, hipHostMallocDefault) !=
        hipSuccess) {
      (void)hipStreamDestroy(stream);
      return fail(FMCA_ERR_OUT_OF_MEMORY, "pinned staging buffer");
    }
    auto download = [&](double *dst, const double *src, int64_t elems) -> hipError_t {
      for (int64_t o = 0; o < elems; o += kStage) {
        const int64_t m = std::min(kStage, elems - o);
        const hipError_t e =
            hipMemcpy(pinned, src + o, sizeof(double) * m, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return e;
        constexpr int64_t kPiece = (int64_t)1 << 17;
        parallel_dynamic((m + kPiece - 1) / kPiece, 1, [&](int64_t p) {
          const int64_t a = p * kPiece, b = std::min(m, a + kPiece);
          std::memcpy(dst + o + a, pinned + a, sizeof(double) * (b - a));
        });
      }
      return hipSuccess;
    };
    M->stats[6] = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    SampletKernel K;
    K.kt = kt;
    K.l = kernel->l;
    K.c = kernel->c;
    K.nu = kernel->nu;
    double gpu_s = 0;
    hipError_t err = hipSuccess;
    const int64_t nj = (int64_t)prog.jobs.size();
    for (int64_t j0 = 0; j0 < nj && err == hipSuccess;) {
      int64_t j1 = j0;
      const int64_t base = prog.jobs[j0].off;
      while (j1 < nj) {
        const SampletJob &jb = prog.jobs[j1];
        const int64_t end = jb.off + (int64_t)S.qdim[jb.r] * S.qdim[jb.c];
        if (end - base > cap && j1 > j0) break;
        ++j1;
      }
      const SampletJob &last = prog.jobs[j1 - 1];
      const int64_t elems = last.off + (int64_t)S.qdim[last.r] * S.qdim[last.c] - base;
      if (elems > cap) {
        err = hipErrorOutOfMemory;  // a single block above the batch size
        break;
      }
      SampletDeviceData dj = d;
      dj.jobs = d.jobs + j0;
      dj.base = base;
      auto tg = std::chrono::steady_clock::now();
      err = (hipError_t)launch_samplet_jobs(dj, K, j1 - j0, reinterpret_cast<double *>(dout),
                                            stream);
      if (err == hipSuccess) err = hipStreamSynchronize(stream);
      gpu_s += seconds_since(tg);
      if (err == hipSuccess)
        err = download(blocks.data() + base, reinterpret_cast<const double *>(dout), elems);
      j0 = j1;
    }
    (void)hipHostFree(pinned);
    (void)hipStreamDestroy(stream);
    if (err != hipSuccess) return from_hip(err, "samplet kernel blocks");
    M->stats[7] = gpu_s;
    M->stats[8] = seconds_since(t0) - gpu_s;
    t0 = std::chrono::steady_clock::now();
    std::vector<double> comb;
    combine_samplet_blocks(prog, S, blocks.data(), comb);
    M->stats[9] = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    samplet_triplets(prog, S, blocks.data(), comb, threshold, M->rows, M->cols, M->vals);
    M->stats[10] = seconds_since(t0);
    M->stats[11] = (double)M->vals.size();
    *out = M.release();
  } catch (const std::bad_alloc &) {
    return fail(FMCA_ERR_OUT_OF_MEMORY, "host memory");
  }
  return FMCA_OK;
}

fmca_status fmca_samplet_matrix_destroy(fmca_samplet_matrix *mat) {
  delete mat;
  return FMCA_OK;
}

fmca_status fmca_samplet_matrix_nnz(const fmca_samplet_matrix *mat, int64_t *nnz) {
  if (!mat || !nnz) return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  *nnz = (int64_t)mat->vals.size();
  return FMCA_OK;
}

fmca_status fmca_samplet_matrix_triplets(const fmca_samplet_matrix *mat, int64_t *rows,
                                         int64_t *cols, double *vals) {
  if (!mat) return fail(FMCA_ERR_INVALID_ARGUMENT, "matrix is null");
  const size_t k = mat->vals.size();
  if (k == 0) return FMCA_OK;
  if (rows) std::memcpy(rows, mat->rows.data(), sizeof(int64_t) * k);
  if (cols) std::memcpy(cols, mat->cols.data(), sizeof(int64_t) * k);
  if (vals) std::memcpy(vals, mat->vals.data(), sizeof(double) * k);
  return FMCA_OK;
}

fmca_status fmca_samplet_matrix_aposteriori(const fmca_samplet_matrix *mat, double thres,
                                            int64_t *count, int64_t *index) {
  if (!mat || !count) return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  try {
    const std::vector<int64_t> keep = samplet_aposteriori_fast(
        mat->rows.data(), mat->cols.data(), mat->vals.data(), mat->vals.size(), thres);
    if (index && !keep.empty()) std::memcpy(index, keep.data(), sizeof(int64_t) * keep.size());
    *count = (int64_t)keep.size();
  } catch (const std::bad_alloc &) {
    return fail(FMCA_ERR_OUT_OF_MEMORY, "host memory");
  }
  return FMCA_OK;
}

fmca_status fmca_samplet_matrix_stats(const fmca_samplet_matrix *mat, double *stats) {
  if (!mat || !stats) return fail(FMCA_ERR_INVALID_ARGUMENT, "null argument");
  std::memcpy(stats, mat->stats, sizeof(mat->stats));
  return FMCA_OK;
}

}  // extern "C"
