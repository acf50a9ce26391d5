// pybind11 module over the C ABI (the Python face of py/FMCA.cpp:209-250,462-468).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>

#include "fmca_amd.h"

namespace py = pybind11;

namespace {

void check(fmca_status s) {
  if (s == FMCA_OK) return;
  const std::string msg = std::string(fmca_status_string(s)) + ": " + fmca_last_error();
  if (s == FMCA_ERR_INVALID_ARGUMENT || s == FMCA_ERR_UNKNOWN_KERNEL || s == FMCA_ERR_NONFINITE)
    throw py::value_error(msg);
  if (s == FMCA_ERR_NOT_SUPPORTED) throw py::value_error(msg);
  throw std::runtime_error(msg);
}

using farr = py::array_t<double, py::array::f_style | py::array::forcecast>;

// float64 only, as FMCA's noconvert() arguments (py/FMCA.cpp:467)
farr as_f64(const py::array &a, const char *what) {
  if (!a.dtype().is(py::dtype::of<double>()))
    throw py::type_error(std::string(what) + " must be a float64 array");
  return farr::ensure(a);
}

struct Actor {
  fmca_h2_actor *h = nullptr;
  int dim = 0;
  int64_t m = 0, n = 0;
  Actor() = default;
  Actor(const Actor &) = delete;
  ~Actor() {
    if (h) fmca_h2_destroy(h);
  }
};

std::unique_ptr<Actor> create(const std::string &ktype, double l, double c, double nu,
                              const py::array &Peval_in, const py::array &P_in, int deg,
                              double eta, int interpolator, int planner) {
  farr Peval = as_f64(Peval_in, "Peval"), P = as_f64(P_in, "P");
  if (Peval.ndim() != 2 || P.ndim() != 2 || Peval.shape(0) != P.shape(0))
    throw py::value_error("points must be (dim, n) arrays of equal dim");
  auto a = std::make_unique<Actor>();
  a->dim = (int)P.shape(0);
  a->m = Peval.shape(1);
  a->n = P.shape(1);
  fmca_kernel_desc k{ktype.c_str(), l, c, nu};
  {
    py::gil_scoped_release nogil;
    check(fmca_h2_create_ex(&k, a->dim, a->m, Peval.data(), a->n, P.data(), deg, eta,
                            interpolator, planner, &a->h));
  }
  return a;
}

py::array action(Actor &a, const py::array &rhs_in, int precision, bool cluster_order) {
  farr rhs = as_f64(rhs_in, "rhs");
  if (rhs.ndim() < 1 || rhs.ndim() > 2 || rhs.shape(0) != a.n)
    throw py::value_error("rhs must have n rows");
  const int64_t nrhs = rhs.ndim() == 2 ? rhs.shape(1) : 1;
  farr out({(py::ssize_t)a.m, (py::ssize_t)nrhs});
  double *po = out.mutable_data();
  const double *pr = rhs.data();
  {
    py::gil_scoped_release nogil;
    if (cluster_order)
      check(fmca_h2_action_cluster_order(a.h, pr, a.n, nrhs, po, a.m, precision));
    else
      check(fmca_h2_action(a.h, pr, a.n, nrhs, po, a.m, precision));
  }
  return out;
}

py::dict info(const Actor &a) {
  double v[13];
  check(fmca_h2_info(a.h, v));
  py::dict d;
  const char *keys[13] = {"dim",        "m",          "n",           "q",
                          "row_clusters", "col_clusters", "far_blocks", "near_blocks",
                          "near_entries", "far_entries",  "row_depth",  "col_depth",
                          "shared_tree"};
  for (int k = 0; k < 13; ++k) d[keys[k]] = v[k];
  return d;
}

py::dict clusters(const Actor &a, int which) {
  double v[13];
  check(fmca_h2_info(a.h, v));
  const py::ssize_t nc = (py::ssize_t)(which == FMCA_ROWS ? v[4] : v[5]);
  py::array_t<int64_t> begin(nc), size(nc);
  py::array_t<int32_t> level(nc), sons({nc, (py::ssize_t)2});
  py::array_t<double> bb({nc, (py::ssize_t)3, (py::ssize_t)a.dim});
  check(fmca_h2_clusters(a.h, which, begin.mutable_data(), size.mutable_data(),
                         level.mutable_data(), sons.mutable_data(), bb.mutable_data()));
  py::dict d;
  d["begin"] = begin;
  d["size"] = size;
  d["level"] = level;
  d["sons"] = sons;
  d["bb"] = bb;
  return d;
}

py::tuple blocks(const Actor &a) {
  double v[13];
  check(fmca_h2_info(a.h, v));
  const py::ssize_t nb = (py::ssize_t)(v[6] + v[7]);
  py::array_t<int32_t> row(nb), col(nb);
  py::array_t<int8_t> lr(nb);
  check(fmca_h2_blocks(a.h, row.mutable_data(), col.mutable_data(), lr.mutable_data()));
  return py::make_tuple(row, col, lr);
}

py::array_t<int32_t> launch_order(const Actor &a, int which) {
  int64_t len = 0;
  check(fmca_h2_launch_order(a.h, which, nullptr, &len));
  py::array_t<int32_t> out({(py::ssize_t)len, (py::ssize_t)4});
  check(fmca_h2_launch_order(a.h, which, out.mutable_data(), &len));
  return out;
}

struct SampletTreeH {
  fmca_samplet_tree *h = nullptr;
  int dim = 0;
  int64_t n = 0;
  SampletTreeH() = default;
  SampletTreeH(const SampletTreeH &) = delete;
  ~SampletTreeH() {
    if (h) fmca_samplet_tree_destroy(h);
  }
};

struct SampletMatrixH {
  fmca_samplet_matrix *h = nullptr;
  int64_t n = 0;
  SampletMatrixH() = default;
  SampletMatrixH(const SampletMatrixH &) = delete;
  ~SampletMatrixH() {
    if (h) fmca_samplet_matrix_destroy(h);
  }
};

std::unique_ptr<SampletTreeH> samplet_tree(const py::array &P_in, int dtilde) {
  farr P = as_f64(P_in, "P");
  if (P.ndim() != 2) throw py::value_error("points must be a (dim, n) array");
  auto t = std::make_unique<SampletTreeH>();
  t->dim = (int)P.shape(0);
  t->n = P.shape(1);
  {
    py::gil_scoped_release nogil;
    check(fmca_samplet_tree_create(t->dim, t->n, P.data(), dtilde, &t->h));
  }
  return t;
}

py::array samplet_transform(const SampletTreeH &t, const py::array &in_, bool inverse) {
  farr in = as_f64(in_, "data");
  if (in.ndim() < 1 || in.ndim() > 2 || in.shape(0) != t.n)
    throw py::value_error("data must have n rows");
  const int64_t k = in.ndim() == 2 ? in.shape(1) : 1;
  farr out({(py::ssize_t)t.n, (py::ssize_t)k});
  double *po = out.mutable_data();
  const double *pi = in.data();
  {
    py::gil_scoped_release nogil;
    check(fmca_samplet_transform(t.h, inverse ? 1 : 0, pi, t.n, k, po, t.n));
  }
  return out;
}

py::dict samplet_clusters(const SampletTreeH &t) {
  int64_t info[7];
  check(fmca_samplet_tree_info(t.h, info));
  const py::ssize_t ncl = info[3];
  py::array_t<int64_t> begin(ncl), size(ncl), start(ncl);
  py::array_t<int32_t> sons({ncl, (py::ssize_t)2}), qdim(ncl), nscal(ncl);
  check(fmca_samplet_tree_clusters(t.h, begin.mutable_data(), size.mutable_data(),
                                   sons.mutable_data(), qdim.mutable_data(),
                                   nscal.mutable_data(), start.mutable_data()));
  py::dict d;
  d["begin"] = begin;
  d["size"] = size;
  d["sons"] = sons;
  d["qdim"] = qdim;
  d["nscal"] = nscal;
  d["start"] = start;
  d["q"] = info[4];
  d["mdtilde"] = info[5];
  d["depth"] = info[6];
  return d;
}

py::tuple samplet_basis(const SampletTreeH &t, int c) {
  int64_t info[7];
  check(fmca_samplet_tree_info(t.h, info));
  if (c < 0 || c >= info[3]) throw py::value_error("cluster out of range");
  int32_t qd = 0;
  std::vector<int32_t> qdim(info[3]);
  check(fmca_samplet_tree_clusters(t.h, nullptr, nullptr, nullptr, qdim.data(), nullptr, nullptr));
  qd = qdim[c];
  farr Q({(py::ssize_t)qd, (py::ssize_t)qd}), V({(py::ssize_t)info[4], (py::ssize_t)qd});
  check(fmca_samplet_tree_basis(t.h, c, Q.mutable_data(), V.mutable_data()));
  return py::make_tuple(Q, V);
}

std::unique_ptr<SampletMatrixH> samplet_compress(const SampletTreeH &t, const std::string &ktype,
                                                 double l, double c, double nu, double eta,
                                                 double thres) {
  auto M = std::make_unique<SampletMatrixH>();
  M->n = t.n;
  fmca_kernel_desc k{ktype.c_str(), l, c, nu};
  {
    py::gil_scoped_release nogil;
    check(fmca_samplet_compress(t.h, &k, eta, thres, &M->h));
  }
  return M;
}

py::tuple samplet_triplets(const SampletMatrixH &M) {
  int64_t nnz = 0;
  check(fmca_samplet_matrix_nnz(M.h, &nnz));
  py::array_t<int64_t> rows(nnz), cols(nnz);
  py::array_t<double> vals(nnz);
  check(fmca_samplet_matrix_triplets(M.h, rows.mutable_data(), cols.mutable_data(),
                                     vals.mutable_data()));
  return py::make_tuple(rows, cols, vals);
}

py::array_t<int64_t> samplet_aposteriori(const SampletMatrixH &M, double thres) {
  int64_t nnz = 0, count = 0;
  check(fmca_samplet_matrix_nnz(M.h, &nnz));
  std::vector<int64_t> idx(nnz);
  check(fmca_samplet_matrix_aposteriori(M.h, thres, &count, idx.data()));
  py::array_t<int64_t> out(count);
  std::copy(idx.begin(), idx.begin() + count, out.mutable_data());
  return out;
}

}  // namespace

PYBIND11_MODULE(_fmca_amd_py, m) {
  m.doc() = "MI355X-native H2-matrix kernel action (C ABI bindings)";
  m.attr("FP64") = (int)FMCA_PRECISION_FP64;
  m.attr("FP32") = (int)FMCA_PRECISION_FP32;
  m.attr("FP32_BF16") = (int)FMCA_PRECISION_FP32_BF16;
  m.attr("ROWS") = (int)FMCA_ROWS;
  m.attr("COLS") = (int)FMCA_COLS;
  m.attr("KERNEL_VALUE") = (int)FMCA_KERNEL_VALUE;
  m.attr("KERNEL_DL") = (int)FMCA_KERNEL_DL;
  m.attr("KERNEL_PHI") = (int)FMCA_KERNEL_PHI;
  m.attr("INTERP_TENSOR_PRODUCT") = (int)FMCA_INTERP_TENSOR_PRODUCT;
  m.attr("INTERP_TOTAL_DEGREE") = (int)FMCA_INTERP_TOTAL_DEGREE;
  m.attr("PLAN_HOST") = (int)FMCA_PLAN_HOST;
  m.attr("PLAN_DEVICE") = (int)FMCA_PLAN_DEVICE;
  py::class_<Actor>(m, "Actor")
      .def(py::init(&create), py::arg("ktype"), py::arg("l"), py::arg("c"), py::arg("nu"),
           py::arg("Peval"), py::arg("P"), py::arg("deg"), py::arg("eta"),
           py::arg("interpolator") = (int)FMCA_INTERP_TENSOR_PRODUCT,
           py::arg("planner") = (int)FMCA_PLAN_HOST)
      .def("assemble",
           [](Actor &a, int precision) {
             int64_t bytes = 0;
             {
               py::gil_scoped_release nogil;
               check(fmca_h2_assemble(a.h, precision, &bytes));
             }
             return bytes;
           })
      .def("release_assembled", [](Actor &a) { check(fmca_h2_release_assembled(a.h)); })
      .def("plan_report",
           [](const Actor &a) {
             double ts = 0, ps = 0;
             int where[3] = {0, 0, 0};
             check(fmca_h2_plan_report(a.h, &ts, &ps, where));
             py::dict d;
             d["tree_s"] = ts;
             d["pattern_s"] = ps;
             d["rows_tree"] = where[0];
             d["cols_tree"] = where[1];
             d["blocks"] = where[2];
             return d;
           })
      .def_readonly("dim", &Actor::dim)
      .def_readonly("m", &Actor::m)
      .def_readonly("n", &Actor::n)
      .def("upload",
           [](Actor &a) {
             py::gil_scoped_release nogil;
             check(fmca_h2_upload(a.h));
           })
      .def("action", &action, py::arg("rhs"), py::arg("precision") = 0,
           py::arg("cluster_order") = false)
      .def(
          "action_device",
          [](Actor &a, uintptr_t rhs, int64_t ld_rhs, int64_t nrhs, uintptr_t out, int64_t ld_out,
             int precision, bool cluster_order, uintptr_t stream) {
            py::gil_scoped_release nogil;
            check(fmca_h2_action_device(a.h, reinterpret_cast<const double *>(rhs), ld_rhs, nrhs,
                                        reinterpret_cast<double *>(out), ld_out, precision,
                                        cluster_order ? 1 : 0, reinterpret_cast<void *>(stream)));
          },
          py::arg("rhs"), py::arg("ld_rhs"), py::arg("nrhs"), py::arg("out"), py::arg("ld_out"),
          py::arg("precision") = 0, py::arg("cluster_order") = false, py::arg("stream") = 0)
      .def(
          "action_transpose_device",
          [](Actor &a, uintptr_t rhs, int64_t ld_rhs, int64_t nrhs, uintptr_t out, int64_t ld_out,
             int precision, uintptr_t stream) {
            py::gil_scoped_release nogil;
            check(fmca_h2_action_transpose_device(
                a.h, reinterpret_cast<const double *>(rhs), ld_rhs, nrhs,
                reinterpret_cast<double *>(out), ld_out, precision,
                reinterpret_cast<void *>(stream)));
          },
          py::arg("rhs"), py::arg("ld_rhs"), py::arg("nrhs"), py::arg("out"), py::arg("ld_out"),
          py::arg("precision") = 0, py::arg("stream") = 0)
      .def(
          "cg_device",
          [](Actor &a, uintptr_t b, int64_t ld_b, int64_t nrhs, uintptr_t x, int64_t ld_x,
             double ridge, double tol, int precision, double refine_factor, int max_iter,
             int max_outer, int method, int max_basis, uintptr_t stream) {
            fmca_cg_options o{ridge,    tol,       precision, refine_factor,
                              max_iter, max_outer, method,    max_basis};
            fmca_cg_result r{};
            fmca_status st;
            {
              py::gil_scoped_release nogil;
              st = fmca_h2_cg_device(a.h, reinterpret_cast<const double *>(b), ld_b, nrhs,
                                     reinterpret_cast<double *>(x), ld_x, &o, &r,
                                     reinterpret_cast<void *>(stream));
            }
            if (st != FMCA_OK && st != FMCA_ERR_NOT_CONVERGED) check(st);
            py::dict d;
            d["converged"] = (bool)r.converged;
            d["iterations"] = r.iterations;
            d["outer"] = r.outer;
            d["matvecs_fp64"] = r.matvecs_fp64;
            d["matvecs_fp32"] = r.matvecs_fp32;
            d["rel_residual"] = r.rel_residual;
            d["indefinite"] = r.indefinite != 0;
            d["seconds"] = r.seconds;
            return d;
          },
          py::arg("b"), py::arg("ld_b"), py::arg("nrhs"), py::arg("x"), py::arg("ld_x"),
          py::arg("ridge"), py::arg("tol"), py::arg("precision"), py::arg("refine_factor"),
          py::arg("max_iter"), py::arg("max_outer"), py::arg("method") = 0,
          py::arg("max_basis") = 0, py::arg("stream") = 0)
      .def("set_kernel",
           [](Actor &a, const std::string &ktype, double l, double c, double nu) {
             fmca_kernel_desc k{ktype.c_str(), l, c, nu};
             check(fmca_h2_set_kernel(a.h, &k));
           })
      .def("set_distance_type",
           [](Actor &a, const std::string &dist) {
             check(fmca_h2_set_distance_type(a.h, dist.c_str()));
           })
      .def("distance_type",
           [](const Actor &a) {
             int g = 0;
             check(fmca_h2_distance_type(a.h, &g));
             return std::string(g ? "GEODESIC" : "EUCLIDEAN");
           })
      .def(
          "action_variants_device",
          [](Actor &a, std::vector<int> variants, bool transpose, uintptr_t rhs, int64_t ld_rhs,
             int64_t nrhs, uintptr_t out, int64_t ld_out, int precision, uintptr_t stream) {
            py::gil_scoped_release nogil;
            check(fmca_h2_action_variants_device(
                a.h, variants.data(), (int)variants.size(), transpose ? 1 : 0,
                reinterpret_cast<const double *>(rhs), ld_rhs, nrhs,
                reinterpret_cast<double *>(out), ld_out, precision,
                reinterpret_cast<void *>(stream)));
          },
          py::arg("variants"), py::arg("transpose"), py::arg("rhs"), py::arg("ld_rhs"),
          py::arg("nrhs"), py::arg("out"), py::arg("ld_out"), py::arg("precision"),
          py::arg("stream") = 0)
      .def(
          "action_variant_device",
          [](Actor &a, int variant, bool transpose, uintptr_t rhs, int64_t ld_rhs, int64_t nrhs,
             uintptr_t out, int64_t ld_out, int precision, uintptr_t stream) {
            py::gil_scoped_release nogil;
            check(fmca_h2_action_variant_device(
                a.h, variant, transpose ? 1 : 0, reinterpret_cast<const double *>(rhs), ld_rhs,
                nrhs, reinterpret_cast<double *>(out), ld_out, precision,
                reinterpret_cast<void *>(stream)));
          },
          py::arg("variant"), py::arg("transpose"), py::arg("rhs"), py::arg("ld_rhs"),
          py::arg("nrhs"), py::arg("out"), py::arg("ld_out"), py::arg("precision") = 0,
          py::arg("stream") = 0)
      .def("indices",
           [](const Actor &a, int which) {
             py::array_t<uint32_t> out(which == FMCA_ROWS ? a.m : a.n);
             check(fmca_h2_indices(a.h, which, out.mutable_data()));
             return out;
           })
      .def("statistics",
           [](const Actor &a) {
             py::array_t<double> out({(py::ssize_t)6, (py::ssize_t)1});
             check(fmca_h2_statistics(a.h, out.mutable_data()));
             return out;
           })
      .def("gradient_device",
           [](Actor &a, uintptr_t alpha, int64_t ld_alpha, uintptr_t g, int64_t ld_g,
              int64_t nrhs, int precision, uintptr_t dalpha, int64_t ld_dalpha,
              uintptr_t dpoints, int64_t ld_points, uintptr_t dl, uintptr_t stream) {
             py::gil_scoped_release nogil;
             check(fmca_h2_gradient_device(
                 a.h, reinterpret_cast<const double *>(alpha), ld_alpha,
                 reinterpret_cast<const double *>(g), ld_g, nrhs, precision,
                 reinterpret_cast<double *>(dalpha), ld_dalpha,
                 reinterpret_cast<double *>(dpoints), ld_points,
                 reinterpret_cast<double *>(dl), reinterpret_cast<void *>(stream)));
           })
      .def("project",
           [](Actor &a, const py::array_t<double, py::array::f_style | py::array::forcecast> &M,
              int precision) {
             if (M.ndim() != 2 || M.shape(0) != a.m || M.shape(1) != a.n)
               throw py::value_error("project: M must be m x n");
             double chk[2] = {0, 0};
             {
               py::gil_scoped_release nogil;
               check(fmca_h2_project(a.h, M.data(), a.m, precision, chk));
             }
             return py::make_tuple(chk[0], chk[1]);
           })
      .def("set_profiling", [](Actor &a, bool on) { check(fmca_h2_set_profiling(a.h, on ? 1 : 0)); })
      .def(
          "phase_times",
          [](Actor &a, bool reset, int precision) {
            double ms[6];
            int64_t cnt[6];
            {
              py::gil_scoped_release nogil;
              check(fmca_h2_phase_times_precision(a.h, precision, ms, cnt, reset ? 1 : 0));
            }
            py::dict d;
            const char *names[6] = {"permute",  "upward_leaf",        "upward_transfer",
                                    "coupling", "downward_transfer", "leaf_output"};
            for (int k = 0; k < 6; ++k) d[names[k]] = py::make_tuple(ms[k], cnt[k]);
            return d;
          },
          py::arg("reset") = false, py::arg("precision") = -1)
      .def("info", &info)
      .def("interpolation",
           [](const Actor &a) {
             int kind = 0, q = 0;
             check(fmca_h2_interpolation(a.h, &kind, &q, nullptr, nullptr, nullptr));
             py::array_t<int32_t> mi({(py::ssize_t)q, (py::ssize_t)a.dim});
             farr Xi({(py::ssize_t)a.dim, (py::ssize_t)q});
             farr invV({(py::ssize_t)q, (py::ssize_t)q});
             check(fmca_h2_interpolation(a.h, nullptr, nullptr, mi.mutable_data(),
                                         Xi.mutable_data(), invV.mutable_data()));
             py::dict d;
             d["interpolator"] = kind;
             d["q"] = q;
             d["multi_indices"] = mi;
             d["Xi"] = Xi;
             d["invV"] = invV;
             return d;
           })
      .def("clusters", &clusters)
      .def("blocks", &blocks)
      .def("launch_order", &launch_order, py::arg("which"));
  m.def(
      "diag_math",
      [](int fn, uintptr_t x, uintptr_t y, int64_t n, uintptr_t stream) {
        py::gil_scoped_release nogil;
        check(fmca_diag_math(fn, reinterpret_cast<const double *>(x),
                             reinterpret_cast<double *>(y), n, reinterpret_cast<void *>(stream)));
      },
      py::arg("fn"), py::arg("x"), py::arg("y"), py::arg("n"), py::arg("stream") = 0);
  m.def("set_device_memory_limit", [](int64_t bytes) { check(fmca_set_device_memory_limit(bytes)); });
  m.def("device_memory", [](bool reset_peak) {
    int64_t live = 0, peak = 0;
    check(fmca_device_memory(&live, &peak, reset_peak ? 1 : 0));
    return py::make_tuple(live, peak);
  }, py::arg("reset_peak") = false,
        "(live, peak) device bytes held by the library's plans and workspaces");
  m.def("status_string", [](int s) { return std::string(fmca_status_string((fmca_status)s)); });
  py::class_<SampletTreeH>(m, "SampletTreeHandle")
      .def(py::init(&samplet_tree), py::arg("P"), py::arg("dtilde"))
      .def_readonly("n", &SampletTreeH::n)
      .def_readonly("dim", &SampletTreeH::dim)
      .def("indices",
           [](const SampletTreeH &t) {
             py::array_t<uint32_t> out(t.n);
             check(fmca_samplet_tree_indices(t.h, out.mutable_data()));
             return out;
           })
      .def("levels",
           [](const SampletTreeH &t) {
             py::array_t<int32_t> out(t.n);
             check(fmca_samplet_tree_levels(t.h, out.mutable_data()));
             return out;
           })
      .def("transform", &samplet_transform, py::arg("data"), py::arg("inverse") = false)
      .def("clusters", &samplet_clusters)
      .def("basis", &samplet_basis, py::arg("cluster"))
      .def("compress", &samplet_compress, py::arg("ktype"), py::arg("l"), py::arg("c"),
           py::arg("nu"), py::arg("eta"), py::arg("threshold"));
  py::class_<SampletMatrixH>(m, "SampletMatrixHandle")
      .def_readonly("n", &SampletMatrixH::n)
      .def("triplets", &samplet_triplets)
      .def("aposteriori", &samplet_aposteriori, py::arg("thres"))
      .def("stats", [](const SampletMatrixH &M) {
        std::vector<double> st(FMCA_SAMPLET_STATS);
        check(fmca_samplet_matrix_stats(M.h, st.data()));
        return st;
      });
}
