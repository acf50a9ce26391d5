// Exercises include/fmca_amd/H2kernelActor.hpp the way FMCA code uses
// FMCA::H2kernelActor: construct from (kernel, Peval, P, deg, eta), call
// action(rhs) on a column-major matrix.  Reads points and rhs from a binary
// file, writes the result; "plan" mode only builds the actor.
//   facade_check plan|action|plan_td|action_td|action_dev|action_td_dev|
//                action_geo|action_geo_set <in.bin> <out.bin>
// (the _td modes instantiate the actor with a Moments type shaped like
// FMCA's NystromMoments<TotalDegreeInterpolator>; action_geo takes the
// facade's CovarianceKernel with setDistanceType("GEODESIC"), action_geo_set
// a kernel class without a distance and the actor's setDistanceType)
// in.bin: int64 dim, n, nrhs, deg; double eta, l; char[16] kernel;
//         double P[dim*n] (column-major); double rhs[n*nrhs]
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "fmca_amd/H2kernelActor.hpp"

// minimal column-major matrix with Eigen's rows()/cols()/data() interface
struct Mat {
  Mat(long r, long c) : r_(r), c_(c), v_((size_t)r * c) {}
  long rows() const { return r_; }
  long cols() const { return c_; }
  const double *data() const { return v_.data(); }
  double *data() { return v_.data(); }
  long r_, c_;
  std::vector<double> v_;
};

// stand-ins with FMCA's names for NystromMoments<TotalDegreeInterpolator> and
// NystromMoments<TensorProductInterpolator> (FMCA's headers need Eigen): the
// facade maps the Moments type by the interpolator it is instantiated with
namespace FMCA {
class TensorProductInterpolator {};
class TotalDegreeInterpolator {};
template <typename Interpolator>
class NystromMoments {};
}  // namespace FMCA

template <typename Actor>
int run(Actor &act, const char *mode, const Mat &rhs, const char *path) {
  std::ofstream out(path, std::ios::binary);
  if (std::strncmp(mode, "plan", 4) == 0) {
    auto idx = act.indices(FMCA_COLS);
    auto st = act.statistics();
    out.write((const char *)idx.data(), sizeof(uint32_t) * idx.size());
    out.write((const char *)st.data(), sizeof(double) * st.size());
    return 0;
  }
  if (std::strstr(mode, "_dev") && act.assemble(fmca_amd::Precision::FP64) <= 0) return 4;
  Mat y = act.action(rhs);
  out.write((const char *)y.data(), sizeof(double) * y.rows() * y.cols());
  return 0;
}

// a kernel type of the caller's own, with the FMCA::CovarianceKernel accessors
struct MyKernel {
  std::string t;
  double l_;
  std::string kernelType() const { return t; }
  double l() const { return l_; }
  double c() const { return 1.0; }
  double nu() const { return 1.0; }
};

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  std::ifstream in(argv[2], std::ios::binary);
  int64_t hdr[4];
  double par[2];
  char kname[16];
  in.read((char *)hdr, sizeof(hdr));
  in.read((char *)par, sizeof(par));
  in.read(kname, 16);
  const long dim = hdr[0], n = hdr[1], nrhs = hdr[2];
  Mat P(dim, n), rhs(n, nrhs);
  in.read((char *)P.data(), sizeof(double) * dim * n);
  in.read((char *)rhs.data(), sizeof(double) * n * nrhs);
  if (!in) return 3;
  try {
    const MyKernel kernel{std::string(kname), par[1]};
    if (std::strcmp(argv[1], "action_geo") == 0) {
      fmca_amd::CovarianceKernel k(std::string(kname), par[1]);
      k.setDistanceType("GEODESIC");
      fmca_amd::H2kernelActor<fmca_amd::CovarianceKernel> act(k, P, P, (unsigned)hdr[3], par[0]);
      return run(act, argv[1], rhs, argv[3]);
    }
    if (std::strcmp(argv[1], "action_geo_set") == 0) {
      fmca_amd::H2kernelActor<MyKernel> act(kernel, P, P, (unsigned)hdr[3], par[0]);
      act.setDistanceType("GEODESIC");
      return run(act, argv[1], rhs, argv[3]);
    }
    const bool td = std::strstr(argv[1], "_td") != nullptr;
    // "_dev": plan on the device and stream assembled blocks
    const auto planner = std::strstr(argv[1], "_dev") ? fmca_amd::Planner::Device
                                                       : fmca_amd::Planner::Host;
    // FMCA's four template arguments: kernel, Moments, tree, evaluator
    if (td) {
      fmca_amd::H2kernelActor<MyKernel, FMCA::NystromMoments<FMCA::TotalDegreeInterpolator>,
                              struct H2ClusterTree,
                              struct Evaluator>
          act(kernel, P, P, (unsigned)hdr[3], par[0], planner);
      return run(act, argv[1], rhs, argv[3]);
    }
    fmca_amd::H2kernelActor<MyKernel, FMCA::NystromMoments<FMCA::TensorProductInterpolator>,
                            struct H2ClusterTree,
                            struct Evaluator>
        act(kernel, P, P, (unsigned)hdr[3], par[0], planner);
    return run(act, argv[1], rhs, argv[3]);
  } catch (const fmca_amd::Error &e) {
    std::cerr << e.what() << std::endl;
    return 10 + (int)e.status();
  }
  return 0;
}
