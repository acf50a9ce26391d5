// C++ facade over the C ABI with FMCA's H2kernelActor interface.
//
// Drop-in for FMCA::H2kernelActor<Covariance, Moments, H2ClusterTree,
// MatrixEvaluator> (FMCA/src/H2Matrix/H2kernelActor.h:17-94): the same
// constructor (kernel, Peval, P, mpole_deg, eta) and `M action(const M &rhs)`.
// The second template argument is the Moments type:
// FMCA::NystromMoments<TensorProductInterpolator> (or
// fmca_amd::TensorProductMoments) selects the tensor-product Chebyshev
// moments, FMCA::NystromMoments<TotalDegreeInterpolator> (or
// fmca_amd::TotalDegreeMoments) the total-degree form; any other Moments type
// (e.g. NystromMoments<MonomialInterpolator>) fails to compile.  The
// arguments after it are accepted and ignored (the tree is a ClusterTree with
// CardinalityBisection, the evaluator unsymmetric Nystrom).
//
// Matrix types: anything column-major with rows(), cols(), data() and a
// (rows, cols) constructor -- Eigen::MatrixXd qualifies.  Kernel types:
// anything with kernelType(), l(), c(), nu() -- FMCA::CovarianceKernel
// (CovarianceKernel.h:154-161) and fmca_amd::CovarianceKernel qualify.
//
// Errors raise fmca_amd::Error (FMCA asserts instead).
#pragma once

#include <cctype>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../fmca_amd.h"

namespace fmca_amd {

class Error : public std::runtime_error {
 public:
  Error(fmca_status s, const std::string &what)
      : std::runtime_error(what + ": " + fmca_status_string(s) + " (" + fmca_last_error() + ")"),
        status_(s) {}
  fmca_status status() const { return status_; }

 private:
  fmca_status status_;
};

inline void check(fmca_status s, const char *what) {
  if (s != FMCA_OK) throw Error(s, what);
}

// FMCA::CovarianceKernel(ktype, l, c, nu) with its parameters held by value.
class CovarianceKernel {
 public:
  CovarianceKernel() : CovarianceKernel("EXPONENTIAL") {}
  explicit CovarianceKernel(const std::string &ktype, double l = 1., double c = 1.,
                            double nu = 1.)
      : ktype_(ktype), l_(l), c_(c), nu_(nu) {
    for (auto &ch : ktype_) ch = (char)std::toupper((unsigned char)ch);
  }
  std::string kernelType() const { return ktype_; }
  const double &l() const { return l_; }
  const double &c() const { return c_; }
  const double &nu() const { return nu_; }
  // CovarianceKernel::setDistanceType (CovarianceKernel.h:123-137):
  // "EUCLIDEAN" or "GEODESIC"; checked when an actor takes the kernel.
  void setDistanceType(const std::string &dist_type) { dist_ = dist_type; }
  const std::string &distanceType() const { return dist_; }

 private:
  std::string ktype_;
  double l_, c_, nu_;
  std::string dist_ = "EUCLIDEAN";
};

// Whether a kernel class says its distance type (this header's kernel does;
// FMCA::CovarianceKernel keeps it private, so an actor built from one uses
// the Euclidean distance).
template <typename K, typename = void>
struct has_distance_type : std::false_type {};
template <typename K>
struct has_distance_type<K, std::void_t<decltype(std::declval<const K &>().distanceType())>>
    : std::true_type {};

enum class Precision : int { FP64 = FMCA_PRECISION_FP64, FP32 = FMCA_PRECISION_FP32 };

// Where the constructor builds the cluster trees and the block tree
// (fmca_h2_create_ex): the host planner, or the current HIP device.
enum class Planner : int { Host = FMCA_PLAN_HOST, Device = FMCA_PLAN_DEVICE };

// Moments tags for callers without FMCA's headers.
struct TensorProductMoments {};
struct TotalDegreeMoments {};

}  // namespace fmca_amd

// FMCA's interpolators and moments class (FMCA/src/Interpolator/
// TensorProductInterpolator.h:31, TotalDegreeInterpolator.h:19,
// FMCA/src/Moments/NystromMoments.h:16-17), declared so that the mapping
// below names them whether or not FMCA's headers are included.
namespace FMCA {
class TensorProductInterpolator;
class TotalDegreeInterpolator;
template <typename Interpolator>
class NystromMoments;
}  // namespace FMCA

namespace fmca_amd {

// FMCA_INTERP_* of an interpolator type; -1 for an interpolator this library
// does not build (MonomialInterpolator, WeightedTotalDegreeInterpolator, ...).
template <typename Interpolator>
struct interpolator_id : std::integral_constant<int, -1> {};
template <>
struct interpolator_id<FMCA::TensorProductInterpolator>
    : std::integral_constant<int, FMCA_INTERP_TENSOR_PRODUCT> {};
template <>
struct interpolator_id<FMCA::TotalDegreeInterpolator>
    : std::integral_constant<int, FMCA_INTERP_TOTAL_DEGREE> {};

// FMCA_INTERP_* of a Moments type: NystromMoments<I> maps through I, the two
// tags above map to themselves; any other Moments type is -1.
template <typename Moments>
struct moments_interpolator : std::integral_constant<int, -1> {};
template <typename Interpolator>
struct moments_interpolator<FMCA::NystromMoments<Interpolator>>
    : interpolator_id<Interpolator> {};
template <>
struct moments_interpolator<TensorProductMoments>
    : std::integral_constant<int, FMCA_INTERP_TENSOR_PRODUCT> {};
template <>
struct moments_interpolator<TotalDegreeMoments>
    : std::integral_constant<int, FMCA_INTERP_TOTAL_DEGREE> {};

template <typename Covariance = CovarianceKernel, typename Moments = TensorProductMoments,
          typename... FmcaPolicies>
class H2kernelActor {
  static_assert(moments_interpolator<Moments>::value >= 0,
                "fmca_amd::H2kernelActor: Moments must be NystromMoments<TensorProductInterpolator>, "
                "NystromMoments<TotalDegreeInterpolator>, TensorProductMoments or "
                "TotalDegreeMoments");

 public:
  // FMCA's constructor (H2kernelActor.h:21-36); the planner argument is an
  // addition with a default, so FMCA's calls compile unchanged.
  template <typename PointMatrix>
  H2kernelActor(const Covariance &kernel, const PointMatrix &Peval, const PointMatrix &P,
                unsigned mpole_deg, double eta, Planner planner = Planner::Host)
      : m_((int64_t)Peval.cols()), n_((int64_t)P.cols()) {
    const std::string kt = kernel.kernelType();
    fmca_kernel_desc desc{kt.c_str(), (double)kernel.l(), (double)kernel.c(),
                          (double)kernel.nu()};
    if (Peval.rows() != P.rows()) throw Error(FMCA_ERR_INVALID_ARGUMENT, "point dimensions");
    fmca_h2_actor *h = nullptr;
    check(fmca_h2_create_ex(&desc, (int)P.rows(), m_, Peval.data(), n_, P.data(),
                            (int)mpole_deg, eta, moments_interpolator<Moments>::value,
                            (int)planner, &h),
          "fmca_h2_create_ex");
    h_.reset(h, [](fmca_h2_actor *p) { fmca_h2_destroy(p); });
    if constexpr (has_distance_type<Covariance>::value)
      check(fmca_h2_set_distance_type(h_.get(), std::string(kernel.distanceType()).c_str()),
            "fmca_h2_set_distance_type");
  }

  // H2kernelActor::action (H2kernelActor.h:38-80): K(Peval, P) * rhs, natural order.
  template <typename M>
  M action(const M &rhs, Precision precision = Precision::FP64) const {
    if ((int64_t)rhs.rows() != n_) throw Error(FMCA_ERR_INVALID_ARGUMENT, "rhs rows");
    M out((decltype(rhs.rows()))m_, rhs.cols());
    check(fmca_h2_action(h_.get(), rhs.data(), n_, (int64_t)rhs.cols(), out.data(), m_,
                         (int)precision),
          "fmca_h2_action");
    return out;
  }

  // H2MatrixBase::action (H2MatrixBase.h:113-152): cluster-ordered rhs and result.
  template <typename M>
  M action_cluster_order(const M &rhs, Precision precision = Precision::FP64) const {
    if ((int64_t)rhs.rows() != n_) throw Error(FMCA_ERR_INVALID_ARGUMENT, "rhs rows");
    M out((decltype(rhs.rows()))m_, rhs.cols());
    check(fmca_h2_action_cluster_order(h_.get(), rhs.data(), n_, (int64_t)rhs.cols(),
                                       out.data(), m_, (int)precision),
          "fmca_h2_action_cluster_order");
    return out;
  }

  // ClusterTreeBase::indices of the row (FMCA_ROWS) or column (FMCA_COLS) tree.
  std::vector<uint32_t> indices(int which) const {
    std::vector<uint32_t> out(which == FMCA_ROWS ? m_ : n_);
    check(fmca_h2_indices(h_.get(), which, out.data()), "fmca_h2_indices");
    return out;
  }

  // H2MatrixBase::statistics (H2MatrixBase.h:70-105).
  std::vector<double> statistics() const {
    std::vector<double> out(6);
    check(fmca_h2_statistics(h_.get(), out.data()), "fmca_h2_statistics");
    return out;
  }

  // H2Matrix::computeH2Matrix (H2MatrixBase.h:222-255): store every block in
  // device memory; later actions of this precision stream them (operator*).
  // Returns the bytes held.
  int64_t assemble(Precision precision = Precision::FP64) {
    int64_t bytes = 0;
    check(fmca_h2_assemble(h_.get(), (int)precision, &bytes), "fmca_h2_assemble");
    return bytes;
  }
  void release_assembled() {
    check(fmca_h2_release_assembled(h_.get()), "fmca_h2_release_assembled");
  }

  // The distance of the kernel entries, "EUCLIDEAN" or "GEODESIC"
  // (CovarianceKernel::setDistanceType, CovarianceKernel.h:123-137), for a
  // kernel class that does not say it (FMCA::CovarianceKernel).
  void setDistanceType(const std::string &dist_type) {
    check(fmca_h2_set_distance_type(h_.get(), dist_type.c_str()), "fmca_h2_set_distance_type");
  }

  int64_t rows() const { return m_; }
  int64_t cols() const { return n_; }
  fmca_h2_actor *handle() const { return h_.get(); }

 private:
  std::shared_ptr<fmca_h2_actor> h_;
  int64_t m_, n_;
};

}  // namespace fmca_amd
