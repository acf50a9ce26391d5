// Projection of a given matrix onto the H2 format of a plan
// (H2MatrixProjector, FMCA/src/H2Matrix/H2MatrixProjector.h:21-75).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "plan.hpp"

namespace fmca_amd {

struct ProjectedMatrix {
  // coupling matrix of every far block, in far-list order, q x q row-major
  // over the planner's tensor order
  std::vector<double> far;
  // every near block in near-list order, |t| x |s| column-major, rows and
  // columns in cluster order; block e at near_off[e]
  std::vector<double> near;
  std::vector<int64_t> near_off;
  // FMCA's dual-basis check (compute_dual_cluster_bases_impl.h:66-86) on the
  // row and column trees: max over clusters of
  // ||V^T - V^T (V V^T)^-1 V V^T||_F / ||V^T||_F (FMCA asserts < 5e-10)
  double dual_check_rows = 0, dual_check_cols = 0;
};

// M: m x n, column-major with leading dimension ld, natural order (rows of
// the row tree's points, columns of the column tree's).  pts_*: the trees'
// points dim-major in cluster order.  False (message in *err) for the
// total-degree interpolator or a singular basis Gram matrix.
bool project_matrix(const Tree &rows, const Tree &cols, bool shared, const Interpolation &ip,
                    const std::vector<double> &trans_rows, const std::vector<double> &trans_cols,
                    const std::vector<double> &pts_rows, const std::vector<double> &pts_cols,
                    const Pattern &pat, const double *M, int64_t ld, ProjectedMatrix &out,
                    std::string *err);

}  // namespace fmca_amd
