// Runs the host planner (tree, pattern, tasks, launch orders, transfers) and
// the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer on
// point sets that exercise the edge paths: one point, fewer points than q,
// duplicates (flat boxes, tie fallback), a flat axis, rectangular target
// sets, both interpolators.  The planner's permutation and block count must
// equal the oracle's.  Exit status 0 on success.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "plan.hpp"

extern "C" {
void *orc_create(int dim, long n, const double *P, long m, const double *Peval, int ktype,
                 double l, double c, double nu, int deg, double eta, int scheme);
void orc_destroy(void *h);
void orc_indices(void *h, int which, uint32_t *out);
long orc_num_blocks(void *h);
}

using namespace fmca_amd;

static std::vector<double> points(int dim, int64_t n, int kind, unsigned seed) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> U(0, 1);
  std::vector<double> P((size_t)dim * n);
  std::vector<double> base((size_t)dim * 7);
  for (auto &v : base) v = U(g);
  for (int64_t j = 0; j < n; ++j)
    for (int i = 0; i < dim; ++i) {
      double v = U(g);
      if (kind == 1) v = base[(size_t)dim * (j % 7) + i];   // duplicates
      if (kind == 2 && i == dim - 1) v = 0.5;                // flat axis
      P[(size_t)dim * j + i] = v;
    }
  return P;
}

int main() {
  struct Case { int dim; int64_t n, m; int kind, deg, scheme; };
  const Case cases[] = {{3, 1, 1, 0, 3, 0},       {3, 40, 30, 0, 4, 1},  {3, 5000, 5000, 1, 3, 0},
                        {3, 5000, 5000, 1, 3, 1}, {3, 8000, 8000, 2, 2, 0}, {2, 6000, 4000, 0, 4, 1},
                        {1, 3000, 3000, 0, 5, 0}, {3, 20000, 20000, 0, 4, 1}};
  int bad = 0;
  for (const Case &cs : cases) {
    const std::vector<double> P = points(cs.dim, cs.n, cs.kind, 1);
    const std::vector<double> Pe = cs.m == cs.n ? P : points(cs.dim, cs.m, cs.kind, 2);
    Interpolation ip;
    ip.init(cs.dim, cs.deg, cs.scheme ? IS_TOTAL_DEGREE : IS_TENSOR);
    Tree R, C;
    build_tree(R, cs.dim, cs.m, Pe.data(), ip.q);
    build_tree(C, cs.dim, cs.n, P.data(), ip.q);
    Pattern pat;
    compute_pattern(pat, R, C, 0.8, ip.q);
    std::vector<double> E;
    if (cs.scheme)
      compute_transfers_dense(E, R, ip);
    else
      compute_transfers(E, R, ip);
    void *h = orc_create(cs.dim, (long)cs.n, P.data(), (long)cs.m, Pe.data(), 0, 0.5, 1.0, 1.0,
                         cs.deg, 0.8, cs.scheme);
    std::vector<uint32_t> ir(cs.m), ic(cs.n);
    orc_indices(h, 0, ir.data());
    orc_indices(h, 1, ic.data());
    const bool ok = ir == R.perm && ic == C.perm && orc_num_blocks(h) == (long)pat.block_row.size();
    orc_destroy(h);
    std::printf("dim %d n %lld m %lld kind %d deg %d scheme %d: %s\n", cs.dim, (long long)cs.n,
                (long long)cs.m, cs.kind, cs.deg, cs.scheme, ok ? "ok" : "MISMATCH");
    bad += !ok;
  }
  return bad ? 1 : 0;
}
