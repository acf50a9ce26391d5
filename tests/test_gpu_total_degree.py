"""GPU parity of the total-degree form (TotalDegreeInterpolator) with the oracle.

FMCA's own actor test (tests/test_H2kernelActor.cpp:20-24) and its Python
binding (py/FMCA.cpp:37) build NystromMoments on TotalDegreeInterpolator.
The GPU applies the coupling invV K invV^T as basis changes of the moments
(invV^T before, invV after) around the nodal coupling kernels, and the
transfers as dense q x q matrices; the oracle follows FMCA's order (coupling
matrices contracted with invV per block).

Bar: FP64 GPU action within 1e-12 relative of the oracle on the same tree;
FP32 within 1e-5 of the FP64 GPU action.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err, uniform_points

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12

# (label, n, m or None (= same points), dim, deg, kernel, l, eta, nrhs)
CASES = [
    ("td-2d-exp-1rhs", 10000, None, 2, 4, "EXPONENTIAL", 1.0, 0.8, 1),
    ("td-3d-matern32-32rhs", 6000, None, 3, 4, "MATERN32", 0.5, 0.8, 32),
    ("td-3d-deg6-gauss-8rhs", 5000, None, 3, 6, "GAUSSIAN", 0.3, 0.8, 8),
    ("td-rect-3d-deg3-5rhs", 3000, 4000, 3, 3, "MATERN52", 0.4, 0.8, 5),
    ("td-1d-deg5-3rhs", 4000, None, 1, 5, "EXPONENTIAL", 0.3, 0.8, 3),
    ("td-3d-deg9-q220-2rhs", 4000, None, 3, 9, "EXPONENTIAL", 0.5, 0.8, 2),
    ("td-2d-65rhs", 5000, None, 2, 5, "MATERN32", 0.3, 0.8, 65),
    ("td-invmq-eta04-16rhs", 3000, 2000, 2, 4, "INVMULTIQUADRIC", 0.7, 0.4, 16),
    ("td-3d-deg7-32rhs-valu-fallback", 4000, None, 3, 7, "MATERN32", 0.4, 0.8, 32),
]


def _setup(case):
    import fmca_amd
    label, n, m, d, deg, ker, l, eta, nrhs = case
    print(f"[case] {label}", flush=True)
    P = uniform_points(n, d, seed=21)
    Peval = P if m is None else uniform_points(m, d, seed=22)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ker, l), Peval, P, deg, eta,
                               interpolator="total_degree")
    A = oracle.OracleH2(P, Peval, ker, l, deg, eta, scheme="total_degree")
    x = np.asfortranarray(np.random.default_rng(23).standard_normal((n, nrhs)))
    return H, A, x


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_total_degree_fp64_action_matches_oracle(case):
    H, A, x = _setup(case)
    assert H.interpolation()["q"] == A.q
    err = rel_err(H.action(x), A.action(x))
    print(f"  fp64 vs oracle rel err {err:.3e}", flush=True)
    assert err <= FP64_TOL


@pytest.mark.parametrize("case", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_total_degree_cluster_order_matches_oracle(case):
    H, A, x = _setup(case)
    assert rel_err(H.action_cluster_order(x), A.action_cluster_order(x)) <= FP64_TOL


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_total_degree_fp32_close_to_fp64(case):
    import fmca_amd
    H, _, x = _setup(case)
    err = rel_err(H.action(x, precision=fmca_amd.FP32), H.action(x))
    print(f"  fp32 vs fp64 rel err {err:.3e}", flush=True)
    assert err <= 1e-5


def test_total_degree_fmca_actor_test_configuration():
    """tests/test_H2kernelActor.cpp: 10000 points in [-1,1]^2, EXPONENTIAL
    l = 1, degree 4, unit-vector columns, eta = 0.8, 0.4, 0.2, 0.1: the GPU
    matches the oracle and the error against the exact column falls with eta."""
    import fmca_amd
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(5)
    n = 10000
    P = rng.uniform(-1, 1, (2, n))
    idx = rng.integers(0, n, 10)
    X = np.zeros((n, 10), order="F")
    X[idx, np.arange(10)] = 1
    C = np.exp(-cdist(P.T, P[:, idx].T))
    errs = []
    for eta in (0.8, 0.4, 0.2, 0.1):
        print(f"[case] actor test eta {eta}", flush=True)
        H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 1.0), P, P, 4, eta,
                                   interpolator="total_degree")
        y = H.action(X)
        yo = oracle.OracleH2(P, P, "EXPONENTIAL", 1.0, 4, eta, scheme="total_degree").action(X)
        assert rel_err(y, yo) <= FP64_TOL
        errs.append(rel_err(y, C))
    print(f"  errors vs exact columns {errs}", flush=True)
    assert errs[0] < 1e-4 and all(b < a for a, b in zip(errs, errs[1:]))


def test_h2matrix_binding_defaults_to_total_degree():
    """fmca_amd.H2Matrix mirrors FMCA's binding, whose Moments use
    TotalDegreeInterpolator (py/FMCA.cpp:37-38)."""
    import fmca_amd
    P = uniform_points(4000, 3, seed=31)
    x = np.random.default_rng(32).standard_normal(4000)
    H = fmca_amd.H2Matrix(fmca_amd.CovarianceKernel("GAUSSIAN", 0.4), P, P, 3, 0.8)
    assert H.interpolation()["interpolator"] == "total_degree"
    ref = oracle.OracleH2(P, P, "GAUSSIAN", 0.4, 3, 0.8, scheme="total_degree").action(x)
    assert rel_err(H.action(x).ravel(), ref.ravel()) <= FP64_TOL


def test_total_degree_transposed_action_matches_oracle():
    import torch
    import fmca_amd
    P = uniform_points(4000, 3, seed=5)
    Pe = uniform_points(2500, 3, seed=6)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN52", 0.4), Pe, P, 4, 0.8,
                               interpolator="total_degree")
    g = np.random.default_rng(7).standard_normal((2500, 5))
    got = H.action_transpose_device(torch.from_numpy(g).cuda()).cpu().numpy()
    ref = oracle.OracleH2(Pe, P, "MATERN52", 0.4, 4, 0.8, scheme="total_degree").action(g)
    assert rel_err(got, ref) <= FP64_TOL


def test_total_degree_gradients_vs_dense_autograd():
    """d/dalpha, d/dl and d/dpoints of 0.5 |K alpha - y|^2 through the
    total-degree H2 matvec against dense FP64 autograd."""
    import torch
    from fmca_amd.autograd import H2Spec, h2_matvec
    from fmca_amd import workloads as W
    print("[case] total-degree autograd", flush=True)
    n, l = 4096, 0.3
    P = torch.from_numpy(W.uniform(n, 2, 0)).cuda()
    alpha = torch.from_numpy(W.normal_rhs(n, 4, 1)).cuda()
    y = torch.from_numpy(W.normal_rhs(n, 4, 2)).cuda()
    spec = H2Spec("MATERN32", 6, 0.8, interpolator="total_degree")
    Pg, ag = P.clone().requires_grad_(True), alpha.clone().requires_grad_(True)
    lt = torch.tensor(l, dtype=torch.float64, requires_grad=True)
    (0.5 * ((h2_matvec(ag, lt, Pg, spec) - y) ** 2).sum()).backward()
    Pd, ad = P.clone().requires_grad_(True), alpha.clone().requires_grad_(True)
    ld = torch.tensor(l, dtype=torch.float64, device="cuda", requires_grad=True)
    diff = Pd[:, :, None] - Pd[:, None, :]
    r2 = (diff ** 2).sum(0)
    zero = r2 == 0  # sqrt has no derivative at 0: the kernel's is 0 there
    r = torch.where(zero, torch.zeros_like(r2),
                    torch.sqrt(torch.where(zero, torch.ones_like(r2), r2)))
    a = np.sqrt(3.0) * r / ld
    K = (1 + a) * torch.exp(-a)
    (0.5 * ((K @ ad - y) ** 2).sum()).backward()
    with torch.no_grad():
        efwd = rel_err(spec.actor(P, None, l).action_device(alpha).cpu().numpy(),
                       (K @ alpha).cpu().numpy())
    ea = rel_err(ag.grad.cpu().numpy(), ad.grad.cpu().numpy())
    el = abs(lt.grad.item() - ld.grad.item()) / abs(ld.grad.item())
    eP = rel_err(Pg.grad.cpu().numpy(), Pd.grad.cpu().numpy())
    print(f"  forward H2 err {efwd:.2e}; rel err alpha {ea:.2e} l {el:.2e} points {eP:.2e}",
          flush=True)
    assert efwd < 1e-5
    bound = 20 * efwd
    assert ea < bound and el < bound and eP < bound


def test_multipole_function_evaluator():
    """MultipoleFunctionEvaluator (MultipoleFunctionEvaluator.h:17-67):
    evaluate(P, Peval, C) = the total-degree H2 action K(Peval, P) C.  At the
    default eta = 0 no block is admissible, so the result is the dense
    product; at eta = 0.8 it is the oracle's H2 action."""
    import fmca_amd
    print("[case] multipole function evaluator", flush=True)
    P = uniform_points(3000, 2, seed=41)
    Pe = uniform_points(1500, 2, seed=42)
    C = np.random.default_rng(43).standard_normal((3000, 3))
    K = fmca_amd.CovarianceKernel("MATERN52", 0.3)
    ev = fmca_amd.MultipoleFunctionEvaluator(K, P, Pe)
    dense = oracle.dense_action(P, Pe, "MATERN52", 0.3, C)
    assert rel_err(ev.evaluate(P, Pe, C), dense) <= 1e-13
    ev = fmca_amd.MultipoleFunctionEvaluator(K, P, Pe, 0.8, 4)
    ref = oracle.OracleH2(P, Pe, "MATERN52", 0.3, 4, 0.8, scheme="total_degree").action(C)
    assert rel_err(ev.evaluate(P, Pe, C), ref) <= FP64_TOL
    # other evaluation points: planned anew
    Pe2 = uniform_points(1000, 2, seed=44)
    ref2 = oracle.OracleH2(P, Pe2, "MATERN52", 0.3, 4, 0.8, scheme="total_degree").action(C)
    assert rel_err(ev.evaluate(P, Pe2, C), ref2) <= FP64_TOL


@pytest.mark.parametrize("transpose", [False, True])
def test_total_degree_variant_groups_equal_separate_actions(transpose):
    """The fused-backward entry (16-column groups with different kernel
    variants in one action) on the total-degree form: each group equals the
    single-variant action of its columns."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] total-degree variant groups T={transpose}", flush=True)
    n, m = 6000, 5000
    P, Pe = W.uniform(n, 3, 31), W.uniform(m, 3, 32)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.4), Pe, P, 4, 0.8,
                               interpolator="total_degree")
    groups = [2, 0, 1, 2]
    rows_in, rows_out = (m, n) if transpose else (n, m)
    ncols = 16 * (len(groups) - 1) + 9
    X = torch.randn(ncols, rows_in, dtype=torch.float64, device="cuda").t()
    stream = torch.cuda.current_stream().cuda_stream
    for prec, tol in ((fmca_amd.FP64, 1e-13), (fmca_amd.FP32, 1e-6)):
        Y = torch.empty(ncols, rows_out, dtype=torch.float64, device="cuda").t()
        H._a.action_variants_device(groups, transpose, X.data_ptr(), rows_in, ncols,
                                    Y.data_ptr(), rows_out, prec, stream)
        for g, v in enumerate(groups):
            c0, c1 = 16 * g, min(16 * g + 16, ncols)
            Xg = X[:, c0:c1].t().contiguous().t()
            Yg = torch.empty(c1 - c0, rows_out, dtype=torch.float64, device="cuda").t()
            H._a.action_variant_device(v, transpose, Xg.data_ptr(), rows_in, c1 - c0,
                                       Yg.data_ptr(), rows_out, prec, stream)
            err = (torch.linalg.norm(Y[:, c0:c1] - Yg) / torch.linalg.norm(Yg)).item()
            assert err <= tol, (prec, g, v, err)


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_total_degree_cg_reaches_1e8(precision):
    """pCG (PCG.h:15-44) on (K + ridge I) with the total-degree form: the FP64
    residual, recomputed through the FP64 action, is <= 1e-8 for every column."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] total-degree cg {precision}", flush=True)
    n, ridge = 20000, 1.0  # well above the H2 error of the deg-4 total-degree form
    P = W.uniform(n, 3, 0)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.3), P, P, 4, 0.8,
                               interpolator="total_degree")
    b = torch.from_numpy(W.normal_rhs(n, 4)).cuda()
    x, info = H.cg(b, x0=torch.zeros_like(b), ridge=ridge, tol=1e-8,
                   precision=getattr(fmca_amd, precision), max_iter=4000)
    print(" ", info, flush=True)
    assert info["converged"]
    r = b - H.action_device(x) - ridge * x
    assert (torch.linalg.norm(r, dim=0) / torch.linalg.norm(b, dim=0)).max().item() <= 1e-8
