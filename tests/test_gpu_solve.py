"""CG solves of (K + ridge I) x = b (FMCA's pCG, PCG.h:15-44) on the GPU.

Bar: the FP64 residual b - (K + ridge I) x, recomputed independently through
the FP64 action, is <= 1e-8 relative for every column -- for plain FP64 CG and
for CG with FP32 matvecs and FP64 reliable updates."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _resid(H, x, b, ridge):
    import torch
    r = b - H.action_device(x) - ridge * x
    return (torch.linalg.norm(r, dim=0) / torch.linalg.norm(b, dim=0)).max().item()


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_cg_uniform_gaussian_reaches_1e8(precision):
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] cg uniform n=20000 {precision}", flush=True)
    n, ridge = 20000, 1e-2
    P = W.uniform(n, 3, 0)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("GAUSSIAN", 0.2), P, P, 4, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, 4)).cuda()
    x, info = H.cg(b, x0=torch.zeros_like(b), ridge=ridge, tol=1e-8,
                   precision=getattr(fmca_amd, precision), max_iter=8000)
    print(" ", info, flush=True)
    assert info["converged"]
    assert _resid(H, x, b, ridge) <= 1e-8


def test_cg_fp32_and_fp64_solutions_agree():
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n, ridge = 8000, 1e-1
    P = W.uniform(n, 2, 3)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.3), P, P, 5, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, 3, 4)).cuda()
    x64, i64 = H.cg(b, ridge=ridge, tol=1e-10, precision=fmca_amd.FP64)
    x32, i32 = H.cg(b, ridge=ridge, tol=1e-10, precision=fmca_amd.FP32)
    print(" ", i64, i32, flush=True)
    assert i64["converged"] and i32["converged"] and i32["matvecs_fp32"] > 0
    assert rel_err(x32.cpu().numpy(), x64.cpu().numpy()) < 1e-8


def test_cg_matches_dense_solve_small():
    """Against numpy's dense solve of the oracle's (exactly assembled) H2 matrix."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n, ridge = 3000, 1e-2
    P = W.uniform(n, 3, 5)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 0.5), P, P, 3, 0.8)
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, 3, 0.8)
    Kh = A.action(np.eye(n))
    bh = W.normal_rhs(n, 2, 6)
    xd = np.linalg.solve(Kh + ridge * np.eye(n), bh)
    x, info = H.cg(torch.from_numpy(bh).cuda(), ridge=ridge, tol=1e-10,
                   precision=fmca_amd.FP32)
    assert info["converged"]
    assert rel_err(x.cpu().numpy(), xd) < 1e-7


def test_cg_single_column_and_zero_rhs():
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n = 4000
    P = W.uniform(n, 3, 7)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN52", 0.3), P, P, 3, 0.8)
    b = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    b[:, 1] = torch.from_numpy(W.normal_rhs(n, 1, 8)[:, 0]).cuda()
    x, info = H.cg(b, x0=torch.zeros_like(b), ridge=1.0, tol=1e-9, precision=fmca_amd.FP32)
    assert info["converged"] and torch.all(x[:, 0] == 0)
    x1, info1 = H.cg(b[:, 1].contiguous(), ridge=1.0, tol=1e-9, precision=fmca_amd.FP64)
    assert info1["converged"] and x1.dim() == 1
    assert _resid(H, x1.reshape(-1, 1), b[:, 1:].contiguous(), 1.0) <= 1e-9


def test_cg_reports_indefinite_system():
    """The multiquadric kernel sqrt(r^2 + c^2) gives a matrix with one positive
    eigenvalue and the rest negative (conditionally negative definite): CG
    must stop with the indefinite flag instead of iterating on."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print("[case] cg indefinite multiquadric", flush=True)
    n = 3000
    P = W.uniform(n, 2, 13)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MULTIQUADRIC", 1.0, 0.5), P, P, 4,
                               0.8)
    b = torch.from_numpy(W.normal_rhs(n, 2, 14)).cuda()
    for prec in (fmca_amd.FP64, fmca_amd.FP32):
        x, info = H.cg(b, ridge=0.0, tol=1e-8, precision=prec, max_iter=500)
        assert info["indefinite"] and not info["converged"], info
        assert info["iterations"] < 500


def test_cg_with_bf16_matvecs_reaches_fp64_solution():
    """CG with reliable updates whose recurrence runs on the BF16 three-term
    matvecs reaches the FP64 residual tolerance and the FP64 CG solution."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print("[case] cg fp32-bf16 n=20000 32 rhs", flush=True)
    n, nrhs = 20000, 32
    P = W.uniform(n, 3, 3)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 3, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, nrhs, 4)).cuda()
    x64, i64 = H.cg(b, x0=torch.zeros_like(b), ridge=20.0, tol=1e-10, precision=fmca_amd.FP64)
    xb, ib = H.cg(b, x0=torch.zeros_like(b), ridge=20.0, tol=1e-10, precision=fmca_amd.FP32_BF16)
    print(f"  iterations fp64 {i64['iterations']} bf16 {ib['iterations']} "
          f"(fp64 matvecs {ib['matvecs_fp64']}), residual {ib['rel_residual']:.2e}", flush=True)
    assert ib["converged"] and ib["rel_residual"] <= 1e-10 and ib["matvecs_fp32"] > 0
    assert _resid(H, xb, b, 20.0) <= 1e-10
    assert (torch.linalg.norm(xb - x64) / torch.linalg.norm(x64)).item() <= 1e-8


def test_clustered_configuration_needs_a_ridge_above_the_h2_error():
    """The clustered configuration (16-Gaussian mixture, sigma 0.02, GAUSSIAN
    l = 0.1, smooth rhs + 1% noise, 64 rhs) at n = 3e4, degree 4, eta 0.8.
    At the configuration's ridge 1e-4 the H2 matrix plus ridge is indefinite --
    its approximation error in the spectral norm exceeds the ridge -- and CG
    reports that instead of iterating (p.(K~ + lambda I)p <= 0).  At ridge 0.1
    (the smallest of 1e-3, 1e-2, 1e-1 at which FP64 CG converges there) both
    precision forms with FP64 reliable updates reach the 1e-8 FP64 residual
    and FP64 CG's solution."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print("[case] clustered cg ridges", flush=True)
    n, nrhs = 30000, 64
    P = W.gaussian_mixture(n, 3, 16, 0.02, 0)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("GAUSSIAN", 0.1), P, P, 4, 0.8)
    y = torch.from_numpy(W.smooth_rhs(P, nrhs)).cuda()
    _, info = H.cg(y, x0=torch.zeros_like(y), ridge=1e-4, tol=1e-8, precision=fmca_amd.FP64,
                   max_iter=1000)
    print(" ridge 1e-4 fp64", info, flush=True)
    assert info["indefinite"] and not info["converged"]
    x64, i64 = H.cg(y, x0=torch.zeros_like(y), ridge=0.1, tol=1e-8, precision=fmca_amd.FP64,
                    max_iter=3000)
    assert i64["converged"] and i64["rel_residual"] <= 1e-8
    for prec in (fmca_amd.FP32, fmca_amd.FP32_BF16):
        x, info = H.cg(y, x0=torch.zeros_like(y), ridge=0.1, tol=1e-8, precision=prec,
                       max_iter=3000)
        print(f" ridge 0.1 precision {prec}", info, flush=True)
        assert info["converged"] and _resid(H, x, y, 0.1) <= 1e-8
        # both residuals <= 1e-8 bound the difference by about cond * 2e-8
        # (cond ~ lambda_max / 0.1 ~ 2e4 here)
        diff = (torch.linalg.norm(x - x64) / torch.linalg.norm(x64)).item()
        print(f"  solution difference {diff:.2e}", flush=True)
        assert diff <= 1e-4


@pytest.mark.parametrize("precision", ["FP64", "FP32", "FP32_BF16"])
def test_gcr_reaches_the_fp64_cg_solution(precision):
    """GCR with each matvec precision reaches the 1e-10 FP64 residual, checked
    again through the FP64 action, and FP64 CG's solution.  With FP32-class
    matvecs it needs about as many iterations as with FP64 ones: every stored
    pair (p, (K + ridge I) p) is consistent as its own matvec computed it,
    whereas CG's recurrence loses conjugacy under the changing rounding."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] gcr n=20000 8 rhs {precision}", flush=True)
    n, nrhs, ridge = 20000, 8, 20.0
    P = W.uniform(n, 3, 3)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 3, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, nrhs, 4)).cuda()
    z = torch.zeros_like(b)
    x64, i64 = H.cg(b, x0=z, ridge=ridge, tol=1e-10, precision=fmca_amd.FP64)
    _, g64 = H.cg(b, x0=z, ridge=ridge, tol=1e-10, precision=fmca_amd.FP64, method="gcr")
    x, info = H.cg(b, x0=z, ridge=ridge, tol=1e-10, precision=getattr(fmca_amd, precision),
                   method="gcr")
    print(f"  cg fp64 {i64['iterations']} it, gcr fp64 {g64['iterations']} it, "
          f"gcr {precision} {info}", flush=True)
    assert info["converged"] and info["rel_residual"] <= 1e-10
    assert _resid(H, x, b, ridge) <= 1e-10
    assert (torch.linalg.norm(x - x64) / torch.linalg.norm(x64)).item() <= 1e-8
    if precision != "FP64":
        assert info["matvecs_fp32"] == info["iterations"] > 0
        assert info["iterations"] <= 1.25 * g64["iterations"] + 2


def test_gcr_window_smaller_than_the_iteration_count():
    """With room for 3 and 8 pairs the window wraps (the oldest pair is
    dropped) and the solve still reaches tol; an oversized window is refused."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print("[case] gcr small windows", flush=True)
    n, ridge = 8000, 1e-1
    P = W.uniform(n, 2, 3)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.3), P, P, 5, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, 3, 4)).cuda()
    x64, i64 = H.cg(b, ridge=ridge, tol=1e-10, precision=fmca_amd.FP64)
    for m in (3, 8):
        x, info = H.cg(b, ridge=ridge, tol=1e-10, precision=fmca_amd.FP32, method="gcr",
                       max_basis=m, max_iter=5000)
        print(f"  window {m}: {info}", flush=True)
        assert info["converged"] and info["iterations"] > m
        assert _resid(H, x, b, ridge) <= 1e-10
        assert rel_err(x.cpu().numpy(), x64.cpu().numpy()) < 1e-8
    with pytest.raises(ValueError):
        H.cg(b, ridge=ridge, precision=fmca_amd.FP32, method="gcr", max_basis=257)
    with pytest.raises(ValueError):
        H.cg(b, ridge=ridge, precision=fmca_amd.FP32, method="minres")


def test_gcr_single_column_and_zero_rhs():
    """A zero column stays zero and stops at once; a 1-D rhs returns 1-D."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n = 4000
    P = W.uniform(n, 3, 7)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN52", 0.3), P, P, 3, 0.8)
    b = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    b[:, 1] = torch.from_numpy(W.normal_rhs(n, 1, 8)[:, 0]).cuda()
    x, info = H.cg(b, x0=torch.zeros_like(b), ridge=1.0, tol=1e-9,
                   precision=fmca_amd.FP32_BF16, method="gcr")
    assert info["converged"] and torch.all(x[:, 0] == 0)
    assert _resid(H, x[:, 1:].contiguous(), b[:, 1:].contiguous(), 1.0) <= 1e-9
    x1, info1 = H.cg(b[:, 1].contiguous(), ridge=1.0, tol=1e-9, precision=fmca_amd.FP64,
                     method="gcr")
    assert info1["converged"] and x1.dim() == 1
    assert _resid(H, x1.reshape(-1, 1), b[:, 1:].contiguous(), 1.0) <= 1e-9


@pytest.mark.parametrize("method", ["cg", "gcr"])
def test_solve_stopped_by_a_cap_reports_the_residual_of_its_x(method):
    """A solve cut off by max_iter is not converged and reports the FP64
    residual of the x it returns, not the last replaced one.  (GCR's residual
    norm never grows; CG's may, as CG minimises the error in the energy norm.)"""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    n, ridge = 6000, 1e-2
    P = W.uniform(n, 3, 9)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.4), P, P, 3, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, 2, 10)).cuda()
    x, info = H.cg(b, x0=torch.zeros_like(b), ridge=ridge, tol=1e-12,
                   precision=fmca_amd.FP32, method=method, max_iter=4)
    print(f"  {method}: {info}", flush=True)
    assert not info["converged"] and info["iterations"] == 4
    r = _resid(H, x, b, ridge)
    assert abs(info["rel_residual"] - r) <= 1e-6 * r
    if method == "gcr":
        assert r < 1
