"""GPU parity of the H2 action against the FP64 CPU oracle (FMCA restatement).

Bar: FP64 GPU action within 1e-12 relative of the oracle's action on the same
tree (identical blocks; only summation order differs).  The FP32 precision
form is checked against the FP64 GPU action with a tolerance stated per test.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err, uniform_points

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12

# (label, n, m or None (= same points), dim, deg, kernel, l, eta, nrhs)
CASES = [
    ("cfg1-exp-3d", 10000, None, 3, 3, "EXPONENTIAL", 0.5, 0.8, 1),
    ("matern32-3d-32rhs", 5000, None, 3, 4, "MATERN32", 0.5, 0.8, 32),
    ("matern32-2d-deg5", 8000, None, 2, 5, "MATERN32", 0.3, 0.8, 8),
    ("gauss-rect", 3000, 5000, 3, 3, "GAUSSIAN", 0.2, 0.8, 3),
    ("matern52-1d", 4000, None, 1, 6, "MATERN52", 0.4, 0.5, 2),
    ("invmq-2d", 3000, 2000, 2, 4, "INVMULTIQUADRIC", 0.7, 0.6, 5),
    ("matern72-3d", 3000, None, 3, 2, "MATERN72", 0.6, 0.8, 1),
    ("matern92-eta04", 3000, None, 2, 3, "MATERN92", 0.6, 0.4, 1),
    ("small-all-dense", 100, None, 3, 4, "EXPONENTIAL", 1.0, 0.8, 2),
    # the kernels that take the run-time switch (CovarianceKernel.h:56-61,103-116),
    # on the VALU (< 8 rhs) and matrix-core (>= 8 rhs) forms
    ("multiquadric-3d-valu", 3000, None, 3, 3, "MULTIQUADRIC", 0.6, 0.8, 3),
    ("multiquadric-rect-mfma", 2500, 3000, 3, 3, "MULTIQUADRIC", 0.6, 0.8, 16),
    ("biharmonic2d-valu", 4000, None, 2, 4, "BIHARMONIC2D", 0.5, 0.8, 1),
    ("biharmonic2d-mfma", 4000, None, 2, 4, "BIHARMONIC2D", 0.5, 0.8, 24),
    ("biharmonic3d-valu", 3000, 2500, 3, 3, "BIHARMONIC3D", 0.5, 0.8, 2),
    ("biharmonic3d-mfma", 3000, None, 3, 3, "BIHARMONIC3D", 0.5, 0.8, 8),
    ("triharmonic3d-valu", 3000, None, 3, 3, "TRIHARMONIC3D", 0.7, 0.8, 1),
    ("triharmonic3d-mfma", 3000, 2000, 3, 3, "TRIHARMONIC3D", 0.7, 0.8, 32),
    ("tps2d-valu", 4000, None, 2, 4, "TPS2D", 1.0, 0.8, 5),
    ("tps2d-mfma", 4000, 3500, 2, 4, "TPS2D", 1.0, 0.8, 16),
    ("rhs-40", 3000, None, 3, 3, "MATERN32", 0.5, 0.8, 40),
    # high degrees in 3-D (q = 512, 729, 1000; leaves of up to 2q - 1 points),
    # eta 1.5 so that n = 3e4 still has far blocks at q = 1000
    ("q512-deg7-1rhs", 30000, None, 3, 7, "GAUSSIAN", 0.3, 1.5, 1),
    ("q512-deg7-12rhs", 30000, None, 3, 7, "GAUSSIAN", 0.3, 1.5, 12),
    ("q729-deg8-3rhs", 30000, None, 3, 8, "MATERN32", 0.4, 1.5, 3),
    ("q1000-deg9-1rhs", 30000, None, 3, 9, "GAUSSIAN", 0.3, 1.5, 1),
    ("q1000-deg9-16rhs", 30000, None, 3, 9, "GAUSSIAN", 0.3, 1.5, 16),
    ("q216-16rhs-two-row-passes", 4000, None, 3, 5, "EXPONENTIAL", 0.5, 0.8, 16),
    ("q343-9rhs-valu-coupling", 3000, None, 3, 6, "GAUSSIAN", 0.3, 0.8, 9),
    ("1d-16rhs", 5000, None, 1, 4, "MATERN32", 0.2, 0.8, 16),
    ("rect-33rhs", 4000, 3000, 3, 3, "MATERN52", 0.4, 0.8, 33),
    ("2d-65rhs", 6000, None, 2, 4, "EXPONENTIAL", 0.3, 0.8, 65),
    # FP64 near field with squared distances from the matrix cores (16-column
    # and 32-column passes, partial last column tile)
    ("gauss-3d-12rhs-gram", 6000, None, 3, 3, "GAUSSIAN", 0.3, 0.8, 12),
    ("matern52-2d-20rhs-gram", 5000, 4000, 2, 4, "MATERN52", 0.3, 0.8, 20),
]


def _setup(case):
    import fmca_amd
    label, n, m, d, deg, ker, l, eta, nrhs = case
    print(f"[case] {label}", flush=True)
    P = uniform_points(n, d, seed=11)
    Peval = P if m is None else uniform_points(m, d, seed=12)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ker, l), Peval, P, deg, eta)
    A = oracle.OracleH2(P, Peval, ker, l, deg, eta)
    x = np.asfortranarray(np.random.default_rng(13).standard_normal((n, nrhs)))
    return H, A, x, P, Peval


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_action_matches_oracle(case):
    H, A, x, _, _ = _setup(case)
    y = H.action(x)
    yo = A.action(x)
    err = rel_err(y, yo)
    print(f"  fp64 vs oracle rel err {err:.3e}", flush=True)
    assert err <= FP64_TOL


@pytest.mark.parametrize("case", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_fp64_cluster_order_matches_oracle(case):
    H, A, x, _, _ = _setup(case)
    y = H.action_cluster_order(x)
    yo = A.action_cluster_order(x)
    assert rel_err(y, yo) <= FP64_TOL


FP32_CASES = CASES[:4] + CASES[-5:]


@pytest.mark.parametrize("case", FP32_CASES, ids=[c[0] for c in FP32_CASES])
def test_fp32_action_close_to_fp64(case):
    import fmca_amd
    H, A, x, _, _ = _setup(case)
    y64 = H.action(x)
    y32 = H.action(x, precision=fmca_amd.FP32)
    err = rel_err(y32, y64)
    print(f"  fp32 vs fp64 rel err {err:.3e}", flush=True)
    assert err <= 1e-5


def test_fp64_action_vs_dense_cfg1():
    """Config 1: H2 error against dense FP64 K*x (states the approximation error)."""
    import fmca_amd
    P = uniform_points(10000, 3, seed=1)
    x = np.random.default_rng(2).uniform(-1, 1, (10000, 1))
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 0.5), P, P, 3, 0.8)
    y = H.action(x)
    yd = oracle.dense_action(P, P, "EXPONENTIAL", 0.5, x)
    err = rel_err(y, yd)
    print(f"  H2 vs dense rel err {err:.3e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("dim,deg,kernel,nrhs", [(3, 3, "GAUSSIAN", 1), (3, 3, "GAUSSIAN", 32),
                                                 (2, 4, "MATERN32", 8), (3, 2, "EXPONENTIAL", 64)])
def test_clustered_points_with_split_targets(dim, deg, kernel, nrhs):
    """Mixture-of-Gaussians points give targets with many times the mean work;
    the planner splits them over several workgroups that add into a zeroed
    output.  The FP64 action still matches the oracle within 1e-12 and the
    FP32 action the FP64 one within 1e-5, and splits do occur."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] clustered d={dim} deg={deg} {kernel} nrhs={nrhs}", flush=True)
    n = 30000
    P = W.gaussian_mixture(n, dim, 16, 0.02, 41)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(kernel, 0.1), P, P, deg, 0.8)
    splits = [(H.launch_order(w)[:, 3] & 1).any() for w in ("coupling", "leaf_output")]
    assert any(splits), splits
    x = np.asfortranarray(np.random.default_rng(42).standard_normal((n, nrhs)))
    ref = oracle.OracleH2(P, P, kernel, 0.1, deg, 0.8).action(x)
    xd = torch.from_numpy(x).cuda()
    y64 = H.action_device(xd, precision=fmca_amd.FP64).cpu().numpy()
    y32 = H.action_device(xd, precision=fmca_amd.FP32).cpu().numpy()
    assert rel_err(y64, ref) <= FP64_TOL
    assert rel_err(y32, y64) <= 1e-5


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_action_replays_in_a_hip_graph(precision):
    """The device action enqueues only kernels on the caller's stream (no
    memsets, no allocation once the workspaces are sized), so it can be
    captured in a graph.  Captured once, replayed three times with eager work
    between the replays (the input changed in place); every replay equals an
    eager action on the same input."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] graph replay {precision}", flush=True)
    prec = getattr(fmca_amd, precision)
    n, nrhs = 6000, 16
    P = W.uniform(n, 3, 51)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.4), P, P, 3, 0.8)
    x = torch.randn(nrhs, n, dtype=torch.float64, device="cuda").t()
    y = torch.empty(nrhs, n, dtype=torch.float64, device="cuda").t()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        H.action_device(x, out=y, precision=prec)  # sizes the workspaces
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        H.action_device(x, out=y, precision=prec)
    for k in range(3):
        x.mul_(-0.5).add_(0.25 * k)          # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        ref = H.action_device(x, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(y, ref) or (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item() <= 1e-15


BF16_TOL = 2e-5  # FP32 entries (~4e-6 here) plus the dropped 2^-16 terms of the split


# every case where the BF16 matrix-core forms run (>= 8 rhs), each kernel type
BF16_CASES = [c for c in CASES if c[8] >= 8]


@pytest.mark.parametrize("case", BF16_CASES, ids=[c[0] for c in BF16_CASES])
def test_fp32_bf16_action_close_to_fp64(case):
    """The FP32 form with BF16 three-term matrix-core contractions (and its
    FP32 fallbacks below 8 rhs) against the FP64 action."""
    import fmca_amd
    H, A, x, _, _ = _setup(case)
    y64 = H.action(x)
    yb = H.action(x, precision=fmca_amd.FP32_BF16)
    y32 = H.action(x, precision=fmca_amd.FP32)
    err, err32 = rel_err(yb, y64), rel_err(y32, y64)
    print(f"  fp32-bf16 vs fp64 rel err {err:.3e} (fp32 {err32:.3e})", flush=True)
    assert err <= BF16_TOL


@pytest.mark.parametrize("dim,kernel,nrhs", [(3, "GAUSSIAN", 32), (2, "MATERN32", 16),
                                             (3, "EXPONENTIAL", 64)])
def test_fp32_bf16_on_clustered_points_with_split_targets(dim, kernel, nrhs):
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] bf16 clustered d={dim} {kernel} nrhs={nrhs}", flush=True)
    n = 30000
    P = W.gaussian_mixture(n, dim, 16, 0.02, 43)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(kernel, 0.1), P, P, 3, 0.8)
    x = torch.from_numpy(np.asfortranarray(np.random.default_rng(44).standard_normal((n, nrhs)))).cuda()
    y64 = H.action_device(x, precision=fmca_amd.FP64).cpu().numpy()
    yb = H.action_device(x, precision=fmca_amd.FP32_BF16).cpu().numpy()
    err = rel_err(yb, y64)
    print(f"  fp32-bf16 vs fp64 rel err {err:.3e}", flush=True)
    assert err <= BF16_TOL
