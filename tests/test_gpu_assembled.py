"""The assembled H2 matrix (H2Matrix::computeH2Matrix + operator*): every
coupling matrix, dense block and leaf basis stored once in device memory and
streamed by later actions.  Its action against the matrix-free action of the
same actor and against the oracle.

Bar: FP64 within 1e-13 of the matrix-free FP64 action (the same entries,
summed in another order) and within 1e-12 of the oracle; FP32 within 1e-5 of
FP64, as the matrix-free FP32 form.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [
    # label, n, m (None: same), d, deg, interpolator, kernel
    ("tensor-3d", 20_000, None, 3, 3, "tensor_product", "MATERN32"),
    ("total-degree-3d", 30_000, None, 3, 4, "total_degree", "EXPONENTIAL"),
    ("rect-2d", 15_000, 12_000, 2, 5, "tensor_product", "GAUSSIAN"),
    ("line-1d", 5_000, None, 1, 6, "tensor_product", "MATERN52"),
]


def _actor(case, planner="auto"):
    import fmca_amd
    label, n, m, d, deg, interp, ker = case
    rng = np.random.default_rng(7)
    P = np.asfortranarray(rng.random((d, n)))
    Pe = P if m is None else np.asfortranarray(rng.random((d, m)))
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ker, 0.4), Pe, P, deg, 0.8, interp,
                               planner=planner)
    return H, P, Pe


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_assembled_action_equals_matrix_free(case):
    import fmca_amd
    print(f"[case] assembled {case[0]}", flush=True)
    H, P, Pe = _actor(case)
    n = P.shape[1]
    rng = np.random.default_rng(3)
    for nrhs in (1, 8, 17, 33, 70):
        x = np.asfortranarray(rng.standard_normal((n, nrhs)))
        ref64 = H.action(x)
        ref64c = H.action_cluster_order(x)
        nb = H.assemble(fmca_amd.FP64)
        assert nb > 0
        y = H.action(x)
        assert rel_err(y, ref64) <= 1e-13, nrhs
        assert rel_err(H.action_cluster_order(x), ref64c) <= 1e-13, nrhs
        H.assemble(fmca_amd.FP32)
        y32 = H.action(x, precision=fmca_amd.FP32)
        assert rel_err(y32, ref64) <= 1e-5, nrhs
        # FP64 actions generate entries again while FP32 blocks are held
        assert rel_err(H.action(x), ref64) <= 1e-14
        H.release_assembled()
        assert rel_err(H.action(x), ref64) <= 1e-14


def test_assembled_action_matches_oracle_and_sizes():
    import fmca_amd
    print("[case] assembled vs oracle", flush=True)
    case = CASES[0]
    H, P, Pe = _actor(case)
    A = oracle.OracleH2(P, Pe, case[6], 0.4, case[4], 0.8)
    x = np.asfortranarray(np.random.default_rng(5).standard_normal((P.shape[1], 4)))
    nb = H.assemble(fmca_amd.FP64)
    assert rel_err(H.action(x), A.action(x)) <= 1e-12
    info = H.info()
    # 8 B per (padded) entry: at least the entries, at most 1.6 x with padding
    entries = info["near_entries"] + info["far_entries"]
    assert 8 * entries <= nb <= 1.6 * 8 * (entries + info["n"] * info["q"])
    assert 4 * entries <= H.assemble(fmca_amd.FP32) < nb


def test_assembled_blocks_follow_the_kernel_and_skip_other_forms():
    """A new lengthscale frees the blocks (the actions follow the new kernel);
    transposed and variant actions never use them."""
    import torch
    import fmca_amd
    print("[case] assembled set_kernel / transpose", flush=True)
    case = ("rect", 8_000, 6_000, 3, 3, "tensor_product", "MATERN32")
    H, P, Pe = _actor(case)
    x = np.asfortranarray(np.random.default_rng(9).standard_normal((P.shape[1], 16)))
    H.assemble(fmca_amd.FP64)
    H.set_kernel(fmca_amd.CovarianceKernel("MATERN32", 0.25))
    A = oracle.OracleH2(P, Pe, "MATERN32", 0.25, 3, 0.8)
    assert rel_err(H.action(x), A.action(x)) <= 1e-12
    H.assemble(fmca_amd.FP64)
    g = np.random.default_rng(2).standard_normal((Pe.shape[1], 16))
    yt = H.action_transpose_device(torch.from_numpy(g).cuda()).cpu().numpy()
    ref = oracle.OracleH2(Pe, P, "MATERN32", 0.25, 3, 0.8).action(g)  # K(P, Pe) = K(Pe, P)^T
    assert rel_err(yt, ref) <= 1e-12


def test_assembled_needs_q_at_most_128():
    import fmca_amd
    P = np.asfortranarray(np.random.default_rng(1).random((3, 3000)))
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.4), P, P, 5, 0.8)
    with pytest.raises(ValueError, match="q <= 128"):
        H.assemble(fmca_amd.FP64)


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_cg_on_assembled_blocks(precision):
    """CG's matvecs of the assembled precision stream the stored blocks (the
    FP32 form's FP64 residual replacements generate entries); the solution
    satisfies the FP64 residual bar and equals the matrix-free solve."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] cg assembled {precision}", flush=True)
    n, ridge = 8000, 1e-1
    P = W.uniform(n, 2, 3)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.3), P, P, 5, 0.8)
    b = torch.from_numpy(W.normal_rhs(n, 3, 4)).cuda()
    prec = getattr(fmca_amd, precision)
    x0, i0 = H.cg(b, ridge=ridge, tol=1e-10, precision=prec)
    H.assemble(prec)
    x1, i1 = H.cg(b, ridge=ridge, tol=1e-10, precision=prec)
    print(" ", i0, i1, flush=True)
    assert i0["converged"] and i1["converged"]
    r = b - H.action_device(x1) - ridge * x1
    assert (torch.linalg.norm(r, dim=0) / torch.linalg.norm(b, dim=0)).max().item() <= 1e-10
    assert rel_err(x1.cpu().numpy(), x0.cpu().numpy()) < 1e-8
