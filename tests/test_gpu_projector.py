"""H2MatrixProjector (H2MatrixProjector.h:21-75) on the GPU actor: a dense
matrix projected onto the plan's H2 format, then applied by the device actions.

Pinned by what fixes an orthogonal projection onto the H2 space, not by a
restatement of it:
  - idempotence: the oracle's H2 matrix K~ (dense, from its action) is
    reproduced (its far blocks lie in the span of the cluster bases);
  - optimality: for the exact kernel matrix K, ||P(K) - K||_F <= ||K~ - K||_F
    (the interpolation K~ is another, oblique, projection onto the same space);
  - orthogonality of the residual on far leaf blocks: V_t (K - P(K))_ts V_s^T = 0;
  - FMCA's own check of the dual cluster bases, < 5e-10
    (compute_dual_cluster_bases_impl.h:66-86)."""
import numpy as np
import pytest

import oracle
from conftest import rel_err, uniform_points

pytestmark = pytest.mark.gpu


def _dense_of(H, n, precision):
    import torch
    I = torch.eye(n, dtype=torch.float64, device="cuda")
    return H.action_device(I, precision=precision).cpu().numpy()


@pytest.mark.parametrize("m,n,d,deg,kernel,l", [(None, 2500, 2, 3, "MATERN32", 0.3),
                                                 (2000, 2600, 3, 2, "EXPONENTIAL", 0.5),
                                                 (None, 3000, 1, 5, "GAUSSIAN", 0.2)])
def test_projection_is_idempotent_on_the_h2_matrix(m, n, d, deg, kernel, l):
    import fmca_amd
    print(f"[case] projector idempotent m={m} n={n} d={d} deg={deg}", flush=True)
    P = uniform_points(n, d, seed=61)
    Pe = P if m is None else uniform_points(m, d, seed=62)
    mm = n if m is None else m
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(kernel, l), Pe, P, deg, 0.8)
    Kt = oracle.OracleH2(P, Pe, kernel, l, deg, 0.8).action(np.eye(n))  # mm x n
    chk = H.project(Kt)
    assert max(chk) < 5e-10, chk
    x = np.asfortranarray(np.random.default_rng(63).standard_normal((n, 5)))
    assert rel_err(H.action(x), Kt @ x) <= 1e-10
    H.project(Kt, precision=fmca_amd.FP32)
    assert rel_err(H.action(x, precision=fmca_amd.FP32), Kt @ x) <= 1e-5
    H.release_assembled()
    assert rel_err(H.action(x), Kt @ x) <= 1e-12  # the kernel again


def test_projection_is_frobenius_optimal_and_orthogonal():
    import fmca_amd
    print("[case] projector optimality", flush=True)
    n, d, deg, kernel, l = 2500, 2, 2, "MATERN32", 0.3
    P = uniform_points(n, d, seed=64)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(kernel, l), P, P, deg, 0.8)
    A = oracle.OracleH2(P, P, kernel, l, deg, 0.8)
    K = oracle.dense_matrix(P, P, kernel, l)
    Kt = A.action(np.eye(n))
    chk = H.project(K)
    assert max(chk) < 5e-10, chk
    PK = _dense_of(H, n, fmca_amd.FP64)
    e_proj, e_interp = np.linalg.norm(PK - K), np.linalg.norm(Kt - K)
    print(f"  ||P(K) - K|| = {e_proj:.3e}, ||K~ - K|| = {e_interp:.3e}", flush=True)
    assert e_proj <= e_interp
    # V_t R V_s^T = 0 on far blocks between leaves
    rows, cols, lr = A.blocks()
    cl = A.clusters(0)
    perm = A.indices(0)
    R = K - PK
    checked = 0
    for t, s, low in zip(rows, cols, lr):
        if not low or cl["sons"][t][0] >= 0 or cl["sons"][s][0] >= 0:
            continue
        bt, nt = cl["begin"][t], cl["size"][t]
        bs, ns = cl["begin"][s], cl["size"][s]
        Vt, Vs = A.leaf_basis(0, t, nt), A.leaf_basis(0, s, ns)
        Rts = R[np.ix_(perm[bt:bt + nt], perm[bs:bs + ns])]
        Kts = K[np.ix_(perm[bt:bt + nt], perm[bs:bs + ns])]
        assert np.linalg.norm(Vt @ Rts @ Vs.T) <= 1e-10 * np.linalg.norm(Vt @ Kts @ Vs.T)
        checked += 1
        if checked == 20:
            break
    assert checked > 0
