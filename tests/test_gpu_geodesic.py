"""The geodesic distance of CovarianceKernel::setDistanceType("GEODESIC")
(CovarianceKernel.h:123-137, SphereClusterTree.h:65-69) on the GPU: kernel
entries acos(clamp(x . y, -1, 1)) in FP64 on the VALU forms, trees and
admissibility Euclidean as in FMCA's H2kernelActor.

Bar: the FP64 action within 1e-12 relative of the oracle with the same
distance (the oracle and the kernels form x . y in the same order, so acos,
which amplifies the rounding of x . y by 1 / r near the diagonal, sees the
same argument), in both orders, transposed, assembled, with the dk/dl
variant against finite differences, and in an FP64 solve; the forms that do
not take the geodesic distance raise."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _sphere(n, seed):
    P = np.random.default_rng(seed).standard_normal((3, n))
    return np.asfortranarray(P / np.linalg.norm(P, axis=0))


# (label, n, m or None, deg, kernel, l, eta, nrhs, interpolator)
CASES = [
    ("exp-l2-1rhs", 20000, None, 3, "EXPONENTIAL", 2.0, 0.8, 1, "tensor_product"),
    ("matern32-8rhs", 20000, None, 4, "MATERN32", 0.5, 0.8, 8, "tensor_product"),
    ("gauss-total-degree-32rhs", 20000, None, 4, "GAUSSIAN", 0.5, 0.8, 32, "total_degree"),
    ("multiquadric-rect-3rhs", 12000, 9000, 3, "MULTIQUADRIC", 0.6, 0.8, 3, "tensor_product"),
    ("matern52-rect-40rhs", 9000, 12000, 3, "MATERN52", 0.5, 0.5, 40, "tensor_product"),
]


def _setup(case):
    import fmca_amd
    label, n, m, deg, ker, l, eta, nrhs, interp = case
    print(f"[case] geodesic {label}", flush=True)
    P = _sphere(n, 41)
    Peval = P if m is None else _sphere(m, 42)
    K = fmca_amd.CovarianceKernel(ker, l)
    K.setDistanceType("GEODESIC")
    H = fmca_amd.H2kernelActor(K, Peval, P, deg, eta, interpolator=interp)
    scheme = "tensor" if interp == "tensor_product" else "total_degree"
    A = oracle.OracleH2(P, Peval, ker, l, deg, eta, scheme=scheme, distance="GEODESIC")
    x = np.asfortranarray(np.random.default_rng(43).standard_normal((n, nrhs)))
    return H, A, x, P, Peval


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_geodesic_fp64_action_matches_oracle(case):
    H, A, x, P, Peval = _setup(case)
    assert H._a.distance_type() == "GEODESIC"
    y = H.action(x)
    yo = A.action(x)
    err = rel_err(y, yo)
    print(f"  fp64 vs oracle rel err {err:.3e}", flush=True)
    assert err <= FP64_TOL
    assert rel_err(H.action_cluster_order(x), A.action_cluster_order(x)) <= FP64_TOL
    # the Euclidean actor on the same plan differs by far more than the bar
    import fmca_amd
    E = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(case[4], case[5]), Peval, P, case[3],
                               case[6], interpolator=case[8])
    assert rel_err(E.action(x), yo) > 1e-4


@pytest.mark.parametrize("case", [CASES[0], CASES[3], CASES[4]],
                         ids=[CASES[i][0] for i in (0, 3, 4)])
def test_geodesic_transposed_action_matches_oracle(case):
    """K^T x through the transposed action against the oracle built with the
    roles of the point sets swapped (the distance is symmetric)."""
    import torch
    H, _, _, P, Peval = _setup(case)
    label, n, m, deg, ker, l, eta, nrhs, interp = case
    m = n if m is None else m
    scheme = "tensor" if interp == "tensor_product" else "total_degree"
    At = oracle.OracleH2(Peval, P, ker, l, deg, eta, scheme=scheme, distance="GEODESIC")
    g = np.asfortranarray(np.random.default_rng(44).standard_normal((m, nrhs)))
    gt = torch.from_numpy(g).cuda()
    yt = H.action_transpose_device(gt).cpu().numpy()
    err = rel_err(yt, At.action(g))
    print(f"  transposed vs oracle rel err {err:.3e}", flush=True)
    assert err <= FP64_TOL


def test_geodesic_assembled_equals_matrix_free_and_fp32_is_refused():
    import fmca_amd
    H, A, x, _, _ = _setup(CASES[1])
    y = H.action(x)
    H.assemble(fmca_amd.FP64)
    ya = H.action(x)
    err = rel_err(ya, y)
    print(f"  assembled vs matrix-free rel err {err:.3e}", flush=True)
    assert err <= 1e-13
    assert rel_err(ya, A.action(x)) <= FP64_TOL
    H.release_assembled()
    for prec in (fmca_amd.FP32, fmca_amd.FP32_BF16):
        with pytest.raises(ValueError, match="geodesic"):
            H.action(x, precision=prec)
    with pytest.raises(ValueError, match="geodesic"):
        H.assemble(fmca_amd.FP32)
    # switching back to the Euclidean distance restores the Euclidean action
    E = fmca_amd.CovarianceKernel(CASES[1][4], CASES[1][5])
    H.set_kernel(E)
    assert H._a.distance_type() == "EUCLIDEAN"
    Ae = oracle.OracleH2(*_sphere_pair(CASES[1]), CASES[1][4], CASES[1][5], CASES[1][3],
                         CASES[1][6])
    assert rel_err(H.action(x), Ae.action(x)) <= FP64_TOL


def _sphere_pair(case):
    P = _sphere(case[1], 41)
    return P, (P if case[2] is None else _sphere(case[2], 42))


def test_geodesic_dl_variant_matches_finite_differences():
    """The dk/dl variant (FMCA_KERNEL_DL) with geodesic entries against the
    centred difference of the FP64 action in l (h = 1e-4 l)."""
    import torch
    import fmca_amd
    from fmca_amd.autograd import KERNEL_DL
    label, n, m, deg, ker, l, eta, nrhs, interp = CASES[1]
    P = _sphere(n, 41)

    def actor(ll):
        K = fmca_amd.CovarianceKernel(ker, ll)
        K.setDistanceType("GEODESIC")
        return fmca_amd.H2kernelActor(K, P, P, deg, eta)

    x = np.asfortranarray(np.random.default_rng(45).standard_normal((n, 2)))
    H = actor(l)
    xt = torch.from_numpy(x).cuda()
    out = torch.empty_like(xt)
    H._a.action_variants_device([KERNEL_DL], False, xt.data_ptr(), n, 2, out.data_ptr(), n,
                                fmca_amd.FP64, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    h = 1e-4 * l
    fd = (actor(l + h).action(x) - actor(l - h).action(x)) / (2 * h)
    err = rel_err(out.cpu().numpy(), fd)
    print(f"  dK/dl vs FD rel err {err:.3e}", flush=True)
    assert err < 1e-6


def test_geodesic_fp64_solve_and_refused_gradient():
    """FMCA's pCG on (K + ridge I) x = b with geodesic entries, FP64 matvecs:
    the residual through the oracle's action meets the tolerance; the
    lower-precision solves and the differentiable point gradient raise."""
    import torch
    import fmca_amd
    # K~ + ridge I stays definite: the H2 error of this kernel (about 1e-4 of
    # ||K|| ~ 400) is well below the ridge
    H, A, _, P, _ = _setup(("gauss-solve", 20000, None, 4, "GAUSSIAN", 0.2, 0.8, 2,
                            "tensor_product"))
    n = P.shape[1]
    b = np.asfortranarray(np.random.default_rng(46).standard_normal((n, 2)))
    ridge = 1.0
    bt = torch.from_numpy(b).cuda()
    x, info = H.cg(bt, ridge=ridge, tol=1e-8, precision=fmca_amd.FP64)
    x = np.asfortranarray(x.cpu().numpy())
    res = b - (A.action(x) + ridge * x)
    err = np.linalg.norm(res, axis=0) / np.linalg.norm(b, axis=0)
    print(f"  FP64 CG residual {err}", flush=True)
    assert err.max() <= 1.1e-8
    with pytest.raises(ValueError, match="geodesic"):
        H.cg(bt, ridge=ridge, tol=1e-8, precision=fmca_amd.FP32)
    a = bt[:, :1].contiguous()
    g = bt[:, 1:].contiguous()
    da = torch.empty_like(a)
    dp = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    dl = torch.empty(1, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="Euclidean"):
        H._a.gradient_device(a.data_ptr(), n, g.data_ptr(), n, 1, fmca_amd.FP64, da.data_ptr(), n,
                             dp.data_ptr(), n, dl.data_ptr(),
                             torch.cuda.current_stream().cuda_stream)


def test_distance_type_names():
    import fmca_amd
    K = fmca_amd.CovarianceKernel("GAUSSIAN", 0.5)
    assert K.distanceType() == "EUCLIDEAN"
    K.setDistanceType("geodesic")
    assert K.distanceType() == "GEODESIC"
    with pytest.raises(ValueError):
        K.setDistanceType("MANHATTAN")
    P = _sphere(500, 47)
    H = fmca_amd.H2kernelActor(K, P, P, 2, 0.8)
    with pytest.raises(ValueError, match="distance"):
        H._a.set_distance_type("MANHATTAN")
    assert H._a.distance_type() == "GEODESIC"
