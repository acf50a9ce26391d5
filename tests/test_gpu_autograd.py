"""Differentiation form: gradients of L = 0.5 ||K(l) alpha - y||^2 with respect
to alpha, the lengthscale l and the point coordinates, through the
torch.autograd.Function over H2 actions with derivative kernels.

Checked against (a) dense FP64 autograd of the same loss (differs by the H2
approximation error) and (b) centred finite differences (h = 1e-4, FP64) of the
H2 forward itself."""
import math

import numpy as np
import pytest

import oracle
from conftest import rel_err, uniform_points

pytestmark = pytest.mark.gpu


def dense_kernel(ktype, T, S, l):
    import torch
    diff = T[:, :, None] - S[:, None, :]
    r2 = (diff ** 2).sum(0)
    zero = r2 == 0
    r = torch.where(zero, torch.zeros_like(r2), torch.sqrt(torch.where(zero, torch.ones_like(r2), r2)))
    if ktype == "MATERN32":
        a = math.sqrt(3.0) * r / l
        return (1 + a) * torch.exp(-a)
    if ktype == "GAUSSIAN":
        return torch.exp(-0.5 * r2 / l ** 2)
    if ktype == "EXPONENTIAL":
        return torch.exp(-r / l)
    raise ValueError(ktype)


def _problem(n=4096, d=2, nrhs=8, ktype="MATERN32", l=0.3, deg=5, seed=0):
    import torch
    from fmca_amd import workloads as W
    P = torch.from_numpy(W.uniform(n, d, seed)).cuda()
    alpha = torch.from_numpy(W.normal_rhs(n, nrhs, seed + 1)).cuda()
    y = torch.from_numpy(W.normal_rhs(n, nrhs, seed + 2)).cuda()
    return P, alpha, y


def _h2_grads(P, alpha, y, ktype, l, deg):
    import torch
    from fmca_amd.autograd import H2Spec, h2_matvec
    spec = H2Spec(ktype, deg, 0.8)
    Pg = P.clone().requires_grad_(True)
    ag = alpha.clone().requires_grad_(True)
    lt = torch.tensor(l, dtype=torch.float64, requires_grad=True)
    loss = 0.5 * ((h2_matvec(ag, lt, Pg, spec) - y) ** 2).sum()
    loss.backward()
    return loss.item(), ag.grad, lt.grad.item(), Pg.grad, spec


def _dense_grads(P, alpha, y, ktype, l):
    import torch
    Pg = P.clone().requires_grad_(True)
    ag = alpha.clone().requires_grad_(True)
    lt = torch.tensor(l, dtype=torch.float64, device=P.device, requires_grad=True)
    K = dense_kernel(ktype, Pg, Pg, lt)
    loss = 0.5 * ((K @ ag - y) ** 2).sum()
    loss.backward()
    return loss.item(), ag.grad, lt.grad.item(), Pg.grad


@pytest.mark.parametrize("ktype,l,deg", [("MATERN32", 0.3, 5), ("GAUSSIAN", 0.2, 5)])
def test_gradients_vs_dense_autograd(ktype, l, deg):
    print(f"[case] dense autograd {ktype}", flush=True)
    P, alpha, y = _problem(ktype=ktype, l=l, deg=deg)
    L1, ga, gl, gP, spec = _h2_grads(P, alpha, y, ktype, l, deg)
    L0, da, dl, dP = _dense_grads(P, alpha, y, ktype, l)
    # H2 approximation error of the forward sets the scale of the comparison
    import torch
    with torch.no_grad():
        yh = spec.actor(P, None, l).action_device(alpha)
        yd = dense_kernel(ktype, P, P, l) @ alpha
    efwd = rel_err(yh.cpu().numpy(), yd.cpu().numpy())
    ea = rel_err(ga.cpu().numpy(), da.cpu().numpy())
    el = abs(gl - dl) / abs(dl)
    eP = rel_err(gP.cpu().numpy(), dP.cpu().numpy())
    print(f"  forward H2 err {efwd:.2e}; loss {L1:.6e} vs {L0:.6e}; rel err alpha {ea:.2e} "
          f"l {el:.2e} points {eP:.2e}", flush=True)
    assert efwd < 1e-5
    bound = 20 * efwd
    assert abs(L1 - L0) / abs(L0) < bound
    assert ea < bound and el < bound and eP < bound


def test_gradients_vs_centred_finite_differences():
    import torch
    from fmca_amd.autograd import H2Spec, h2_matvec
    print("[case] finite differences", flush=True)
    ktype, l, deg, h = "MATERN32", 0.3, 5, 1e-4
    P, alpha, y = _problem(n=3000, nrhs=4, seed=3)
    _, ga, gl, gP, _ = _h2_grads(P, alpha, y, ktype, l, deg)

    def loss(Pv, av, lv):
        spec = H2Spec(ktype, deg, 0.8)
        with torch.no_grad():
            return 0.5 * ((h2_matvec(av, torch.tensor(lv, dtype=torch.float64), Pv, spec) - y) ** 2).sum().item()
    fl = (loss(P, alpha, l + h) - loss(P, alpha, l - h)) / (2 * h)
    assert abs(fl - gl) / abs(gl) < 1e-7
    v = torch.from_numpy(np.random.default_rng(9).standard_normal(alpha.shape)).cuda()
    fa = (loss(P, alpha + h * v, l) - loss(P, alpha - h * v, l)) / (2 * h)
    assert abs(fa - (ga * v).sum().item()) / abs(fa) < 1e-8
    # point coordinates: the finite difference of the H2 forward also moves the
    # point inside its leaf basis V_t(x_i) and can move a tight box; it agrees
    # with the compressed exact gradient to the accuracy of the derivative of
    # the interpolant (measured ~1.4e-4 relative at degree 5)
    rng = np.random.default_rng(4)
    checked = 0
    for idx in rng.choice(P.shape[1], 40, replace=False):
        if P[:, idx].min() < 0.05 or P[:, idx].max() > 0.95:
            continue
        for k in range(P.shape[0]):
            Pp, Pm = P.clone(), P.clone()
            Pp[k, idx] += h
            Pm[k, idx] -= h
            fp = (loss(Pp, alpha, l) - loss(Pm, alpha, l)) / (2 * h)
            g = gP[k, idx].item()
            assert abs(fp - g) <= 1e-3 * abs(g) + 1e-4 * gP.abs().max().item(), (idx, k, fp, g)
            checked += 1
        if checked >= 12:
            break
    assert checked >= 6


@pytest.mark.parametrize("ktype", oracle.KERNELS)
def test_derivative_kernels_vs_finite_differences(ktype):
    """FMCA_KERNEL_DL and FMCA_KERNEL_PHI against centred differences of the
    oracle's kernel values (small problem: every block is dense, so the H2
    action is the dense product)."""
    import torch
    import fmca_amd
    from fmca_amd.autograd import KERNEL_DL, KERNEL_PHI, _variant
    n, l, c = 200, 0.7, 0.8
    P = uniform_points(n, 2, seed=1) * 2.0
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ktype, l, c), P, P, 7, 0.8)
    assert H.info()["far_blocks"] == 0
    x = np.random.default_rng(2).standard_normal((n, 2))
    xd = torch.from_numpy(x).cuda()
    diff = P[:, :, None] - P[:, None, :]
    r = np.sqrt((diff ** 2).sum(0))
    assert r[~np.eye(n, dtype=bool)].min() > 1e-3   # finite differences below stay valid
    hl = 1e-5 * l
    Kp = oracle.kernel_eval(ktype, r.ravel(), l=l + hl, c=c).reshape(n, n)
    Km = oracle.kernel_eval(ktype, r.ravel(), l=l - hl, c=c).reshape(n, n)
    dK = (Kp - Km) / (2 * hl)
    got = _variant(H, KERNEL_DL, xd, False, fmca_amd.FP64).cpu().numpy()
    # TPS2D does not depend on l: dK/dl = 0 there, and the variant gives 0
    if np.abs(dK).max() == 0:
        assert np.abs(got).max() == 0
    else:
        assert rel_err(got, dK @ x) < 1e-6
    hr = 1e-6
    with np.errstate(divide="ignore", invalid="ignore"):
        kp = oracle.kernel_eval(ktype, (r + hr).ravel(), l=l, c=c).reshape(n, n)
        km = oracle.kernel_eval(ktype, np.maximum(r - hr, 0).ravel(), l=l, c=c).reshape(n, n)
        phi = np.where(r > 1e-3, (kp - km) / (2 * hr) / r, 0.0)
    got = _variant(H, KERNEL_PHI, xd, False, fmca_amd.FP64).cpu().numpy()
    # the coincident diagonal holds phi(0): the limit for smooth kernels, 0
    # where k'(r)/r is singular; compare off the diagonal
    e0 = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    e0[0, 0] = 1.0
    phi0 = _variant(H, KERNEL_PHI, e0, False, fmca_amd.FP64)[0, 0].item()
    np.fill_diagonal(phi, 0.0)
    assert rel_err(got - phi0 * x, phi @ x) < 1e-5


def test_transposed_action_matches_oracle():
    import torch
    import fmca_amd
    P = uniform_points(4000, 3, seed=5)
    Pe = uniform_points(2500, 3, seed=6)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN52", 0.4), Pe, P, 3, 0.8)
    g = np.random.default_rng(7).standard_normal((2500, 5))
    got = H.action_transpose_device(torch.from_numpy(g).cuda()).cpu().numpy()
    ref = oracle.OracleH2(Pe, P, "MATERN52", 0.4, 3, 0.8).action(g)  # K(P, Pe) = K(Pe, P)^T
    assert rel_err(got, ref) <= 1e-12


def test_gradients_rectangular_targets_vs_dense():
    import torch
    from fmca_amd.autograd import H2Spec, h2_matvec
    from fmca_amd import workloads as W
    S = torch.from_numpy(W.uniform(3000, 2, 11)).cuda()
    T = torch.from_numpy(W.uniform(2000, 2, 12)).cuda()
    a = torch.from_numpy(W.normal_rhs(3000, 3, 13)).cuda()
    y = torch.from_numpy(W.normal_rhs(2000, 3, 14)).cuda()
    Sg, Tg = S.clone().requires_grad_(True), T.clone().requires_grad_(True)
    lt = torch.tensor(0.25, dtype=torch.float64, requires_grad=True)
    loss = 0.5 * ((h2_matvec(a, lt, Sg, H2Spec("GAUSSIAN", 5, 0.8), targets=Tg) - y) ** 2).sum()
    loss.backward()
    Sd, Td = S.clone().requires_grad_(True), T.clone().requires_grad_(True)
    ld = torch.tensor(0.25, dtype=torch.float64, device="cuda", requires_grad=True)
    lossd = 0.5 * ((dense_kernel("GAUSSIAN", Td, Sd, ld) @ a - y) ** 2).sum()
    lossd.backward()
    assert rel_err(Sg.grad.cpu().numpy(), Sd.grad.cpu().numpy()) < 1e-5
    assert rel_err(Tg.grad.cpu().numpy(), Td.grad.cpu().numpy()) < 1e-5
    assert abs(lt.grad.item() - ld.grad.item()) / abs(ld.grad.item()) < 1e-5


@pytest.mark.parametrize("kernel,dim,deg,transpose,groups", [
    ("MATERN32", 2, 5, False, [2, 0, 1, 2, 2]),   # fused backward layout, one MFMA pass of 4 + 1
    ("MATERN32", 3, 4, True, [0, 2]),
    ("GAUSSIAN", 3, 3, False, [1, 0, 1]),
    ("INVMULTIQUADRIC", 2, 4, False, [0, 1, 2]),  # run-time kernel switch
    ("EXPONENTIAL", 3, 6, False, [0, 2, 2]),      # q = 343: coupling on VALU, split per group
])
def test_variant_groups_equal_separate_actions(kernel, dim, deg, transpose, groups):
    """fmca_h2_action_variants_device: the 16-column groups of one action carry
    different kernel variants; each group must equal the single-variant action
    of its columns (FP64 within 1e-13, FP32 within 1e-6 relative)."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    print(f"[case] variant groups {kernel} d={dim} deg={deg} T={transpose} {groups}", flush=True)
    n, m = 6000, 5000
    P, Pe = W.uniform(n, dim, 31), W.uniform(m, dim, 32)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(kernel, 0.4), Pe, P, deg, 0.8)
    rows_in, rows_out = (m, n) if transpose else (n, m)
    ncols = 16 * (len(groups) - 1) + 11
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


def test_spec_replans_for_fresh_points_at_a_reused_address():
    """One H2Spec across steps whose points tensor is freshly computed each
    time (P = P0 + shift): the caching allocator may hand the new tensor the
    freed one's address with version 0, so the plan cache must not key on the
    address.  Every step's value equals a fresh spec's on the same points."""
    import torch
    from fmca_amd.autograd import H2Spec, h2_matvec
    print("[case] spec reuse with fresh points tensors", flush=True)
    P0, alpha, _ = _problem(n=3000, d=2, nrhs=4)
    l = torch.tensor(0.3, dtype=torch.float64)
    spec = H2Spec("MATERN32", 4, 0.8)
    for k in range(3):
        P = P0 + 0.05 * k  # freed after the step; the next one may reuse its memory
        y = h2_matvec(alpha, l, P, spec)
        ref = h2_matvec(alpha, l, P, H2Spec("MATERN32", 4, 0.8))
        assert rel_err(y.cpu().numpy(), ref.cpu().numpy()) <= 1e-13, k
        del P, y
    # in-place change behind the version counter (through .data)
    P = P0.clone()
    y0 = h2_matvec(alpha, l, P, spec)
    P.data.mul_(0.5)
    y1 = h2_matvec(alpha, l, P, spec)
    ref = h2_matvec(alpha, l, P0 * 0.5, H2Spec("MATERN32", 4, 0.8))
    assert rel_err(y1.cpu().numpy(), ref.cpu().numpy()) <= 1e-13
    assert rel_err(y0.cpu().numpy(), y1.cpu().numpy()) > 1e-3


def test_backward_uses_its_own_forward_lengthscale():
    """Two forwards of one spec with different lengthscales, then the backward
    of the first: its gradients equal those of a lone forward at that l."""
    import torch
    from fmca_amd.autograd import H2Spec, h2_matvec
    print("[case] backward after a second forward", flush=True)
    P, alpha, y = _problem(n=3000, d=2, nrhs=4)

    def grads(l1, second):
        spec = H2Spec("MATERN32", 4, 0.8)
        Pg = P.clone().requires_grad_(True)
        ag = alpha.clone().requires_grad_(True)
        lt = torch.tensor(l1, dtype=torch.float64, requires_grad=True)
        loss = 0.5 * ((h2_matvec(ag, lt, Pg, spec) - y) ** 2).sum()
        actor0 = [spec._actor]
        if second:
            # the same points object: the spec reuses its actor and sets l = 0.9 on it
            h2_matvec(ag.detach(), torch.tensor(0.9, dtype=torch.float64), Pg, spec)
            assert spec._actor is actor0[0]
        loss.backward()
        return ag.grad, lt.grad, Pg.grad

    a0, l0, p0 = grads(0.3, False)
    a1, l1, p1 = grads(0.3, True)
    assert torch.allclose(a1, a0, rtol=0, atol=1e-13 * a0.abs().max().item())
    assert abs(l1.item() - l0.item()) <= 1e-12 * abs(l0.item())
    assert torch.allclose(p1, p0, rtol=0, atol=1e-12 * p0.abs().max().item())


def _interior_points(H, P, count, seed):
    """Points strictly inside their leaf's box on every axis: a small move of
    such a point changes no box, no split and no block, so the H2 forward is a
    smooth function of it and finite differences see its exact derivative."""
    cl = H.clusters(0)
    perm = H.indices(0)
    cand = []
    for c in np.nonzero(cl["sons"][:, 0] < 0)[0]:
        b, s = cl["begin"][c], cl["size"][c]
        for j in perm[b:b + s]:
            p = P[:, j]
            if np.all(p > cl["bbmin"][c]) and np.all(p < cl["bbmax"][c]):
                cand.append(j)
    rng = np.random.default_rng(seed)
    return rng.choice(np.array(cand), count, replace=False)


@pytest.mark.parametrize("d,ktype,l,deg,nrhs,prec", [
    (2, "MATERN32", 0.3, 5, 8, "FP64"), (3, "EXPONENTIAL", 0.5, 3, 3, "FP64"),
    (2, "GAUSSIAN", 0.2, 4, 16, "FP64"), (1, "MATERN52", 0.2, 6, 1, "FP64"),
    (3, "MATERN32", 0.5, 3, 12, "FP32")])
def test_gradient_kernels_are_the_derivative_of_the_h2_forward(d, ktype, l, deg, nrhs, prec):
    """The shared-tree gradient (fmca_h2_gradient_device, used by the backward)
    against (a) the oracle's K~ G for dL/dalpha, (b) a central difference in
    l of L = <G, K~(l) alpha> for dL/dl (the blocks do not depend on l),
    (c) central differences of L in the coordinates of points inside their
    leaf boxes (h = 1e-5).  FP64 to 1e-6 of the largest gradient entry; the
    FP32 form against the FP64 gradient to 1e-4."""
    import torch
    import fmca_amd
    from fmca_amd import workloads as W
    from fmca_amd.autograd import _gradient
    print(f"[case] gradient kernels d={d} {ktype} deg={deg} nrhs={nrhs} {prec}", flush=True)
    n = 4000
    P = W.uniform(n, d, 21)
    K = fmca_amd.CovarianceKernel(ktype, l)
    H = fmca_amd.H2kernelActor(K, P, P, deg, 0.8)
    A = torch.from_numpy(W.normal_rhs(n, nrhs, 22)).cuda()
    G = torch.from_numpy(W.normal_rhs(n, nrhs, 23)).cuda()
    out64 = _gradient(H, A, G, fmca_amd.FP64)
    assert out64 is not None
    dal, dpo, dl = (t.cpu().numpy() for t in out64)
    An, Gn = A.cpu().numpy(), G.cpu().numpy()
    if prec == "FP32":
        dal32, dpo32, dl32 = (t.cpu().numpy() for t in _gradient(H, A, G, fmca_amd.FP32))
        assert rel_err(dal32, dal) <= 1e-4 and rel_err(dpo32, dpo) <= 1e-4
        assert abs(dl32[0] - dl[0]) <= 1e-4 * abs(dl[0])
        return
    ref = oracle.OracleH2(P, P, ktype, l, deg, 0.8)
    assert rel_err(dal, ref.action(Gn)) <= 1e-12

    def L_of(Pm, lm):
        Hm = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ktype, lm), Pm, Pm, deg, 0.8)
        return float(np.sum(Gn * Hm.action(An))), Hm

    h = 1e-5 * l
    fd_l = (L_of(P, l + h)[0] - L_of(P, l - h)[0]) / (2 * h)
    assert abs(fd_l - dl[0]) <= 1e-7 * abs(dl[0])
    scale = np.abs(dpo).max()
    perm0 = H.indices(0)
    for j in _interior_points(H, P, 6, 24):
        for k in range(d):
            hp = 1e-5
            Pp, Pm = P.copy(order="F"), P.copy(order="F")
            Pp[k, j] += hp
            Pm[k, j] -= hp
            Lp, Hp = L_of(Pp, l)
            Lm, Hm = L_of(Pm, l)
            assert np.array_equal(Hp.indices(0), perm0) and np.array_equal(Hm.indices(0), perm0)
            fd = (Lp - Lm) / (2 * hp)
            assert abs(fd - dpo[k, j]) <= 1e-6 * scale, (j, k, fd, dpo[k, j])


def test_exponential_3d_deg4_16rhs_forward_gradients_and_refined_solve():
    """The n = 8e6 configuration's shape (uniform 3-D, EXPONENTIAL l = 0.25,
    degree 4, eta 0.8, 16 rhs) at n = 4096, where the oracle and dense FP64
    autograd finish in seconds.
    Forward: FP64 vs the oracle <= 1e-12, FP32 vs FP64 <= 1e-5.
    Backward (d alpha, d l; no point gradients, as at n = 8e6, where the
    exponential's k'(r)/r is singular at coincident points anyway), FP64 and
    FP32: against dense FP64 autograd (bound tied to the forward's H2 error)
    and centred differences (h = 1e-4) of the H2 forward; FP32 vs FP64 <= 1e-4.
    Refinement leg: GCR on BF16-split FP32 matvecs with FP64 residual
    replacement reaches an FP64 residual <= 1e-8 (rechecked through the FP64
    action) and FP64 CG's solution, at ridge 1e-4 n as at n = 8e6."""
    import torch
    import fmca_amd
    from fmca_amd.autograd import H2Spec, h2_matvec
    print("[case] exponential 3-D deg 4, 16 rhs, n = 4096", flush=True)
    n, nrhs, ktype, l, deg = 4096, 16, "EXPONENTIAL", 0.25, 4
    P, alpha, y = _problem(n=n, d=3, nrhs=nrhs, seed=40)
    Pn, an = P.cpu().numpy(), alpha.cpu().numpy()
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ktype, l), Pn, Pn, deg, 0.8)
    assert H.info()["far_blocks"] > 0
    y64 = H.action_device(alpha, precision=fmca_amd.FP64)
    y32 = H.action_device(alpha, precision=fmca_amd.FP32)
    ref = oracle.OracleH2(Pn, Pn, ktype, l, deg, 0.8).action(an)
    assert rel_err(y64.cpu().numpy(), ref) <= 1e-12
    assert rel_err(y32.cpu().numpy(), ref) <= 1e-5
    with torch.no_grad():
        efwd = rel_err(ref, (dense_kernel(ktype, P, P, l) @ alpha).cpu().numpy())
    assert efwd < 1e-3
    _, da, dl, _ = _dense_grads(P, alpha, y, ktype, l)

    def h2_grads(prec):
        spec = H2Spec(ktype, deg, 0.8, precision=prec)
        ag = alpha.clone().requires_grad_(True)
        lt = torch.tensor(l, dtype=torch.float64, requires_grad=True)
        loss = 0.5 * ((h2_matvec(ag, lt, P, spec) - y) ** 2).sum()
        loss.backward()
        return ag.grad, lt.grad.item()

    ga, gl = h2_grads(fmca_amd.FP64)
    ga32, gl32 = h2_grads(fmca_amd.FP32)
    ea = rel_err(ga.cpu().numpy(), da.cpu().numpy())
    el = abs(gl - dl) / abs(dl)
    print(f"  forward H2 err {efwd:.2e}; vs dense: alpha {ea:.2e} l {el:.2e}; "
          f"fp32 vs fp64: alpha {rel_err(ga32.cpu().numpy(), ga.cpu().numpy()):.2e} "
          f"l {abs(gl32 - gl) / abs(gl):.2e}", flush=True)
    assert ea < 20 * efwd and el < 20 * efwd
    assert rel_err(ga32.cpu().numpy(), ga.cpu().numpy()) <= 1e-4
    assert abs(gl32 - gl) <= 1e-4 * abs(gl)
    h = 1e-4
    spec = H2Spec(ktype, deg, 0.8)

    def loss(av, lv):
        with torch.no_grad():
            return 0.5 * ((h2_matvec(av, torch.tensor(lv, dtype=torch.float64), P, spec) - y)
                          ** 2).sum().item()
    fl = (loss(alpha, l + h) - loss(alpha, l - h)) / (2 * h)
    assert abs(fl - gl) <= 1e-6 * abs(gl)
    v = torch.from_numpy(np.random.default_rng(41).standard_normal(alpha.shape)).cuda()
    fa = (loss(alpha + h * v, l) - loss(alpha - h * v, l)) / (2 * h)
    assert abs(fa - (ga * v).sum().item()) <= 1e-8 * abs(fa)
    # refinement leg
    ridge = 1e-4 * n
    z = torch.zeros_like(y)
    x64, i64 = H.cg(y, x0=z, ridge=ridge, tol=1e-8, precision=fmca_amd.FP64)
    xr, ir = H.cg(y, x0=z, ridge=ridge, tol=1e-8, precision=fmca_amd.FP32_BF16,
                  method="gcr", refine_factor=1e-3)
    print(f"  fp64 cg {i64['iterations']} it; refined gcr {ir}", flush=True)
    assert i64["converged"] and ir["converged"] and ir["matvecs_fp32"] > 0
    r = y - H.action_device(xr) - ridge * xr
    assert (torch.linalg.norm(r, dim=0) / torch.linalg.norm(y, dim=0)).max().item() <= 1e-8
    # both residuals <= 1e-8 bound the difference by about cond * 2e-8, cond ~ 1e3 here
    assert (torch.linalg.norm(xr - x64) / torch.linalg.norm(x64)).item() <= 1e-4
