"""Pins the FP64 CPU oracle (FMCA restatement) without importing FMCA.

FMCA cannot be compiled here (Eigen3 is absent), and its tests hold no
golden vectors for the H2 action, so every oracle function is pinned by
checks that a plausible mistake would fail: closed forms and independent
formulas, interpolation identities, tree and pattern invariants, FMCA's own
nestedness check, a hand-worked tiny tree, and convergence to dense K*x.
"""
import itertools
import math

import numpy as np
import pytest
from scipy import special, stats

import oracle
from conftest import rel_err, uniform_points

KERNELS_WITH_K0_ONE = ["EXPONENTIAL", "MATERN32", "MATERN52", "MATERN72", "MATERN92", "GAUSSIAN"]


# --- kernels (CovarianceKernel.h:56-116) -------------------------------------
def matern_nu(r, l, nu):
    """Textbook Matern: 2^(1-nu)/Gamma(nu) (sqrt(2 nu) r/l)^nu K_nu(sqrt(2 nu) r/l)."""
    z = np.sqrt(2 * nu) * r / l
    with np.errstate(invalid="ignore"):
        v = 2 ** (1 - nu) / special.gamma(nu) * z ** nu * special.kv(nu, z)
    return np.where(z == 0, 1.0, v)


@pytest.mark.parametrize("name,nu", [("EXPONENTIAL", 0.5), ("MATERN32", 1.5), ("MATERN52", 2.5),
                                     ("MATERN72", 3.5), ("MATERN92", 4.5)])
def test_matern_family_matches_bessel_form(name, nu):
    r = np.concatenate([[0.0], np.linspace(1e-3, 4, 200)])
    for l in (0.3, 1.0, 2.5):
        got = oracle.kernel_eval(name, r, l=l)
        assert np.allclose(got, matern_nu(r, l, nu), rtol=1e-12, atol=1e-15)


def test_gaussian_uses_half_convention():
    r = np.linspace(0, 3, 50)
    for l in (0.1, 0.5, 2.0):
        ref = stats.norm.pdf(r, scale=l) / stats.norm.pdf(0, scale=l)
        assert np.allclose(oracle.kernel_eval("GAUSSIAN", r, l=l), ref, rtol=1e-13)


@pytest.mark.parametrize("name", KERNELS_WITH_K0_ONE)
def test_kernel_at_zero_is_one(name):
    assert oracle.kernel_eval(name, [0.0], l=0.7)[0] == 1.0


def test_other_kernels_hand_values():
    # INVMULTIQUADRIC 1/sqrt((r/l)^2 + c^2), MULTIQUADRIC sqrt(.), TPS2D r^2 log r
    assert oracle.kernel_eval("INVMULTIQUADRIC", [3.0], l=1.0, c=4.0)[0] == pytest.approx(0.2)
    assert oracle.kernel_eval("MULTIQUADRIC", [6.0], l=2.0, c=4.0)[0] == pytest.approx(5.0)
    assert oracle.kernel_eval("TPS2D", [math.e], l=1.0)[0] == pytest.approx(math.e ** 2)
    assert oracle.kernel_eval("TPS2D", [0.0])[0] == 0.0
    assert oracle.kernel_eval("BIHARMONIC3D", [3.0], l=2.0)[0] == pytest.approx(1.5)
    assert oracle.kernel_eval("TRIHARMONIC3D", [3.0], l=1.5)[0] == pytest.approx(8.0)
    assert oracle.kernel_eval("BIHARMONIC2D", [2 * math.e], l=2.0)[0] == pytest.approx(math.e ** 2)
    assert oracle.kernel_eval("BIHARMONIC2D", [0.0])[0] == 0.0
    with pytest.raises(ValueError):
        oracle.kernel_id("MATERNNU")


# --- interpolation (TensorProductInterpolator.h:39-81, MultiIndexSet.h) ------
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5, 8])
def test_chebyshev_nodes_are_roots_and_descending(deg):
    xi, w, _, _ = oracle.interp(1, deg)
    # roots of T_{deg+1}(2x - 1) on [0,1], strictly decreasing
    T = np.polynomial.chebyshev.Chebyshev.basis(deg + 1)
    assert np.max(np.abs(T(2 * xi - 1))) < 1e-12
    assert np.all(np.diff(xi) < 0) and xi.min() > 0 and xi.max() < 1
    # barycentric weights alternate in sign and are symmetric in magnitude
    assert np.all(np.sign(w[:-1]) != np.sign(w[1:])) if deg > 0 else True
    assert np.allclose(np.abs(w), np.abs(w[::-1]))


@pytest.mark.parametrize("dim,deg", [(1, 3), (2, 2), (2, 4), (3, 3)])
def test_multi_index_order_graded_lex(dim, deg):
    _, _, mi, Xi = oracle.interp(dim, deg)
    allidx = sorted(itertools.product(range(deg + 1), repeat=dim), key=lambda a: (sum(a), a))
    assert [tuple(r) for r in mi] == allidx
    xi, _, _, _ = oracle.interp(1, deg)
    assert np.array_equal(Xi, xi[mi.T])


@pytest.mark.parametrize("dim,deg", [(1, 5), (2, 3), (3, 2), (3, 4)])
def test_lagrange_delta_partition_and_reproduction(dim, deg):
    _, _, mi, Xi = oracle.interp(dim, deg)
    q = Xi.shape[1]
    # delta property at the nodes
    assert np.allclose(oracle.eval_polynomials(dim, deg, Xi), np.eye(q), atol=1e-13)
    x = np.random.default_rng(0).random((dim, 200))
    L = oracle.eval_polynomials(dim, deg, x)
    assert np.allclose(L.sum(axis=0), 1.0, atol=1e-12)
    # reproduction of a random polynomial of per-axis degree <= deg
    coef = np.random.default_rng(1).standard_normal(q)

    def f(pts):
        return sum(c * np.prod(pts ** a[:, None], axis=0) for c, a in zip(coef, mi))
    assert np.allclose(f(Xi) @ L, f(x), atol=1e-10 * np.abs(coef).sum())


# --- cluster tree (ClusterTreeInitializer_ClusterTree.h, ClusterTreeSplitter.h)
def test_tiny_tree_by_hand():
    P = np.array([[0.9, 0.1, 0.5, 0.3, 0.7, 0.2, 0.8, 0.4]])
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 1.0, 1, 0.8)  # q = 2: leaves of <= 3 points
    assert A.q == 2
    assert A.indices(1).tolist() == [1, 5, 3, 7, 2, 4, 6, 0]
    cl = A.clusters(1)
    assert cl["size"].tolist() == [8, 4, 4, 2, 2, 2, 2]
    assert cl["begin"].tolist() == [0, 0, 4, 0, 2, 4, 6]
    assert cl["level"].tolist() == [0, 1, 1, 2, 2, 2, 2]
    assert cl["sons"].tolist() == [[1, 2], [3, 4], [5, 6]] + [[-1, -1]] * 4
    # tight boxes after shrink-to-fit
    assert cl["bbmin"][:, 0].tolist() == [0.1, 0.1, 0.5, 0.1, 0.3, 0.5, 0.8]
    assert cl["bbmax"][:, 0].tolist() == [0.9, 0.4, 0.9, 0.2, 0.4, 0.7, 0.9]


# Hand-worked trees in 2-D and 3-D that pin the split-axis rule of
# CardinalityBisection (ClusterTreeSplitter.h:60-81): the axis is the longest
# edge of the node's *split box* (the root's box widened by eps |coordinate|,
# ClusterTreeInitializer_ClusterTree.h:42-52, then cut at the boundary points
# of every split, :72-80), not of the tight point box, and maxCoeff takes the
# FIRST axis on equal edges (ClusterTreeSplitter.h:67).  degree 0 gives q = 1,
# so nodes split down to single points.  Every expected value below was worked
# out by hand from the points; the comments give the steps.
#
# 2-D: A (0.0, .55)  B (.2, .45)  C (.4, .5)  D (.7, 0.)  E (.9, 1.)  F (1., .6),
# stored in the order E B F A D C (E = 0, B = 1, F = 2, A = 3, D = 4, C = 5).
#  root  split box x [0, 1+eps], y [0, 1+eps]: equal edges -> x (first).
#        by x: A B C | D E F
#  c1 {A B C}  split box x [0, .4], y [0, 1+eps] -> y (the tight box
#        x .4 > y .1 would say x).  by y: B | C A
#  c2 {D E F}  split box x [.7, 1+eps], y [0, 1+eps] -> y.  by y: D | F E
#  c12 {C A}   split box x [0, .4], y [.5, 1+eps] -> y (tight: x .4 > y .05).
#        by y: C | A   (by x it would be A | C)
#  c22 {F E}   split box x [.7, 1+eps], y [.6, 1+eps] -> y.  by y: F | E
#        (by x it would be E | F)
#  level order: 0 root, 1 c1, 2 c2, 3 c11, 4 c12, 5 c21, 6 c22, 7-8 sons of c12,
#  9-10 sons of c22; leaves left to right B C A D F E.
HAND_2D = dict(
    P=[[0.9, 1.0], [0.2, 0.45], [1.0, 0.6], [0.0, 0.55], [0.7, 0.0], [0.4, 0.5]],
    perm=[1, 5, 3, 4, 2, 0],
    size=[6, 3, 3, 1, 2, 1, 2, 1, 1, 1, 1],
    begin=[0, 0, 3, 0, 1, 3, 4, 1, 2, 4, 5],
    level=[0, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3],
    sons=[[1, 2], [3, 4], [5, 6], [-1, -1], [7, 8], [-1, -1], [9, 10]] + [[-1, -1]] * 4)
# 3-D: A (0., .45, 0.)  B (.1, .4, 1.)  C (.6, 0., .3)  D (1., 1., .2),
# stored in the order C A D B (C = 0, A = 1, D = 2, B = 3).
#  root  split box [0, 1+eps] on every axis: three equal edges -> x.
#        by x: A B | C D
#  c1 {A B}  split box x [0, .1], y [0, 1+eps], z [0, 1+eps]: y and z equal
#        -> y (first).  The tight box (x .1, y .05, z 1) would say z, and the
#        last of the equal edges would also be z.  by y: B | A (by z: A | B)
#  c2 {C D}  split box x [.6, 1+eps], y, z [0, 1+eps] -> y.  by y: C | D
#        (by z it would be D | C)
#  leaves left to right B A C D.
HAND_3D = dict(
    P=[[0.6, 0.0, 0.3], [0.0, 0.45, 0.0], [1.0, 1.0, 0.2], [0.1, 0.4, 1.0]],
    perm=[3, 1, 0, 2],
    size=[4, 2, 2, 1, 1, 1, 1],
    begin=[0, 0, 2, 0, 1, 2, 3],
    level=[0, 1, 1, 2, 2, 2, 2],
    sons=[[1, 2], [3, 4], [5, 6]] + [[-1, -1]] * 4)


@pytest.mark.parametrize("hand", [HAND_2D, HAND_3D], ids=["2d", "3d"])
def test_split_axis_rule_by_hand(hand):
    P = np.asfortranarray(np.array(hand["P"], dtype=np.float64).T)
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 1.0, 0, 0.8)
    assert A.q == 1
    assert A.indices(1).tolist() == hand["perm"]
    cl = A.clusters(1)
    for k in ("size", "begin", "level", "sons"):
        assert cl[k].tolist() == hand[k], k
    # single-point leaves: the flat-box rule widens each axis by 1e2 eps
    eps = np.finfo(float).eps
    for c in range(len(hand["size"])):
        if hand["sons"][c][0] < 0:
            p = P[:, hand["perm"][hand["begin"][c]]]
            assert np.array_equal(cl["bbmin"][c], p - 1e2 * eps)
            assert np.array_equal(cl["bbmax"][c], p + 1e2 * eps)
    _check_tree(A, 1, P, A.q)


def _check_tree(A, which, P, q):
    n = P.shape[1]
    perm = A.indices(which)
    assert np.array_equal(np.sort(perm), np.arange(n))
    cl = A.clusters(which)
    nc = len(cl["size"])
    assert cl["size"][0] == n and cl["begin"][0] == 0
    assert np.all(np.diff(cl["level"]) >= 0)  # level order
    for c in range(nc):
        b, s = cl["begin"][c], cl["size"][c]
        pts = P[:, perm[b:b + s]]
        s0, s1 = cl["sons"][c]
        if s0 < 0:
            assert s <= 2 * q - 1 and (s >= q or c == 0)
            lo, hi = pts.min(axis=1), pts.max(axis=1)
        else:
            assert s0 > c and s1 == s0 + 1
            assert cl["begin"][s0] == b and cl["size"][s0] == s // 2
            assert cl["begin"][s1] == b + s // 2 and cl["size"][s1] == s - s // 2
            p0 = P[:, perm[b:b + s // 2]]
            p1 = P[:, perm[b + s // 2:b + s]]
            # some axis separates the halves (the cardinality split)
            assert np.any(p0.max(axis=1) <= p1.min(axis=1))
            lo = np.minimum(cl["bbmin"][s0], cl["bbmin"][s1])
            hi = np.maximum(cl["bbmax"][s0], cl["bbmax"][s1])
        flat = hi - lo < 2e2 * np.finfo(float).eps
        lo = np.where(flat, lo - 1e2 * np.finfo(float).eps, lo)
        hi = np.where(flat, hi + 1e2 * np.finfo(float).eps, hi)
        assert np.array_equal(cl["bbmin"][c], lo) and np.array_equal(cl["bbmax"][c], hi)
        assert np.array_equal(cl["bbsize"][c], hi - lo)


@pytest.mark.parametrize("n,d,deg", [(5000, 3, 3), (3000, 2, 4), (2000, 1, 5), (7, 3, 2)])
def test_tree_invariants(n, d, deg):
    P = uniform_points(n, d, seed=n)
    A = oracle.OracleH2(P, P, "MATERN32", 0.5, deg, 0.8)
    _check_tree(A, 1, P, A.q)


def test_tree_with_duplicates_and_flat_axis():
    rng = np.random.default_rng(3)
    P = rng.random((3, 1500))
    P[:, 700:900] = P[:, 700:701]   # 200 copies of one point
    P[2, :] = 0.25                  # a flat axis
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, 2, 0.8)
    _check_tree(A, 1, P, A.q)


@pytest.mark.parametrize("d,deg", [(3, 3), (2, 5), (3, 4), (1, 6)])
def test_transfer_nestedness(d, deg):
    """FMCA's check_transfer_matrices KAT (compute_cluster_bases_impl.h:59-80)."""
    P = uniform_points(4000, d, seed=7)
    A = oracle.OracleH2(P, uniform_points(3000, d, seed=8), "EXPONENTIAL", 0.5, deg, 0.8)
    assert A.check_transfer(0) < 1e-13
    assert A.check_transfer(1) < 1e-13


# --- block pattern (H2MatrixBase.h:160-191, Admissibility.h:95-115) -----------
def test_pattern_covers_every_pair_once_and_admissible():
    P = uniform_points(1500, 2, seed=2)
    Pe = uniform_points(1200, 2, seed=3)
    eta = 0.6
    A = oracle.OracleH2(P, Pe, "EXPONENTIAL", 0.5, 2, eta)
    rows, cols = A.clusters(0), A.clusters(1)
    r, c, lr = A.blocks()
    cover = np.zeros((1200, 1500), dtype=np.int32)
    for t, s, low in zip(r, c, lr):
        cover[rows["begin"][t]:rows["begin"][t] + rows["size"][t],
              cols["begin"][s]:cols["begin"][s] + cols["size"][s]] += 1
        rt = 0.5 * np.linalg.norm(rows["bbsize"][t])
        rs = 0.5 * np.linalg.norm(cols["bbsize"][s])
        ct = 0.5 * (rows["bbmin"][t] + rows["bbmax"][t])
        cs = 0.5 * (cols["bbmin"][s] + cols["bbmax"][s])
        dist = max(np.linalg.norm(ct - cs) - rt - rs, 0.0)
        if low:
            assert max(rt, rs) <= eta * dist
            assert A.q ** 2 < rows["size"][t] * cols["size"][s]
        else:
            leaf = rows["sons"][t][0] < 0 or cols["sons"][s][0] < 0
            assert leaf or max(rt, rs) <= eta * dist
    assert np.all(cover == 1)


# --- whole action --------------------------------------------------------------
def test_action_error_vs_dense_decreases_with_degree_and_eta():
    P = uniform_points(3000, 3, seed=4)
    x = np.random.default_rng(5).uniform(-1, 1, (3000, 2))
    yd = oracle.dense_action(P, P, "EXPONENTIAL", 0.5, x)
    errs = [rel_err(oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, deg, 0.8).action(x), yd)
            for deg in (1, 2, 3, 4)]
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-4
    errs_eta = [rel_err(oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, 2, eta).action(x), yd)
                for eta in (0.8, 0.4, 0.2)]
    assert errs_eta[2] < errs_eta[0], errs_eta


def test_action_forms_agree():
    P = uniform_points(2500, 2, seed=6)
    Pe = uniform_points(1800, 2, seed=9)
    A = oracle.OracleH2(P, Pe, "MATERN32", 0.4, 3, 0.8)
    x = np.random.default_rng(1).standard_normal((2500, 3))
    y = A.action(x)
    ci, ri = A.indices(1), A.indices(0)
    yc = A.action_cluster_order(x[ci])
    assert np.allclose(yc, y[ri], rtol=0, atol=1e-13 * np.abs(y).max())
    rows = A.clusters(0)
    leaves = np.flatnonzero(rows["sons"][:, 0] < 0)[[0, 3, -1]]
    yl = A.action_leaves(x, leaves)
    sel = np.concatenate([ri[rows["begin"][t]:rows["begin"][t] + rows["size"][t]] for t in leaves])
    assert np.allclose(yl, y[sel], rtol=0, atol=1e-13 * np.abs(y).max())


def test_action_linearity_and_symmetry():
    P = uniform_points(2000, 3, seed=10)
    A = oracle.OracleH2(P, P, "MATERN52", 0.5, 3, 0.8)
    rng = np.random.default_rng(2)
    x, z = rng.standard_normal((2000, 1)), rng.standard_normal((2000, 1))
    Ax, Az = A.action(x), A.action(z)
    assert rel_err(A.action(x + 2 * z), Ax + 2 * Az) < 1e-14
    # same tree for rows and columns: the H2 matrix is symmetric
    assert abs((z.T @ Ax - x.T @ Az).item()) < 1e-12 * abs((z.T @ Ax).item())


def test_pcg_solves_regularised_system():
    P = uniform_points(1500, 2, seed=12)
    A = oracle.OracleH2(P, P, "GAUSSIAN", 0.2, 4, 0.8)
    b = np.random.default_rng(3).standard_normal(1500)
    x, it, err = A.pcg(b, ridge=1e-2, tol=1e-8)
    r = b - A.action(x)[:, 0] - 1e-2 * x
    assert err <= 1e-8 and np.linalg.norm(r) / np.linalg.norm(b) <= 1.1e-8 and it > 1


def test_dense_action_matches_scipy_pairwise_distances():
    """dense_action / dense_matrix against K built with scipy's cdist and the
    closed-form kernels (FMCA's CovarianceKernel.h:64-70,93-94)."""
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(21)
    P, Pe = rng.random((3, 700)), rng.random((3, 500))
    x = rng.standard_normal((700, 2))
    R = cdist(Pe.T, P.T)
    forms = {"EXPONENTIAL": lambda r, l: np.exp(-r / l),
             "MATERN32": lambda r, l: (1 + np.sqrt(3) * r / l) * np.exp(-np.sqrt(3) * r / l),
             "GAUSSIAN": lambda r, l: np.exp(-r * r / (2 * l * l))}
    for name, f in forms.items():
        K = f(R, 0.4)
        assert np.allclose(oracle.dense_matrix(P, Pe, name, 0.4), K, rtol=1e-13, atol=1e-15)
        assert rel_err(oracle.dense_action(P, Pe, name, 0.4, x), K @ x) < 1e-13


@pytest.mark.parametrize("dim,deg,kernel", [(2, 3, "MATERN32"), (3, 2, "EXPONENTIAL")])
def test_action_equals_blockwise_assembly(dim, deg, kernel):
    """The oracle's recursive action (upward transfers, coupling, downward
    transfers, near field) equals the H2 matrix assembled block by block in
    numpy from its own block list and boxes:
        y = sum_dense K(t, s) x_s + sum_low-rank V_t^T K(xi_t, xi_s) V_s x_s
    with V_c[a, p] = L_a((x_p - bbmin_c) / bbsize_c) evaluated directly for
    every cluster (no transfer matrices) and xi_c = Xi * bbsize_c + bbmin_c
    (unsymmetricNystromEvaluator.h:42-73, H2kernelActor.h:38-80)."""
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(22)
    n, m = 2400, 2000
    P, Pe = rng.random((dim, n)), rng.random((dim, m))
    x = rng.standard_normal((n, 2))
    l = 0.4
    A = oracle.OracleH2(P, Pe, kernel, l, deg, 0.8)
    f = {"EXPONENTIAL": lambda r: np.exp(-r / l),
         "MATERN32": lambda r: (1 + np.sqrt(3) * r / l) * np.exp(-np.sqrt(3) * r / l)}[kernel]
    _, _, _, Xi = oracle.interp(dim, deg)
    rows, cols = A.clusters(0), A.clusters(1)
    ri, ci = A.indices(0), A.indices(1)

    def pts(tree, idx, Pm, c):
        b, s = tree["begin"][c], tree["size"][c]
        return Pm[:, idx[b:b + s]], idx[b:b + s]

    def basis(tree, idx, Pm, c):
        X, _ = pts(tree, idx, Pm, c)
        U = (X - tree["bbmin"][c][:, None]) / tree["bbsize"][c][:, None]
        return oracle.eval_polynomials(dim, deg, U)

    y = np.zeros((m, 2))
    row, col, lr = A.blocks()
    assert lr.any() and (~lr).any()
    for t, s, low in zip(row, col, lr):
        Xt, it = pts(rows, ri, Pe, t)
        Xs, is_ = pts(cols, ci, P, s)
        if low:
            xt = Xi * rows["bbsize"][t][:, None] + rows["bbmin"][t][:, None]
            xs = Xi * cols["bbsize"][s][:, None] + cols["bbmin"][s][:, None]
            S = f(cdist(xt.T, xs.T))
            y[it] += basis(rows, ri, Pe, t).T @ (S @ (basis(cols, ci, P, s) @ x[is_]))
        else:
            y[it] += f(cdist(Xt.T, Xs.T)) @ x[is_]
    assert rel_err(A.action(x), y) < 1e-12


def _sphere(n, seed):
    P = np.random.default_rng(seed).standard_normal((3, n))
    return P / np.linalg.norm(P, axis=0)


def test_geodesic_distance_closed_forms():
    """setDistanceType("GEODESIC") (CovarianceKernel.h:123-137): the distance
    is acos(clamp(x . y, -1, 1)) (SphereClusterTree.h:65-69).  BIHARMONIC3D
    with l = 1 is k(r) = r, so the kernel matrix is the distance itself."""
    def dist(X, Y):
        return oracle.dense_matrix(np.asarray(Y, float), np.asarray(X, float), "BIHARMONIC3D",
                                   1.0, distance="GEODESIC")

    e = np.eye(3)
    D = dist(e, np.hstack([e, -e]))
    assert np.allclose(D[:, :3], np.pi / 2 * (1 - np.eye(3)), atol=0, rtol=1e-15)
    assert np.allclose(np.diag(D[:, 3:]), np.pi, rtol=1e-15)
    # great-circle distance of unit vectors: 2 asin(chord / 2), where acos is
    # well conditioned (angles above 1e-3)
    X, Y = _sphere(300, 30), _sphere(300, 31)
    G = dist(X, Y)
    chord = np.linalg.norm(X[:, :, None] - Y[:, None, :], axis=0)
    ref = 2 * np.arcsin(chord / 2)
    ok = ref > 1e-3
    assert ok.mean() > 0.99
    assert np.abs(G - ref)[ok].max() < 1e-12
    # off the sphere the dot product is clamped: x . y >= 1 is distance 0,
    # x . y <= -1 is pi
    far = np.array([[2.0, 0.0], [0.0, 0.0], [0.0, 0.0]])
    assert dist(far[:, :1], far[:, :1])[0, 0] == 0.0
    assert dist(far[:, :1], -far[:, :1])[0, 0] == np.pi
    # the Euclidean default is the chord
    assert np.allclose(oracle.dense_matrix(Y, X, "BIHARMONIC3D", 1.0), chord, rtol=1e-13)


@pytest.mark.parametrize("kernel,l", [("GAUSSIAN", 0.5), ("MATERN32", 0.5)])
def test_geodesic_action_error_vs_dense_falls_with_degree(kernel, l):
    """The H2 action with geodesic entries (near blocks between points, far
    blocks between interpolation nodes off the sphere) against the dense
    geodesic product formed in numpy (arccos of clipped P^T P): small, and
    falling in the degree.  The Euclidean chord in either part would leave an
    error near the 15 % by which the two dense products differ."""
    n = 20000
    P = _sphere(n, 32)
    x = np.random.default_rng(33).standard_normal((n, 2))
    f = {"GAUSSIAN": lambda r: np.exp(-0.5 * r * r / (l * l)),
         "MATERN32": lambda r: (1 + np.sqrt(3) * r / l) * np.exp(-np.sqrt(3) * r / l)}[kernel]
    yd = np.zeros((n, 2))
    for b in range(0, n, 2000):
        yd[b:b + 2000] = f(np.arccos(np.clip(P[:, b:b + 2000].T @ P, -1, 1))) @ x
    yo = oracle.dense_action(P, P, kernel, l, x, distance="GEODESIC")
    # the two dense products differ by the rounding of x . y near the diagonal
    assert rel_err(yo, yd) < 1e-8
    assert rel_err(oracle.dense_action(P, P, kernel, l, x), yd) > 1e-2
    errs = []
    for deg in (2, 3, 4):
        A = oracle.OracleH2(P, P, kernel, l, deg, 0.8, distance="GEODESIC")
        assert A.entry_counts()["far_blocks"] > 1000
        errs.append(rel_err(A.action(x), yd))
    assert errs[0] < 1e-2 and errs[2] < 2e-4 and errs[1] < errs[0] and errs[2] < errs[1], errs
