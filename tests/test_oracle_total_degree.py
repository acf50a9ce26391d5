"""Pins the oracle's total-degree interpolation (TotalDegreeInterpolator).

FMCA's own actor test and Python binding use NystromMoments with
TotalDegreeInterpolator (tests/test_H2kernelActor.cpp:20-24, py/FMCA.cpp:37).
FMCA cannot be built here (Eigen3 is absent) and its tests hold no golden
vectors, so each piece is pinned by an independent route:
  * Legendre values: scipy's eval_legendre and Gauss-Legendre orthonormality;
  * index set: count binom(deg+d, d), membership, graded-lexicographic order;
  * DeMarchi points: the Halton radical inverses in exact rational arithmetic
    and LAPACK's column-pivoted QR (scipy.linalg.qr, xGEQP3) on them;
  * invV: V invV = I, the Lagrange property and polynomial reproduction;
  * transfers: FMCA's nestedness check (compute_cluster_bases_impl.h:59-80);
  * the whole action: a numpy assembly from the block list with scipy
    Legendre bases and numpy's inverse, and convergence to dense K x.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy import linalg, special
from scipy.spatial.distance import cdist

import oracle
from conftest import rel_err

CASES = [(1, 5), (2, 4), (2, 6), (3, 3), (3, 4), (3, 6)]


def radical_inverse(k: int, b: int) -> float:
    x, f = Fraction(0), Fraction(1, b)
    while k:
        x += (k % b) * f
        f /= b
        k //= b
    return float(x)


def legendre_basis(mi, U):
    """P_alpha(u) = prod_i sqrt(2 a_i + 1) L_{a_i}(2 u_i - 1); U: (dim, npts)."""
    out = np.ones((len(mi), U.shape[1]))
    for k, a in enumerate(mi):
        for i, ai in enumerate(a):
            out[k] *= np.sqrt(2 * ai + 1) * special.eval_legendre(ai, 2 * U[i] - 1)
    return out


def test_legendre_values_and_orthonormality():
    deg = 7
    u = np.linspace(0, 1, 41)[None, :]
    got = oracle.eval_polynomials(1, deg, u, "total_degree")
    ref = legendre_basis([(a,) for a in range(deg + 1)], u)
    assert np.abs(got - ref).max() < 1e-13
    # L2([0,1]) orthonormality by Gauss-Legendre quadrature
    x, w = np.polynomial.legendre.leggauss(deg + 2)
    G = oracle.eval_polynomials(1, deg, (0.5 * (x + 1))[None, :], "total_degree")
    assert np.abs((G * (0.5 * w)) @ G.T - np.eye(deg + 1)).max() < 1e-13


@pytest.mark.parametrize("dim,deg", CASES)
def test_total_degree_index_set(dim, deg):
    mi, Xi, V, invV = oracle.interp_total_degree(dim, deg)
    q = math.comb(deg + dim, dim)
    assert mi.shape == (q, dim) and Xi.shape == (dim, q)
    assert (mi.sum(axis=1) <= deg).all() and len({tuple(a) for a in mi}) == q
    keys = [(int(a.sum()), tuple(a)) for a in mi]
    assert keys == sorted(keys)  # MultiIndexSet.h:87-100


@pytest.mark.parametrize("dim,deg", CASES)
def test_demarchi_points_are_lapack_pivots_of_halton_points(dim, deg):
    mi, Xi, V, invV = oracle.interp_total_degree(dim, deg)
    q = len(mi)
    primes = [2, 3, 5][:dim]
    # HaltonSet<100>: the points after 100 warm-up steps (DeMarchiPoints.h:27-33)
    cand = np.array([[radical_inverse(100 + j, p) for p in primes]
                     for j in range(10 * q)]).T
    VT = legendre_basis(mi, cand)
    _, _, piv = linalg.qr(VT, pivoting=True, mode="economic")
    assert np.array_equal(Xi, cand[:, piv[:q]])


@pytest.mark.parametrize("dim,deg", CASES)
def test_invV_lagrange_property_and_reproduction(dim, deg):
    mi, Xi, V, invV = oracle.interp_total_degree(dim, deg)
    q = len(mi)
    assert np.abs(V - legendre_basis(mi, Xi).T).max() < 1e-13
    assert np.abs(V @ invV - np.eye(q)).max() < 1e-12
    # l(x) = invV^T P(x) is the Lagrange basis of the nodes
    L = invV.T @ oracle.eval_polynomials(dim, deg, Xi, "total_degree")
    assert np.abs(L - np.eye(q)).max() < 1e-12
    # a polynomial of total degree <= deg is reproduced by its interpolant
    rng = np.random.default_rng(dim * 10 + deg)
    coef = rng.standard_normal(q)
    U = rng.random((dim, 50))

    def poly(X):  # monomials of the same index set
        return sum(c * np.prod([X[i] ** a[i] for i in range(dim)], axis=0)
                   for c, a in zip(coef, mi))
    interp = poly(Xi) @ (invV.T @ oracle.eval_polynomials(dim, deg, U, "total_degree"))
    assert np.abs(interp - poly(U)).max() < 1e-10 * max(1, np.abs(poly(U)).max())


@pytest.mark.parametrize("dim,deg", [(2, 4), (3, 3)])
def test_total_degree_transfers_nested(dim, deg):
    rng = np.random.default_rng(7)
    P = rng.random((dim, 3000))
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, deg, 0.8, scheme="total_degree")
    assert A.check_transfer(0) < 1e-12


@pytest.mark.parametrize("dim,deg,kernel", [(2, 4, "EXPONENTIAL"), (3, 3, "MATERN32")])
def test_total_degree_action_equals_blockwise_numpy(dim, deg, kernel):
    """y = sum_dense K(t, s) x_s + sum_low-rank P_t^T invV K(xi_t, xi_s) invV^T P_s x_s
    (unsymmetricNystromEvaluator.h:42-58, NystromMoments.h:24-34) with scipy
    Legendre bases evaluated directly per cluster and numpy's inverse of V."""
    rng = np.random.default_rng(23)
    n, m = 2400, 2000
    P, Pe = rng.random((dim, n)), rng.random((dim, m))
    x = rng.standard_normal((n, 2))
    l = 0.4
    A = oracle.OracleH2(P, Pe, kernel, l, deg, 0.8, scheme="total_degree")
    f = {"EXPONENTIAL": lambda r: np.exp(-r / l),
         "MATERN32": lambda r: (1 + np.sqrt(3) * r / l) * np.exp(-np.sqrt(3) * r / l)}[kernel]
    mi, Xi, _, _ = oracle.interp_total_degree(dim, deg)
    iV = np.linalg.inv(legendre_basis(mi, Xi).T)
    rows, cols = A.clusters(0), A.clusters(1)
    ri, ci = A.indices(0), A.indices(1)

    def pts(tree, idx, Pm, c):
        b, s = tree["begin"][c], tree["size"][c]
        return Pm[:, idx[b:b + s]], idx[b:b + s]

    def basis(tree, idx, Pm, c):
        X, _ = pts(tree, idx, Pm, c)
        return legendre_basis(mi, (X - tree["bbmin"][c][:, None]) / tree["bbsize"][c][:, None])

    y = np.zeros((m, 2))
    row, col, lr = A.blocks()
    assert lr.any() and (~lr).any()
    for t, s, low in zip(row, col, lr):
        Xt, it = pts(rows, ri, Pe, t)
        Xs, is_ = pts(cols, ci, P, s)
        if low:
            xt = Xi * rows["bbsize"][t][:, None] + rows["bbmin"][t][:, None]
            xs = Xi * cols["bbsize"][s][:, None] + cols["bbmin"][s][:, None]
            S = iV @ f(cdist(xt.T, xs.T)) @ iV.T
            y[it] += basis(rows, ri, Pe, t).T @ (S @ (basis(cols, ci, P, s) @ x[is_]))
        else:
            y[it] += f(cdist(Xt.T, Xs.T)) @ x[is_]
    assert rel_err(A.action(x), y) < 1e-12


def test_fmca_actor_test_configuration():
    """FMCA's tests/test_H2kernelActor.cpp: 2-D points in [-1,1]^2, EXPONENTIAL
    l = 1, degree 4, total degree, columns of K for unit vectors, eta halved
    from 0.8 (4000 points instead of 10000).  The error falls with eta."""
    rng = np.random.default_rng(5)
    n = 4000
    P = rng.uniform(-1, 1, (2, n))
    idx = rng.integers(0, n, 10)
    X = np.zeros((n, 10))
    X[idx, np.arange(10)] = 1
    C = np.exp(-cdist(P.T, P[:, idx].T))
    errs = []
    for eta in (0.8, 0.4, 0.2, 0.1):
        A = oracle.OracleH2(P, P, "EXPONENTIAL", 1.0, 4, eta, scheme="total_degree")
        assert A.q == 15
        errs.append(rel_err(A.action(X), C))
    assert errs[0] < 1e-4 and all(b < a for a, b in zip(errs, errs[1:]))
