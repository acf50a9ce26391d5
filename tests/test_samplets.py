"""Samplet tree, samplet transforms and the samplet compression: the oracle's
pins and the product's host parts (tree, bases, coefficient order,
transforms) against the oracle.  The compression itself needs the GPU
(tests/test_gpu_samplets.py)."""
import itertools
import math

import numpy as np
import pytest

import oracle
from oracle.samplets import (OracleSampletCompressor, OracleSampletTree, aposteriori_fast,
                             householder_qr, monomials, multinomial_matrix, shift_matrix,
                             total_degree_set)


def _points(dim, n, seed):
    return np.asfortranarray(np.random.default_rng(seed).random((dim, n)))


def test_householder_qr_matches_lapack_up_to_signs():
    """Eigen's convention (beta = -sign(c0) ||x||) against LAPACK's geqrf: the
    same reflections up to the sign of each column of Q and row of R."""
    A = np.random.default_rng(1).standard_normal((23, 7))
    Q, R = householder_qr(A)
    Q2, R2 = np.linalg.qr(A, mode="complete")
    assert np.allclose(Q.T @ Q, np.eye(23), atol=1e-14)
    assert np.allclose(Q[:, :7] @ R, A, atol=1e-13)
    s = np.sign(np.diag(R)) * np.sign(np.diag(R2[:7]))
    assert np.allclose(Q[:, :7], Q2[:, :7] * s, atol=1e-13)
    assert np.allclose(R, R2[:7] * s[:, None], atol=1e-13)
    # Eigen's sign rule: beta = -sign(c0) ||x|| for the first column
    assert R[0, 0] == pytest.approx(-np.sign(A[0, 0]) * np.linalg.norm(A[:, 0]), rel=1e-15)


def test_multi_index_order_and_moment_shift():
    idx = total_degree_set(2, 2)
    assert idx == [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]
    assert len(total_degree_set(3, 3)) == math.comb(6, 3)
    C = multinomial_matrix(idx)
    rng = np.random.default_rng(2)
    X = rng.random((2, 50))
    c_old, c_new = np.array([0.3, 0.6]), np.array([0.1, 0.2])
    M_old = monomials(idx, X - c_old[:, None])
    M_new = monomials(idx, X - c_new[:, None])
    S = shift_matrix(idx, C, c_old - c_new)
    assert np.allclose(S @ M_old, M_new, atol=1e-14)
    assert np.array_equal(shift_matrix(idx, C, np.zeros(2)), np.eye(len(idx)))


@pytest.mark.parametrize("dim,n,dtilde", [(1, 400, 4), (2, 1500, 3), (3, 1200, 2), (2, 600, 1)])
def test_oracle_samplet_basis_is_orthonormal_with_vanishing_moments(dim, n, dtilde):
    """The two quantities FMCA's tests/test_SampletBasis.cpp prints: the
    transform is orthogonal, and samplets annihilate polynomials of degree
    < dtilde (all coefficients but the root's scaling functions vanish)."""
    O = OracleSampletTree(_points(dim, n, 3), dtilde)
    T = O.transform_matrix()
    assert np.abs(T @ T.T - np.eye(n)).max() < 1e-13
    Pol = monomials(O.idx, O.pts).T
    C = O.transform(Pol)
    assert np.abs(C[O.m:]).max() < 1e-12 * np.abs(C).max()
    assert np.allclose(O.inverse_transform(C), Pol, atol=1e-12)
    # leaves have |t| in [q, 2q - 1] scaling/samplet functions, inner clusters 2 m
    assert O.q == math.comb(dim + 2 * (dtilde - 1), dim) and O.m == math.comb(dim + dtilde - 1, dim)
    assert int(O.start[-1] + O.nsamp[-1]) == n


@pytest.mark.parametrize("dim,n,dtilde,kernel", [(2, 1200, 3, "EXPONENTIAL"),
                                                 (3, 900, 2, "MATERN32"),
                                                 (1, 500, 4, "GAUSSIAN")])
def test_oracle_compression_with_no_admissible_block_is_the_transformed_matrix(dim, n, dtilde,
                                                                               kernel):
    """eta = 0: no block is admissible, every recursion ends in dense leaf
    blocks, so the compressed upper triangle equals T K T^T exactly --
    independently of the pattern, recycling and storeBlock windows."""
    O = OracleSampletTree(_points(dim, n, 4), dtilde)
    cmp_ = OracleSampletCompressor(O, kernel, 0.7, 0.0)
    r, c, v = cmp_.triplets()
    T = O.transform_matrix()
    K = oracle.dense_matrix(O.pts, O.pts, kernel, 0.7)
    U = np.triu(T @ K @ T.T)
    M = np.zeros((n, n))
    M[r, c] = v
    assert len(v) == n * (n + 1) // 2
    assert np.abs(M - U).max() <= 1e-13 * np.abs(U).max()


def test_oracle_compression_error_falls_with_dtilde():
    """The error FMCA's tests/test_H2SampletCompressor.cpp prints, on the full
    transformed matrix: small, and falling as dtilde (hence the multipole
    degree 2 (dtilde - 1)) grows."""
    n = 1500
    P = _points(2, n, 5)
    errs = []
    for dt in (2, 3, 4):
        O = OracleSampletTree(P, dt)
        r, c, v = OracleSampletCompressor(O, "EXPONENTIAL", 1.0, 0.8).triplets()
        T = O.transform_matrix()
        KT = T @ oracle.dense_matrix(O.pts, O.pts, "EXPONENTIAL", 1.0) @ T.T
        M = np.zeros((n, n))
        M[r, c] = v
        M = M + np.triu(M, 1).T
        errs.append(np.linalg.norm(M - KT) / np.linalg.norm(KT))
    assert errs[0] < 1e-3 and errs[1] < errs[0] and errs[2] < errs[1], errs


def test_aposteriori_fast_keeps_decades_and_the_diagonal():
    rows = np.array([0, 0, 1, 2, 1, 3])
    cols = np.array([0, 1, 1, 3, 2, 3])
    vals = np.array([1.0, 0.5, 2e-9, 1e-3, 3e-12, 1e-20])
    keep = aposteriori_fast(vals, rows, cols, 1e-6)
    # decades 0, 0 (0.5 -> bucket 1), 3, 9, 12, 16: dropping everything below
    # 1e-3 leaves sqrt(4e-18 + 9e-24 + 1e-40) / |all| < 1e-6 of the norm
    assert sorted(keep.tolist()) == [0, 1, 2, 3, 5]
    keep0 = aposteriori_fast(vals, rows, cols, 0.0)
    assert sorted(keep0.tolist()) == list(range(6))


@pytest.mark.parametrize("dim,n,dtilde", [(2, 1500, 3), (3, 1200, 2), (1, 700, 4), (2, 800, 1),
                                          (3, 40, 3)])
def test_product_samplet_tree_matches_oracle(dim, n, dtilde):
    """The product's tree, coefficient order, bases Q and multiscale bases V,
    and both transforms (host code, no GPU) against the oracle."""
    import fmca_amd
    P = _points(dim, n, 6)
    T = fmca_amd.SampletTree(P, dtilde)
    O = OracleSampletTree(P, dtilde)
    cl = T.clusters()
    assert np.array_equal(T.indices(), O.perm)
    assert np.array_equal(cl["start"], O.start)
    assert np.array_equal(cl["nscal"], O.nscal)
    for c in range(O.ncl):
        Q, V = T.basis(c)
        assert np.abs(Q - O.Q[c]).max() <= 1e-13
        assert np.abs(V - O.V[c]).max() <= 1e-12 * max(1.0, np.abs(O.V[c]).max())
    x = np.random.default_rng(7).standard_normal((n, 3))
    y = fmca_amd.sampletTransform(T, x)
    assert np.abs(y - O.transform(x)).max() <= 1e-13 * np.abs(x).max()
    assert np.abs(fmca_amd.inverseSampletTransform(T, y) - x).max() <= 1e-13 * np.abs(x).max()
    # cluster / natural order round trip and the coefficient levels
    z = np.random.default_rng(8).standard_normal((n, 2))
    assert np.array_equal(T.toNaturalOrder(T.toClusterOrder(z)), z)
    # sampletLevelMapper: round(-log2(diam / root diam)) of each coefficient's cluster
    diam = np.linalg.norm(O.bbsize, axis=1)
    lvl = np.rint(-np.log(diam / diam[0]) / np.log(2.0)).astype(int)
    owner = np.concatenate([np.zeros(int(O.nscal[0]), dtype=np.int64),
                            np.repeat(np.arange(O.ncl), O.nsamp.astype(np.int64))])
    assert np.array_equal(T.levels(), lvl[owner])
    # coefficient -> owning cluster's points
    assert set(T.coeff2indices(0).tolist()) == set(range(n))


def test_samplet_tree_errors_and_compress_without_device():
    import fmca_amd
    with pytest.raises(ValueError):
        fmca_amd.SampletTree(np.zeros((4, 10)), 2)
    with pytest.raises(ValueError):
        fmca_amd.SampletTree(np.random.default_rng(0).random((2, 100)), 6)
    P = np.random.default_rng(0).random((2, 100))
    P[0, 3] = np.nan
    with pytest.raises(ValueError):
        fmca_amd.SampletTree(P, 2)
    T = fmca_amd.SampletTree(_points(2, 300, 9), 0)   # dtilde < 1 is taken as 1
    assert T.clusters()["mdtilde"] == 1
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            fmca_amd.SampletKernelCompressor(T, fmca_amd.CovarianceKernel("EXPONENTIAL", 1.0),
                                             None, 0.8, 0.0)
