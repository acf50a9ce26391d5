"""Samplet compression with its kernel blocks evaluated on the GPU, against
the numpy oracle (oracle/samplets.py): the same triplets (positions exactly,
values to 1e-12 of the largest entry), the same a-posteriori selection, and
at n = 1e5 the compression error of FMCA's own test program
(tests/test_H2SampletCompressor.cpp: columns of K through the compressed
matrix and the samplet transforms)."""
import numpy as np
import pytest

import oracle
from oracle.samplets import OracleSampletCompressor, OracleSampletTree, aposteriori_fast

pytestmark = pytest.mark.gpu

# dim, n, dtilde, kernel, l, eta
CASES = [
    (2, 2000, 3, "EXPONENTIAL", 1.0, 0.8),
    (2, 1500, 3, "EXPONENTIAL", 1.0, 0.0),      # no admissible block: T K T^T
    (3, 1500, 2, "MATERN32", 0.5, 0.5),
    (1, 800, 4, "GAUSSIAN", 0.3, 0.8),
    (3, 1200, 3, "MATERN52", 0.7, 0.8),          # q = 35: leaves of up to 69 points
    (2, 3000, 5, "EXPONENTIAL", 1.0, 0.8),       # q = 45, mdtilde = 15
    (3, 3000, 4, "EXPONENTIAL", 1.0, 1.5),       # q = 84, mdtilde = 20
    (3, 4000, 5, "EXPONENTIAL", 1.0, 3.0),       # q = 165: leaves of up to 250 points, 3 tiles
    (2, 600, 1, "INVMULTIQUADRIC", 0.5, 0.8),   # one-point leaves
]


def _sorted(r, c, v):
    o = np.lexsort((c, r))
    return r[o], c[o], v[o]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-n{c[1]}-dt{c[2]}-{c[3]}-eta{c[5]}"
                                             for c in CASES])
def test_compressed_triplets_match_oracle(case):
    import fmca_amd
    dim, n, dtilde, kernel, l, eta = case
    print(f"[case] samplets {case}", flush=True)
    P = np.asfortranarray(np.random.default_rng(dim * 100 + dtilde).random((dim, n)))
    T = fmca_amd.SampletTree(P, dtilde)
    K = fmca_amd.CovarianceKernel(kernel, l)
    comp = fmca_amd.SampletKernelCompressor(T, K, P, eta, 0.0)
    r, c, v = _sorted(*comp.triplets())
    O = OracleSampletTree(P, dtilde)
    ro, co, vo = _sorted(*OracleSampletCompressor(O, kernel, l, eta).triplets())
    # threshold 0 keeps |v| > 0, so an entry that rounds to exactly zero on
    # one side only is absent there; compare on the union of positions
    key, keyo = r * n + c, ro * n + co
    allk = np.union1d(key, keyo)
    a = np.zeros(len(allk))
    b = np.zeros(len(allk))
    a[np.searchsorted(allk, key)] = v
    b[np.searchsorted(allk, keyo)] = vo
    scale = np.abs(vo).max()
    err = np.abs(a - b).max() / scale
    st = comp.statistics()
    print(f"  nnz {len(v)} ({len(v) / n:.0f} per row), oracle {len(vo)}, max diff {err:.2e}, "
          f"dense jobs {st['dense_jobs']:.0f}, far jobs {st['far_jobs']:.0f}", flush=True)
    assert err <= 1e-12
    assert np.all(r <= c)
    assert np.count_nonzero(r == c) == n
    if eta == 0.0:
        assert st["far_jobs"] == 0 and len(v) == n * (n + 1) // 2
    else:
        assert st["far_jobs"] > 0
    # a-posteriori selection: the product's against the oracle's rule on the
    # same triplets
    rr, cc, vv = comp.triplets()
    for thres in (1e-6, 1e-3):
        keep = np.sort(comp._m.aposteriori(thres))
        assert np.array_equal(keep, np.sort(aposteriori_fast(vv, rr, cc, thres)))
    # the a-priori threshold drops |v| <= thres, keeps the diagonal
    comp2 = fmca_amd.SampletKernelCompressor(T, K, P, eta, 1e-6 * scale)
    r2, c2, v2 = _sorted(*comp2.triplets())
    sel = (np.abs(v) > 1e-6 * scale) | (r == c)
    assert np.array_equal(r2, r[sel]) and np.array_equal(c2, c[sel])
    assert np.array_equal(v2, v[sel])


def test_compression_error_at_n_1e5():
    """FMCA's test program's measure (tests/test_H2SampletCompressor.cpp:
    96-121): ten columns of the kernel matrix, through the transformed unit
    vector, the symmetric compressed matrix and the inverse transform, against
    the columns evaluated directly.  2-D, n = 1e5, EXPONENTIAL l = 1,
    dtilde 3, eta 0.8, and after a-posteriori thresholding at 1e-3."""
    import fmca_amd
    import scipy.sparse as sp
    n = 100_000
    print("[case] samplet compression n=1e5 2-D", flush=True)
    P = np.asfortranarray(np.random.default_rng(11).random((2, n)))
    T = fmca_amd.SampletTree(P, 3)
    comp = fmca_amd.SampletKernelCompressor(T, fmca_amd.CovarianceKernel("EXPONENTIAL", 1.0), P,
                                            0.8, 0.0)
    st = comp.statistics()
    print("  ", {k: round(v, 3) for k, v in st.items()}, flush=True)
    perm = T.indices()
    rng = np.random.default_rng(12)
    idx = rng.choice(n, 10, replace=False)
    X = np.zeros((n, 10))
    X[idx, np.arange(10)] = 1.0                          # unit vectors, cluster order
    Y1 = oracle.dense_matrix(P[:, perm], P[:, perm[idx]], "EXPONENTIAL", 1.0).T  # n x 10

    def through(r, c, v):
        S = sp.csr_matrix((v, (r, c)), shape=(n, n))
        S = S + sp.triu(S, 1).T
        return fmca_amd.inverseSampletTransform(T, S @ fmca_amd.sampletTransform(T, X))

    r, c, v = comp.triplets()
    err = np.linalg.norm(through(r, c, v) - Y1) / np.linalg.norm(Y1)
    ra, ca, va = comp.aposteriori_triplets_fast(1e-3)
    err_a = np.linalg.norm(through(ra, ca, va) - Y1) / np.linalg.norm(Y1)
    print(f"  nnz per row {len(v) / n:.0f} (a posteriori {len(va) / n:.0f}); "
          f"compression error {err:.2e} (a posteriori {err_a:.2e})", flush=True)
    # the interpolation error of the far blocks at degree 4 (2-D, eta 0.8)
    assert err < 1e-4
    assert err_a < 5e-3 and len(va) < len(v)
