"""Edge cases of the GPU action against the oracle: single points, one target
or one source, heavy duplication (empty and zero-extent boxes), a flat axis,
point sets smaller than q (all near field), zero right-hand sides, and
right-hand-side counts that leave partial column tiles.  Both interpolators.

Bar as elsewhere: FP64 within 1e-12 of the oracle, FP32 within 1e-5 of FP64.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu

SCHEMES = {"tensor_product": "tensor", "total_degree": "total_degree"}


def _points(kind, n, d, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        P = rng.random((d, n))
    elif kind == "duplicates":  # 7 distinct locations, each repeated
        P = rng.random((d, 7))[:, rng.integers(0, 7, n)]
    elif kind == "flat":  # last axis constant
        P = rng.random((d, n))
        P[-1] = 0.5
    elif kind == "line":  # all points on one segment
        t = rng.random(n)
        P = np.outer(np.linspace(0.2, 0.8, d), t)
    return np.asfortranarray(P)


CASES = [
    # label, kind, n, m (None: same), d, deg, kernel, nrhs
    ("one-point", "uniform", 1, None, 3, 3, "EXPONENTIAL", 2),
    ("one-target", "uniform", 4000, 1, 3, 3, "MATERN32", 3),
    ("one-source", "uniform", 1, 3000, 2, 4, "GAUSSIAN", 1),
    ("duplicates", "duplicates", 3000, None, 3, 3, "MATERN52", 9),
    ("flat-axis", "flat", 5000, None, 3, 3, "MATERN32", 16),
    ("line", "line", 4000, None, 3, 2, "EXPONENTIAL", 1),
    ("fewer-than-q", "uniform", 50, 40, 3, 4, "INVMULTIQUADRIC", 33),
    ("rhs-17", "uniform", 3000, None, 2, 4, "MATERN32", 17),
]


def _run(case, interp):
    import fmca_amd
    label, kind, n, m, d, deg, ker, nrhs = case
    print(f"[case] {label} {interp}", flush=True)
    P = _points(kind, n, d, 1)
    Pe = P if m is None else _points(kind, m, d, 2)
    x = np.asfortranarray(np.random.default_rng(3).standard_normal((n, nrhs)))
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel(ker, 0.4), Pe, P, deg, 0.8,
                               interpolator=interp)
    A = oracle.OracleH2(P, Pe, ker, 0.4, deg, 0.8, scheme=SCHEMES[interp])
    return H, A, x


@pytest.mark.parametrize("interp", list(SCHEMES))
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_edge_case_matches_oracle(case, interp):
    import fmca_amd
    H, A, x = _run(case, interp)
    y = H.action(x)
    yo = A.action(x)
    assert y.shape == yo.shape
    err = rel_err(y, yo)
    print(f"  fp64 vs oracle {err:.2e}", flush=True)
    assert err <= 1e-12
    assert rel_err(H.action(x, precision=fmca_amd.FP32), y) <= 1e-5


def test_zero_rhs_columns():
    import torch
    import fmca_amd
    P = _points("uniform", 2000, 3, 4)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.4), P, P, 3, 0.8)
    y = H.action(np.zeros((2000, 0), order="F"))
    assert y.shape == (2000, 0)
    x = torch.zeros((2000, 0), dtype=torch.float64, device="cuda")
    assert tuple(H.action_device(x).shape) == (2000, 0)
    # a zero right-hand side gives exactly zero
    assert not H.action(np.zeros((2000, 2), order="F")).any()
