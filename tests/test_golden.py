"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from seeded inputs and the FP64 oracle): regression pins of the planner and of
the action.  CPU: the oracle and the planner still reproduce them.  GPU: the
FP64 action matches them within 1e-12 relative, the FP32 action within 1e-5."""
import glob
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT, rel_err

FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))
IDS = [os.path.basename(f)[:-4] for f in FIXTURES]


def load(path):
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    g["interpolator"] = str(g.get("interpolator", "tensor_product"))
    return g


SCHEME = {"tensor_product": "tensor", "total_degree": "total_degree"}


def actor(g):
    import fmca_amd
    K = fmca_amd.CovarianceKernel(str(g["kernel"]), float(g["l"]))
    return fmca_amd.H2kernelActor(K, g["Peval"], g["P"], int(g["deg"]), float(g["eta"]),
                                  interpolator=g["interpolator"])


def test_fixtures_present():
    assert len(FIXTURES) >= 5
    assert {load(f)["interpolator"] for f in FIXTURES} == {"tensor_product", "total_degree"}


@pytest.mark.parametrize("path", FIXTURES, ids=IDS)
def test_oracle_reproduces_fixture(path):
    g = load(path)
    A = oracle.OracleH2(g["P"], g["Peval"], str(g["kernel"]), float(g["l"]), int(g["deg"]),
                        float(g["eta"]), scheme=SCHEME[g["interpolator"]])
    assert np.array_equal(A.indices(0), g["rindices"])
    assert np.array_equal(A.indices(1), g["cindices"])
    row, col, lr = A.blocks()
    assert np.array_equal(row, g["block_row"]) and np.array_equal(col, g["block_col"])
    assert np.array_equal(lr, g["block_lr"])
    assert rel_err(A.action(g["rhs"]), g["action"]) <= 1e-14
    # every fixture exercises the far field; its H2 error against dense K*x is small
    assert lr.sum() > 0
    assert 0 < rel_err(g["action"], g["dense"]) <= 5e-3


@pytest.mark.parametrize("path", FIXTURES, ids=IDS)
def test_planner_reproduces_fixture(path):
    g = load(path)
    H = actor(g)
    if g["interpolator"] == "total_degree":
        # DeMarchi nodes bit for bit, invV to rounding
        ip = H.interpolation()
        assert np.array_equal(ip["Xi"], g["Xi"])
        assert np.abs(ip["invV"] - g["invV"]).max() <= 1e-13
    assert np.array_equal(H.rindices(), g["rindices"])
    assert np.array_equal(H.cindices(), g["cindices"])
    row, col, lr = H.blocks()
    assert np.array_equal(row, g["block_row"]) and np.array_equal(col, g["block_col"])
    assert np.array_equal(lr, g["block_lr"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=IDS)
def test_gpu_action_matches_fixture(path):
    import torch
    import fmca_amd
    g = load(path)
    print(f"[case] golden {os.path.basename(path)}", flush=True)
    H = actor(g)
    x = torch.from_numpy(g["rhs"]).cuda()
    y64 = H.action_device(x, precision=fmca_amd.FP64).cpu().numpy()
    y32 = H.action_device(x, precision=fmca_amd.FP32).cpu().numpy()
    assert rel_err(y64, g["action"]) <= 1e-12
    assert rel_err(y32, g["action"]) <= 1e-5
