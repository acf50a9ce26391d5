"""Failure handling forced from the host: device allocations beyond a cap set
with fmca_set_device_memory_limit fail with FMCA_ERR_OUT_OF_MEMORY (as a full
HBM would) before any kernel is launched, and the actor keeps working once
the cap is lifted.  No kernel is made to fault."""
import numpy as np
import pytest

import oracle
from conftest import rel_err, uniform_points


@pytest.mark.gpu
def test_out_of_memory_is_reported_and_recoverable():
    import torch
    import fmca_amd
    print("[case] device memory cap", flush=True)
    P = uniform_points(20000, 3, seed=51)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.4), P, P, 3, 0.8)
    H.upload()
    x = torch.from_numpy(np.random.default_rng(52).standard_normal((20000, 64))).cuda()
    live, _ = fmca_amd.device_memory()
    fmca_amd.set_device_memory_limit(live + (1 << 20))  # 64-rhs workspaces need far more
    try:
        with pytest.raises(RuntimeError, match="out of device memory"):
            H.action_device(x)
        with pytest.raises(RuntimeError, match="out of device memory"):
            H.action_device(x, precision=fmca_amd.FP32)
    finally:
        fmca_amd.set_device_memory_limit(0)
    y = H.action_device(x).cpu().numpy()
    ref = oracle.OracleH2(P, P, "MATERN32", 0.4, 3, 0.8).action(x.cpu().numpy())
    assert rel_err(y, ref) <= 1e-12


def test_memory_limit_argument_check():
    import fmca_amd
    with pytest.raises(ValueError):
        fmca_amd.set_device_memory_limit(-1)
    fmca_amd.set_device_memory_limit(0)


@pytest.mark.gpu
def test_device_planner_falls_back_to_host_under_the_memory_cap():
    """The device planner's transient buffers count against the cap; with no
    room left, both trees and the block expansion are planned on the host
    (reported as host-fallback), the plan is the host's, and the accounting
    returns to where it was."""
    import fmca_amd
    from fmca_amd import workloads as W
    print("[case] device planner under a memory cap", flush=True)
    P = W.uniform(50_000, 3, 4)
    ker = fmca_amd.CovarianceKernel("MATERN32", 0.5)
    Hh = fmca_amd.H2kernelActor(ker, P, P, 3, 0.8, planner="host")
    live, _ = fmca_amd.device_memory()
    fmca_amd.set_device_memory_limit(live + 1024)  # below any buffer of the build
    try:
        Hd = fmca_amd.H2kernelActor(ker, P, P, 3, 0.8, planner="device")
    finally:
        fmca_amd.set_device_memory_limit(0)
    rep = Hd.plan_report()
    assert rep["rows_tree"] == "host-fallback" and rep["blocks"] == "host-fallback"
    assert np.array_equal(Hd.indices(0), Hh.indices(0))
    for a, b in zip(Hd.blocks(), Hh.blocks()):
        assert np.array_equal(a, b)
    assert fmca_amd.device_memory()[0] == live
    x = np.asfortranarray(np.random.default_rng(1).standard_normal((50_000, 4)))
    assert rel_err(Hd.action(x), Hh.action(x)) <= 1e-14
