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
