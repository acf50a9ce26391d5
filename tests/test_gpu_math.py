"""Accuracy of the device exp / sqrt used by the kernel-entry loops.

FP64 routines are checked against numpy (correctly rounded to within an ulp on
glibc): a few ulp are allowed, far below the 1e-12 action tolerance.  FP32
routines use the hardware transcendental instructions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(fn, x):
    import torch
    import fmca_amd
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    fmca_amd._C.diag_math(fn, xd.data_ptr(), yd.data_ptr(), xd.numel(),
                          torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _args():
    rng = np.random.default_rng(5)
    return np.concatenate([-rng.exponential(3.0, 200000), -rng.uniform(0, 740, 20000),
                           np.array([0.0, -1e-300, -1e-17, -0.3465, -0.3466, -700.0, -744.9,
                                     -746.0, -1e6])])


def test_fp64_exp_ulp():
    x = _args()
    y = _run(0, x)
    ref = np.exp(x)
    mask = ref > 1e-300
    rel = np.abs(y[mask] - ref[mask]) / ref[mask]
    print("max rel err", rel.max())
    assert rel.max() < 5e-16
    assert np.all(y[~mask] <= 1e-300) and np.all(y >= 0)


def test_fp64_sqrt_ulp():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(0, 3, 200000), 10.0 ** rng.uniform(-30, 3, 50000),
                        np.array([0.0, 1.0, 2.0, 1e-300, 4.0])])
    y = _run(1, x)
    ref = np.sqrt(x)
    nz = ref > 0
    rel = np.abs(y[nz] - ref[nz]) / ref[nz]
    print("max rel err", rel.max())
    assert rel.max() < 3e-16
    assert y[x == 0][0] == 0.0


def test_fp32_exp_sqrt():
    x = _args()
    x32 = x.astype(np.float32).astype(np.float64)
    y = _run(2, x)
    ref = np.exp(x32)
    mask = ref > 1e-30
    # x * log2(e) is rounded to FP32 before v_exp_f32: relative error grows with |x|
    tol = 3e-7 + 1.2e-7 * np.abs(x32[mask])
    assert np.all(np.abs(y[mask] - ref[mask]) / ref[mask] < tol)
    s = np.random.default_rng(7).uniform(0, 3, 100000)
    s32 = s.astype(np.float32).astype(np.float64)
    ys = _run(3, s)
    assert (np.abs(ys - np.sqrt(s32)) / np.maximum(np.sqrt(s32), 1e-30)).max() < 3e-7
