"""fmca_h2_action_device_noalloc through the C ABI: the same result as
fmca_h2_action_device (bit for bit: the same kernels, only the workspace
differs) on two calls in a row, the second with a smaller input; no device
memory is allocated by it; oversize workloads and a plan not yet on the
device are refused with a status."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "fmca_amd", "libfmca_amd.so")


class Desc(ctypes.Structure):
    _fields_ = [("ktype", ctypes.c_char_p), ("l", ctypes.c_double), ("c", ctypes.c_double),
                ("nu", ctypes.c_double)]


def _lib():
    L = ctypes.CDLL(LIB)
    L.fmca_last_error.restype = ctypes.c_char_p
    return L


def _create(L, P, deg, upload=True):
    d, n = P.shape
    h = ctypes.c_void_p()
    desc = Desc(b"MATERN32", 0.5, 1.0, 1.0)
    st = L.fmca_h2_create_ex(ctypes.byref(desc), d, ctypes.c_int64(n),
                             P.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(n),
                             P.ctypes.data_as(ctypes.c_void_p), deg, ctypes.c_double(0.8),
                             0, 1, ctypes.byref(h))
    assert st == 0, L.fmca_last_error()
    if upload:
        assert L.fmca_h2_upload(h) == 0, L.fmca_last_error()
    return h


def _live(L):
    live, peak = ctypes.c_int64(), ctypes.c_int64()
    assert L.fmca_device_memory(ctypes.byref(live), ctypes.byref(peak), 0) == 0
    return live.value


def _call(L, fn, h, x, y, precision, stream):
    n, nrhs = x.shape[1], x.shape[0]   # x, y: (nrhs, rows) row-major = column-major (rows, nrhs)
    return getattr(L, fn)(h, ctypes.c_void_p(x.data_ptr()), ctypes.c_int64(x.shape[1]),
                          ctypes.c_int64(nrhs), ctypes.c_void_p(y.data_ptr()),
                          ctypes.c_int64(y.shape[1]), precision, 0, ctypes.c_void_p(stream))


@pytest.mark.parametrize("precision", [0, 1], ids=["FP64", "FP32"])
def test_noalloc_action_equals_the_allocating_one(precision):
    import torch
    L = _lib()
    torch.cuda.init()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(5)
    cases = [(20000, 16), (20000, 3), (6000, 1)]  # the later calls are smaller
    actors = {}
    for n, nrhs in cases:
        print(f"[case] noalloc precision={precision} n={n} nrhs={nrhs}", flush=True)
        if n not in actors:
            actors[n] = _create(L, np.asfortranarray(rng.random((3, n))), 3)
        h = actors[n]
        x = torch.randn(nrhs, n, dtype=torch.float64, device="cuda")
        y_ref = torch.empty(nrhs, n, dtype=torch.float64, device="cuda")
        y_na = torch.full((nrhs, n), float("nan"), dtype=torch.float64, device="cuda")
        assert _call(L, "fmca_h2_action_device", h, x, y_ref, precision, stream) == 0
        torch.cuda.synchronize()
        live0, torch0 = _live(L), torch.cuda.memory_allocated()
        st = _call(L, "fmca_h2_action_device_noalloc", h, x, y_na, precision, stream)
        assert st == 0, L.fmca_last_error()
        torch.cuda.synchronize()
        assert _live(L) == live0 and torch.cuda.memory_allocated() == torch0
        assert torch.equal(y_na, y_ref)
    for h in actors.values():
        L.fmca_h2_destroy(h)


def test_noalloc_refuses_oversize_and_unplanned_actors():
    import torch
    L = _lib()
    torch.cuda.init()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(6)
    P = np.asfortranarray(rng.random((3, 20000)))
    print("[case] noalloc oversize / not uploaded", flush=True)
    h = _create(L, P, 3)
    # FP64 at 120 rhs needs about 80 MiB of workspace (> 64 MiB): refused before any launch
    x = torch.zeros(120, 20000, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    assert _call(L, "fmca_h2_action_device_noalloc", h, x, y, 0, stream) == 3
    assert b"scratch" in L.fmca_last_error()
    L.fmca_h2_destroy(h)
    h = _create(L, P, 3, upload=False)
    assert _call(L, "fmca_h2_action_device_noalloc", h, x[:2], y[:2], 0, stream) == 1
    assert b"not on the device" in L.fmca_last_error()
    L.fmca_h2_destroy(h)
