"""The cluster trees built on the GPU (planner="device") against the host
planner, which is itself pinned bit for bit against the oracle's restatement
of FMCA's ClusterTreeInitializer (tests/test_planner.py).

Bar: bit-exact.  The permutation, every cluster's range, level and sons, and
the boxes compared as raw float64 bits; the block pattern as a set; the FP64
action of both actors within 1e-14 (the same plan; split targets add into
their output with atomics, so the last bits may differ).  Equal coordinates at
a split point or inside a leaf, and -0.0 coordinates, must send the build to
the host ("host-fallback") with the same result.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("begin", "size", "level", "sons")
BOXES = ("bbmin", "bbmax", "bbsize")


def _points(kind, n, d, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        P = rng.random((d, n))
    elif kind == "signed":       # negative coordinates and an offset axis
        P = rng.uniform(-1, 1, (d, n))
        P[0] += 1e3
    elif kind == "mixture":      # 16 tight Gaussian clusters
        c = rng.random((d, 16))
        P = c[:, rng.integers(0, 16, n)] + 0.02 * rng.standard_normal((d, n))
    elif kind == "anisotropic":  # one long axis: the split axis stays the same
        P = rng.random((d, n)) * np.array([100.0, 1.0, 0.01][:d])[:, None]
    elif kind == "grid-ties":    # equal coordinates everywhere
        P = rng.integers(0, 25, (d, n)).astype(float) / 25
    elif kind == "negzero":      # one -0.0 coordinate among distinct values
        P = rng.uniform(-1, 1, (d, n))
        P[0, n // 3] = -0.0
    return np.asfortranarray(P)


def _same_tree(Hd, Hh, which):
    assert np.array_equal(Hd.indices(which), Hh.indices(which))
    cd, ch = Hd.clusters(which), Hh.clusters(which)
    for k in FIELDS:
        assert np.array_equal(cd[k], ch[k]), k
    for k in BOXES:
        assert np.array_equal(np.ascontiguousarray(cd[k]).view(np.uint64),
                              np.ascontiguousarray(ch[k]).view(np.uint64)), k


def _same_blocks(Hd, Hh):
    for a, b in zip(Hd.blocks(), Hh.blocks()):
        assert np.array_equal(a, b)


CASES = [
    # kind, n, m (None: same points), d, deg, interpolator
    ("uniform", 10_000, None, 3, 3, "tensor_product"),
    ("uniform", 50, None, 3, 3, "tensor_product"),         # root is the only leaf
    ("uniform", 128, None, 3, 3, "tensor_product"),        # n = 2q: one split
    ("uniform", 40_000, None, 2, 5, "tensor_product"),
    ("uniform", 5_000, None, 1, 6, "tensor_product"),
    ("signed", 30_000, 20_000, 3, 4, "tensor_product"),    # two trees on the device
    ("mixture", 100_000, None, 3, 4, "total_degree"),
    ("anisotropic", 60_000, None, 3, 2, "tensor_product"),
    ("uniform", 300_000, None, 3, 4, "tensor_product"),    # 74 chunks per root segment
    ("uniform", 20_000, None, 3, 9, "tensor_product"),     # q = 1000: leaves up to 1999 points
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}-d{c[3]}-deg{c[4]}" for c in CASES])
def test_device_tree_equals_host_tree(case):
    import fmca_amd
    kind, n, m, d, deg, interp = case
    print(f"[case] device tree {kind} n={n} m={m} d={d} deg={deg} {interp}", flush=True)
    P = _points(kind, n, d, 1)
    Pe = P if m is None else _points(kind, m, d, 2)
    ker = fmca_amd.CovarianceKernel("MATERN32", 0.4)
    Hh = fmca_amd.H2kernelActor(ker, Pe, P, deg, 0.8, interp)
    Hd = fmca_amd.H2kernelActor(ker, Pe, P, deg, 0.8, interp, planner="device")
    rep = Hd.plan_report()
    print(f"  {rep}", flush=True)
    assert rep["rows_tree"] == "device" and rep["cols_tree"] == "device"
    assert rep["blocks"] == "device"
    for which in (0, 1):
        _same_tree(Hd, Hh, which)
    _same_blocks(Hd, Hh)
    if n <= 100_000 and deg <= 6:
        x = np.asfortranarray(np.random.default_rng(3).standard_normal((n, 8)))
        assert rel_err(Hd.action(x), Hh.action(x)) <= 1e-14


@pytest.mark.parametrize("kind", ["grid-ties", "negzero"])
def test_device_tree_falls_back_to_host_on_ties(kind):
    import fmca_amd
    print(f"[case] device tree fallback {kind}", flush=True)
    P = _points(kind, 20_000, 3, 5)
    ker = fmca_amd.CovarianceKernel("EXPONENTIAL", 0.5)
    Hh = fmca_amd.H2kernelActor(ker, P, P, 3, 0.8)
    Hd = fmca_amd.H2kernelActor(ker, P, P, 3, 0.8, planner="device")
    assert Hd.plan_report()["rows_tree"] == "host-fallback"
    _same_tree(Hd, Hh, 0)
    _same_blocks(Hd, Hh)


@pytest.mark.parametrize("eta,where", [(2.5, "host-fallback"), (0.8, "device"), (0.0, "device")])
def test_device_blocks_on_an_integer_grid(eta, where):
    """Points 0, 1, ..., n-1 on a line: boxes, radii and distances are
    (half-)integers.  At eta = 2.5, 2,046 of the pairs the expansion tests have
    max(r_t, r_s) == eta * dist exactly (counted on the CPU from the host
    tree), so the device expansion must hand over to the host.  At eta = 0.8
    none does, and eta = 0 (MultipoleFunctionEvaluator's) never can."""
    import fmca_amd
    print(f"[case] device blocks integer grid eta={eta}", flush=True)
    P = np.asfortranarray(np.arange(6000, dtype=float)[None, :])
    ker = fmca_amd.CovarianceKernel("EXPONENTIAL", 50.0)
    Hh = fmca_amd.H2kernelActor(ker, P, P, 3, eta)
    Hd = fmca_amd.H2kernelActor(ker, P, P, 3, eta, planner="device")
    rep = Hd.plan_report()
    print(f"  {rep}", flush=True)
    assert rep["rows_tree"] == "device" and rep["blocks"] == where
    _same_tree(Hd, Hh, 0)
    _same_blocks(Hd, Hh)


def test_device_tree_at_one_million_points():
    """The tree of the 1e6-point workload (U[0,1]^3, degree 4): 8,191
    clusters, depth 12, built on the GPU, equal to the host's."""
    import fmca_amd
    from fmca_amd import workloads as W
    print("[case] device tree n=1e6", flush=True)
    P = W.uniform(1_000_000, 3, 0)
    ker = fmca_amd.CovarianceKernel("MATERN32", 0.5)
    Hh = fmca_amd.H2kernelActor(ker, P, P, 4, 0.8)
    Hd = fmca_amd.H2kernelActor(ker, P, P, 4, 0.8, planner="device")
    print(f"  host {Hh.plan_report()}  device {Hd.plan_report()}", flush=True)
    assert Hd.plan_report()["rows_tree"] == "device"
    assert Hd.plan_report()["blocks"] == "device"
    _same_tree(Hd, Hh, 0)
    _same_blocks(Hd, Hh)
    assert Hd.info()["far_blocks"] == Hh.info()["far_blocks"]
    assert Hd.info()["near_blocks"] == Hh.info()["near_blocks"]
