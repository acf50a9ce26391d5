"""Host-side product checks that need no GPU: the planner reproduces the
oracle's (FMCA's) tree, boxes and block pattern exactly, the C-ABI library
exports everything include/fmca_amd.h declares, and invalid input is reported
through status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from conftest import ROOT, gpu_available, uniform_points

import fmca_amd

LIB = os.path.join(ROOT, "fmca_amd", "libfmca_amd.so")
HEADER = os.path.join(ROOT, "include", "fmca_amd.h")


def _block_set(r, c, lr):
    return set(zip(np.asarray(r).tolist(), np.asarray(c).tolist(), np.asarray(lr).astype(bool).tolist()))


CASES = [
    (10000, None, 3, 3, 0.8),
    (20000, None, 2, 5, 0.8),
    (6000, 4500, 3, 4, 0.8),
    (3000, None, 1, 4, 0.5),
    (5000, 2000, 2, 2, 0.3),
    (100, None, 3, 4, 0.8),
]


@pytest.mark.parametrize("n,m,d,deg,eta", CASES)
def test_planner_matches_oracle(n, m, d, deg, eta):
    P = uniform_points(n, d, seed=n + d)
    Pe = P if m is None else uniform_points(m, d, seed=m + 1)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), Pe, P, deg, eta)
    A = oracle.OracleH2(P, Pe, "MATERN32", 0.5, deg, eta)
    for which in (0, 1):
        assert np.array_equal(H.indices(which), A.indices(which))
        cp, co = H.clusters(which), A.clusters(which)
        for k in ("begin", "size", "level", "sons", "bbmin", "bbmax", "bbsize"):
            assert np.array_equal(cp[k], co[k]), k
    assert _block_set(*H.blocks()) == _block_set(*A.blocks())
    info = H.info()
    counts = A.entry_counts()
    assert info["near_entries"] == counts["near_entries"]
    assert info["far_entries"] == counts["far_entries"]
    assert info["far_blocks"] == counts["far_blocks"] and info["near_blocks"] == counts["near_blocks"]


def test_planner_ties_and_duplicates():
    rng = np.random.default_rng(11)
    P = np.round(rng.random((3, 6000)) * 40) / 40   # many equal coordinates
    P[:, 100:400] = P[:, 50:51]
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 0.3), P, P, 3, 0.8)
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.3, 3, 0.8)
    assert np.array_equal(H.indices(1), A.indices(1))
    assert _block_set(*H.blocks()) == _block_set(*A.blocks())


@pytest.mark.parametrize("kind", ["uniform", "grid-ties", "1d-ties"])
@pytest.mark.parametrize("scheme", ["tensor", "total_degree"])
def test_planner_fast_build_and_tie_fallback(kind, scheme):
    """The planner partitions with nth_element and sorts each leaf once; on
    equal coordinates at a split point or inside a leaf it rebuilds with
    FMCA's full sorts.  Either way the permutation and blocks are FMCA's."""
    rng = np.random.default_rng(9)
    if kind == "uniform":
        P = rng.random((3, 20000))
    elif kind == "grid-ties":
        P = rng.integers(0, 30, (3, 20000)).astype(float) / 30
    else:
        P = rng.integers(0, 500, (1, 5000)).astype(float) / 500
    P = np.asfortranarray(P)
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, 3, 0.8, scheme=scheme)
    interp = "tensor_product" if scheme == "tensor" else "total_degree"
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 0.5), P, P, 3, 0.8,
                               interp)
    assert np.array_equal(H.indices(0), A.indices(0))
    assert _block_set(*H.blocks()) == _block_set(*A.blocks())


def test_statistics_semantics():
    """H2MatrixBase::statistics (H2MatrixBase.h:70-105)."""
    P = uniform_points(8000, 3, seed=3)
    H = fmca_amd.H2Matrix(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 3, 0.8)
    q = H.interpolation()["q"]
    assert q == 20  # FMCA's binding: TotalDegreeInterpolator, binom(3 + 3, 3)
    st = H.statistics().ravel()
    r, c, lr = H.blocks()
    cl = H.clusters(0)
    mem = lr.sum() * q ** 2 + sum(cl["size"][t] * cl["size"][s] for t, s, l in zip(r, c, lr) if not l)
    assert st[0] == 8000 and st[1] == 8000
    assert st[2] == lr.sum() and st[3] == (~lr).sum()
    assert st[4] == round(mem / 8000) and st[5] == pytest.approx(mem * 8 / 1e9)


def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    text = open(HEADER).read()
    names = re.findall(r"^(?:fmca_status|const char \*)\s*(fmca_\w+)\s*\(", text, re.M)
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    lib.fmca_status_string.restype = ctypes.c_char_p
    assert lib.fmca_status_string(0) == b"ok"
    assert lib.fmca_status_string(2) == b"unknown kernel type"


def test_c_abi_create_plan_without_device():
    """The plan is host-side; create / queries / destroy work with plain pointers."""
    lib = ctypes.CDLL(LIB)

    class Desc(ctypes.Structure):
        _fields_ = [("ktype", ctypes.c_char_p), ("l", ctypes.c_double), ("c", ctypes.c_double),
                    ("nu", ctypes.c_double)]
    P = uniform_points(3000, 3, seed=1)
    h = ctypes.c_void_p()
    desc = Desc(b"matern32", 0.5, 1.0, 1.0)
    ptr = P.ctypes.data_as(ctypes.c_void_p)
    st = lib.fmca_h2_create(ctypes.byref(desc), 3, ctypes.c_int64(3000), ptr, ctypes.c_int64(3000),
                            ptr, 3, ctypes.c_double(0.8), ctypes.byref(h))
    assert st == 0 and h.value
    info = (ctypes.c_double * 13)()
    assert lib.fmca_h2_info(h, info) == 0
    assert info[0] == 3 and info[3] == 64 and info[12] == 1.0
    perm = np.empty(3000, dtype=np.uint32)
    assert lib.fmca_h2_indices(h, 0, perm.ctypes.data_as(ctypes.c_void_p)) == 0
    assert np.array_equal(np.sort(perm), np.arange(3000))
    assert lib.fmca_h2_destroy(h) == 0
    # status codes instead of FMCA's asserts
    bad = Desc(b"MATERNNU", 0.5, 1.0, 1.0)
    assert lib.fmca_h2_create(ctypes.byref(bad), 3, ctypes.c_int64(3000), ptr,
                              ctypes.c_int64(3000), ptr, 3, ctypes.c_double(0.8),
                              ctypes.byref(h)) == 2
    assert lib.fmca_h2_create(ctypes.byref(desc), 4, ctypes.c_int64(10), ptr, ctypes.c_int64(10),
                              ptr, 3, ctypes.c_double(0.8), ctypes.byref(h)) == 3


def test_c_abi_total_degree_plan_without_device():
    """fmca_h2_create_interp with FMCA_INTERP_TOTAL_DEGREE: q = binom(deg+d, d),
    the tree's minimum cluster size is that q, and fmca_h2_interpolation
    returns the oracle's DeMarchi nodes and invV."""
    lib = ctypes.CDLL(LIB)

    class Desc(ctypes.Structure):
        _fields_ = [("ktype", ctypes.c_char_p), ("l", ctypes.c_double), ("c", ctypes.c_double),
                    ("nu", ctypes.c_double)]
    P = uniform_points(3000, 3, seed=4)
    h = ctypes.c_void_p()
    desc = Desc(b"EXPONENTIAL", 0.5, 1.0, 1.0)
    ptr = P.ctypes.data_as(ctypes.c_void_p)
    assert lib.fmca_h2_create_interp(ctypes.byref(desc), 3, ctypes.c_int64(3000), ptr,
                                     ctypes.c_int64(3000), ptr, 4, ctypes.c_double(0.8), 1,
                                     ctypes.byref(h)) == 0
    kind, q = ctypes.c_int(), ctypes.c_int()
    assert lib.fmca_h2_interpolation(h, ctypes.byref(kind), ctypes.byref(q), None, None,
                                     None) == 0
    assert kind.value == 1 and q.value == 35
    mi = np.empty((35, 3), dtype=np.int32)
    Xi = np.empty((3, 35), order="F")
    invV = np.empty((35, 35), order="F")
    assert lib.fmca_h2_interpolation(h, None, None, mi.ctypes.data_as(ctypes.c_void_p),
                                     Xi.ctypes.data_as(ctypes.c_void_p),
                                     invV.ctypes.data_as(ctypes.c_void_p)) == 0
    omi, oXi, _, oinvV = oracle.interp_total_degree(3, 4)
    assert np.array_equal(mi, omi) and np.array_equal(Xi, oXi)
    assert np.abs(invV - oinvV).max() < 1e-13
    A = oracle.OracleH2(P, P, "EXPONENTIAL", 0.5, 4, 0.8, scheme="total_degree")
    perm = np.empty(3000, dtype=np.uint32)
    assert lib.fmca_h2_indices(h, 0, perm.ctypes.data_as(ctypes.c_void_p)) == 0
    assert np.array_equal(perm, A.indices(0))
    assert lib.fmca_h2_destroy(h) == 0
    # unknown interpolator id
    assert lib.fmca_h2_create_interp(ctypes.byref(desc), 3, ctypes.c_int64(3000), ptr,
                                     ctypes.c_int64(3000), ptr, 4, ctypes.c_double(0.8), 7,
                                     ctypes.byref(h)) == 1


def test_python_errors():
    P = uniform_points(500, 2, seed=1)
    K = fmca_amd.CovarianceKernel("GAUSSIAN", 0.3)
    with pytest.raises(ValueError):
        fmca_amd.CovarianceKernel("NOPE")
    with pytest.raises(TypeError):
        fmca_amd.H2Matrix(K, P.astype(np.float32), P)
    with pytest.raises(ValueError):
        fmca_amd.H2Matrix(K, P, P, 10, 0.8)      # degree > 9
    Q = P.copy()
    Q[0, 3] = np.nan
    with pytest.raises(ValueError):
        fmca_amd.H2Matrix(K, Q, Q, 3, 0.8)
    with pytest.raises(ValueError):
        fmca_amd.H2Matrix(fmca_amd.CovarianceKernel("GAUSSIAN", -1.0), P, P, 3, 0.8)
    H = fmca_amd.H2Matrix(K, P, P, 3, 0.8)
    with pytest.raises(TypeError):
        H.action(np.ones(500, dtype=np.float32))   # FMCA's noconvert(): float64 only
    with pytest.raises(ValueError):
        H.action(np.ones(499))


@pytest.mark.skipif(gpu_available(), reason="checks the no-device status")
def test_action_without_device_reports_status():
    P = uniform_points(500, 2, seed=1)
    H = fmca_amd.H2Matrix(fmca_amd.CovarianceKernel("GAUSSIAN", 0.3), P, P, 3, 0.8)
    with pytest.raises(RuntimeError, match="no HIP device"):
        H.action(np.ones(500))


def test_moment_order_table_matches_fmca():
    """The GPU keeps moments in tensor order; FMCA's graded-lex order maps onto it."""
    from itertools import product
    for d, deg in [(2, 3), (3, 4)]:
        _, _, mi, _ = oracle.interp(d, deg)
        P1 = deg + 1
        lin = [sum(a[i] * P1 ** i for i in range(d)) for a in mi]
        assert sorted(lin) == list(range(P1 ** d))
        assert lin[0] == 0 and lin[1:d + 1] == sorted(P1 ** i for i in range(d))[::-1]


@pytest.mark.parametrize("points", ["uniform", "clustered"])
def test_launch_orders_cover_every_item_once_with_balanced_xcd_pieces(points):
    """The coupling / leaf-output work lists: every target's far / near list
    covered exactly once by contiguous pieces, a target split only above three
    times the mean work (and then flagged), one far-field piece per leaf, -1 rows
    only for idle workgroups, and per tree level 8 contiguous XCD pieces of
    work within 1.35x of the level mean where a level has 256 items or more."""
    from fmca_amd import workloads as W
    n = 40000
    P = W.uniform(n, 3, 11) if points == "uniform" else W.gaussian_mixture(n, 3, 16, 0.02, 11)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 3, 0.8)
    cl = H.clusters(fmca_amd.ROWS)
    ncl = len(cl["level"])
    row, col, lr = H.blocks()
    nfar = np.bincount(row[lr], minlength=ncl)
    leaves = np.flatnonzero(cl["sons"][:, 0] < 0)
    for which in ("coupling", "leaf_output"):
        order = H.launch_order(which)
        assert len(order) % 8 == 0
        idle = order[:, 0] < 0
        assert np.all(order[idle] == -1)
        tasks = order[~idle]
        ntarget = ncl if which == "coupling" else len(leaves)
        assert np.array_equal(np.unique(tasks[:, 0]), np.arange(ntarget))
        # per target: pieces contiguous and covering its list; split flag iff several
        by = {}
        for t, b, e, f in tasks:
            by.setdefault(int(t), []).append((int(b), int(e), int(f)))
        for t, pieces in by.items():
            pieces.sort()
            assert all(pieces[k][1] == pieces[k + 1][0] for k in range(len(pieces) - 1))
            assert all(((f & 1) == 1) == (len(pieces) > 1) for _, _, f in pieces)
            if which == "coupling":
                assert pieces[-1][1] - pieces[0][0] == nfar[t]
            else:
                assert sum((f & 2) == 2 for _, _, f in pieces) == 1
        if points == "uniform":
            assert not (tasks[:, 3] & 1).any()  # uniform points: nothing to split
        if which != "coupling":
            continue
        xcd = np.flatnonzero(~idle) % 8
        work = tasks[:, 2] - tasks[:, 1] + 0.05
        lev = cl["level"][tasks[:, 0]]
        for L in np.unique(lev):
            sel = lev == L
            # contiguous pieces: the XCD index never decreases along the level
            pos = np.flatnonzero(sel)
            order_in_level = np.argsort(tasks[sel, 0] * 1000000 + tasks[sel, 1], kind="stable")
            assert np.all(np.diff(xcd[pos][order_in_level]) >= 0)
            if sel.sum() >= 256:
                w = np.bincount(xcd[sel], weights=work[sel], minlength=8)
                assert w.max() <= 1.35 * w.mean(), (L, w)


def test_variant_groups_argument_checks_without_device():
    """fmca_h2_action_variants_device validates its arguments before it needs
    a device: group count in [1, 8], variants in {0, 1, 2}, and nrhs filling
    the last 16-column group."""
    lib = ctypes.CDLL(LIB)

    class Desc(ctypes.Structure):
        _fields_ = [("ktype", ctypes.c_char_p), ("l", ctypes.c_double), ("c", ctypes.c_double),
                    ("nu", ctypes.c_double)]
    P = uniform_points(500, 2, seed=1)
    h = ctypes.c_void_p()
    desc = Desc(b"MATERN32", 0.5, 1.0, 1.0)
    ptr = P.ctypes.data_as(ctypes.c_void_p)
    assert lib.fmca_h2_create(ctypes.byref(desc), 2, ctypes.c_int64(500), ptr,
                              ctypes.c_int64(500), ptr, 3, ctypes.c_double(0.8),
                              ctypes.byref(h)) == 0
    buf = (ctypes.c_double * 1)()

    def call(variants, nrhs):
        v = (ctypes.c_int * max(1, len(variants)))(*variants)
        return lib.fmca_h2_action_variants_device(h, v, len(variants), 0, buf, ctypes.c_int64(500),
                                                  ctypes.c_int64(nrhs), buf, ctypes.c_int64(500),
                                                  0, None)
    assert call([0] * 9, 130) == 1       # more than 8 groups
    assert call([], 1) == 1              # no group
    assert call([0, 3], 20) == 1         # unknown variant
    assert call([0, 1], 33) == 1         # more columns than groups
    assert call([0, 1], 16) == 1         # last group empty
    assert lib.fmca_h2_destroy(h) == 0


def test_planner_choice_without_device():
    """planner="auto" plans on the host when no GPU is visible; "device" then
    fails with a status (no crash); the C ABI rejects an unknown planner."""
    P = uniform_points(3000, 3, seed=4)
    ker = fmca_amd.CovarianceKernel("MATERN32", 0.5)
    H = fmca_amd.H2kernelActor(ker, P, P, 3, 0.8)
    rep = H.plan_report()
    assert rep["rows_tree"] == rep["cols_tree"] == rep["blocks"] == "host"
    assert rep["tree_s"] >= 0 and rep["pattern_s"] >= 0
    with pytest.raises(RuntimeError, match="device"):
        fmca_amd.H2kernelActor(ker, P, P, 3, 0.8, planner="device")
    with pytest.raises(ValueError):
        fmca_amd.H2kernelActor(ker, P, P, 3, 0.8, planner="gpu")
    lib = ctypes.CDLL(LIB)

    class Desc(ctypes.Structure):
        _fields_ = [("ktype", ctypes.c_char_p), ("l", ctypes.c_double), ("c", ctypes.c_double),
                    ("nu", ctypes.c_double)]
    desc = Desc(b"matern32", 0.5, 1.0, 1.0)
    h = ctypes.c_void_p()
    ptr = P.ctypes.data_as(ctypes.c_void_p)
    args = (ctypes.byref(desc), 3, ctypes.c_int64(3000), ptr, ctypes.c_int64(3000), ptr, 3,
            ctypes.c_double(0.8), 0)
    assert lib.fmca_h2_create_ex(*args, 7, ctypes.byref(h)) == 1
    assert lib.fmca_h2_create_ex(*args, 0, ctypes.byref(h)) == 0 and h.value
    where = (ctypes.c_int * 3)(9, 9, 9)
    assert lib.fmca_h2_plan_report(h, None, None, where) == 0 and list(where) == [0, 0, 0]
    assert lib.fmca_h2_destroy(h) == 0


def test_assemble_without_device_reports_status():
    """fmca_h2_assemble needs the device: without one it returns a status
    (no crash), and the actor stays usable for host queries."""
    P = uniform_points(2000, 3, seed=8)
    H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("MATERN32", 0.5), P, P, 3, 0.8)
    with pytest.raises(RuntimeError):
        H.assemble(fmca_amd.FP64)
    H.release_assembled()
    assert H.info()["n"] == 2000
    lib = ctypes.CDLL(LIB)
    assert lib.fmca_h2_assemble(None, 0, None) == 1
    assert lib.fmca_h2_release_assembled(None) == 1


@pytest.mark.parametrize("hand", ["HAND_2D", "HAND_3D"])
def test_planner_split_axis_rule_by_hand(hand):
    """The product planner (host and, with a GPU, device) reproduces the
    hand-worked 2-D / 3-D trees of test_oracle.py, which pin the split-box
    longest-edge rule with first-axis ties (ClusterTreeSplitter.h:60-81)."""
    import test_oracle
    h = getattr(test_oracle, hand)
    P = np.asfortranarray(np.array(h["P"], dtype=np.float64).T)
    planners = ["host"] + (["device"] if gpu_available() else [])
    for planner in planners:
        H = fmca_amd.H2kernelActor(fmca_amd.CovarianceKernel("EXPONENTIAL", 1.0), P, P, 0, 0.8,
                                   planner=planner)
        assert H.indices(1).tolist() == h["perm"], planner
        cl = H.clusters(1)
        for k in ("size", "begin", "level", "sons"):
            assert np.asarray(cl[k]).tolist() == h[k], (planner, k)


@pytest.mark.parametrize("d,deg,n,m", [(2, 3, 3000, None), (3, 2, 2500, 2000), (1, 6, 2000, None)])
def test_projector_dual_bases_meet_fmcas_bar(d, deg, n, m):
    """fmca_h2_project computes the inverse Gram matrices of the cluster bases
    on the host before it needs the device; FMCA's check of its dual cluster
    bases (compute_dual_cluster_bases_impl.h:66-86) must stay < 5e-10 on both
    trees.  Without a device the call then reports FMCA_ERR_NO_DEVICE."""
    lib = ctypes.CDLL(LIB)
    lib.fmca_last_error.restype = ctypes.c_char_p

    class Desc(ctypes.Structure):
        _fields_ = [("ktype", ctypes.c_char_p), ("l", ctypes.c_double), ("c", ctypes.c_double),
                    ("nu", ctypes.c_double)]
    P = uniform_points(n, d, seed=71)
    Pe = P if m is None else uniform_points(m, d, seed=72)
    mm = Pe.shape[1]
    M = np.asfortranarray(np.random.default_rng(73).standard_normal((mm, n)))
    desc = Desc(b"EXPONENTIAL", 0.5, 1.0, 1.0)
    h = ctypes.c_void_p()
    st = lib.fmca_h2_create_ex(ctypes.byref(desc), d, ctypes.c_int64(mm),
                               Pe.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(n),
                               P.ctypes.data_as(ctypes.c_void_p), deg, ctypes.c_double(0.8), 0, 0,
                               ctypes.byref(h))
    assert st == 0
    chk = (ctypes.c_double * 2)(-1.0, -1.0)
    st = lib.fmca_h2_project(h, M.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(mm), 0, chk)
    assert st in (0, 4), lib.fmca_last_error()  # ok, or no device after the host part
    assert 0 <= chk[0] < 5e-10 and 0 <= chk[1] < 5e-10, (chk[0], chk[1])
    lib.fmca_h2_destroy(h)
