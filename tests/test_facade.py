"""The C++ facade (include/fmca_amd/H2kernelActor.hpp) used as FMCA code uses
FMCA::H2kernelActor: constructor (kernel, Peval, P, deg, eta), action(rhs)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, rel_err, uniform_points

HERE = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(HERE, "_build", "facade_check")


def _binary():
    src = [os.path.join(HERE, "facade_check.cpp"),
           os.path.join(ROOT, "include", "fmca_amd", "H2kernelActor.hpp")]
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < max(map(os.path.getmtime, src)):
        subprocess.run(["bash", os.path.join(HERE, "build.sh")], check=True)
    return BIN


def _write(path, P, rhs, deg, eta, kernel, l):
    dim, n = P.shape
    with open(path, "wb") as f:
        np.array([dim, n, rhs.shape[1], deg], dtype=np.int64).tofile(f)
        np.array([eta, l], dtype=np.float64).tofile(f)
        f.write(kernel.encode().ljust(16, b"\0"))
        np.asfortranarray(P).ravel(order="F").tofile(f)
        np.asfortranarray(rhs).ravel(order="F").tofile(f)


def test_facade_plan_matches_oracle(tmp_path):
    P = uniform_points(5000, 3, seed=21)
    rhs = np.zeros((5000, 1))
    _write(tmp_path / "in.bin", P, rhs, 3, 0.8, "matern32", 0.5)
    r = subprocess.run([_binary(), "plan", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    idx = np.frombuffer(raw[:4 * 5000], dtype=np.uint32)
    st = np.frombuffer(raw[4 * 5000:], dtype=np.float64)
    A = oracle.OracleH2(P, P, "MATERN32", 0.5, 3, 0.8)
    assert np.array_equal(idx, A.indices(1))
    counts = A.entry_counts()
    assert st[2] == counts["far_blocks"] and st[3] == counts["near_blocks"]


def test_facade_total_degree_moments_select_total_degree(tmp_path):
    """A Moments argument shaped like NystromMoments<TotalDegreeInterpolator>
    gives the total-degree plan (leaves of q = 20 .. 39 points at deg 3)."""
    P = uniform_points(5000, 3, seed=21)
    _write(tmp_path / "in.bin", P, np.zeros((5000, 1)), 3, 0.8, "matern32", 0.5)
    r = subprocess.run([_binary(), "plan_td", str(tmp_path / "in.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    idx = np.frombuffer((tmp_path / "out.bin").read_bytes()[:4 * 5000], dtype=np.uint32)
    A = oracle.OracleH2(P, P, "MATERN32", 0.5, 3, 0.8, scheme="total_degree")
    assert np.array_equal(idx, A.indices(1))
    assert not np.array_equal(idx, oracle.OracleH2(P, P, "MATERN32", 0.5, 3, 0.8).indices(1))


def test_facade_reports_errors(tmp_path):
    P = uniform_points(100, 3, seed=1)
    _write(tmp_path / "in.bin", P, np.zeros((100, 1)), 3, 0.8, "NOSUCHKERNEL", 0.5)
    r = subprocess.run([_binary(), "plan", str(tmp_path / "in.bin"), str(tmp_path / "o.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 10 + 2 and "unknown kernel" in r.stderr


@pytest.mark.gpu
def test_facade_action_matches_oracle(tmp_path):
    P = uniform_points(6000, 3, seed=22)
    rhs = np.asfortranarray(np.random.default_rng(23).standard_normal((6000, 5)))
    _write(tmp_path / "in.bin", P, rhs, 4, 0.8, "exponential", 0.4)
    r = subprocess.run([_binary(), "action", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(5, 6000).T
    ref = oracle.OracleH2(P, P, "EXPONENTIAL", 0.4, 4, 0.8).action(rhs)
    assert rel_err(y, ref) <= 1e-12


@pytest.mark.gpu
def test_facade_total_degree_action_matches_oracle(tmp_path):
    P = uniform_points(5000, 2, seed=24)
    rhs = np.asfortranarray(np.random.default_rng(25).standard_normal((5000, 3)))
    _write(tmp_path / "in.bin", P, rhs, 4, 0.8, "exponential", 1.0)
    r = subprocess.run([_binary(), "action_td", str(tmp_path / "in.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(3, 5000).T
    ref = oracle.OracleH2(P, P, "EXPONENTIAL", 1.0, 4, 0.8, scheme="total_degree").action(rhs)
    assert rel_err(y, ref) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode,scheme", [("action_dev", "tensor"), ("action_td_dev", "total_degree")])
def test_facade_device_planner_and_assembled_blocks(tmp_path, mode, scheme):
    """Planner::Device at construction and assemble() before the action: the
    same result as FMCA's restated action."""
    P = uniform_points(6000, 3, seed=26)
    rhs = np.asfortranarray(np.random.default_rng(27).standard_normal((6000, 4)))
    _write(tmp_path / "in.bin", P, rhs, 3, 0.8, "matern32", 0.5)
    r = subprocess.run([_binary(), mode, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(4, 6000).T
    ref = oracle.OracleH2(P, P, "MATERN32", 0.5, 3, 0.8, scheme=scheme).action(rhs)
    assert rel_err(y, ref) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["action_geo", "action_geo_set"])
def test_facade_geodesic_distance_matches_oracle(tmp_path, mode):
    """setDistanceType("GEODESIC") through the facade: on its own kernel
    class, and on the actor for a kernel class that does not carry a distance
    (FMCA's keeps it private)."""
    P = np.random.default_rng(28).standard_normal((3, 6000))
    P /= np.linalg.norm(P, axis=0)
    rhs = np.asfortranarray(np.random.default_rng(29).standard_normal((6000, 3)))
    _write(tmp_path / "in.bin", P, rhs, 3, 0.8, "matern32", 0.5)
    r = subprocess.run([_binary(), mode, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(3, 6000).T
    ref = oracle.OracleH2(P, P, "MATERN32", 0.5, 3, 0.8, distance="GEODESIC").action(rhs)
    assert rel_err(y, ref) <= 1e-12


_MOMENTS_SNIPPET = r"""
#include "fmca_amd/H2kernelActor.hpp"
namespace FMCA {
class TensorProductInterpolator {};
class TotalDegreeInterpolator {};
class MonomialInterpolator {};
template <typename I> class NystromMoments {};
}
struct K { std::string kernelType() const { return "EXPONENTIAL"; }
           double l() const { return 1; } double c() const { return 1; }
           double nu() const { return 1; } };
int main() { return sizeof(fmca_amd::H2kernelActor<K, MOMENTS>); }
"""


@pytest.mark.parametrize("moments,interp", [
    ("FMCA::NystromMoments<FMCA::TensorProductInterpolator>", 0),
    ("FMCA::NystromMoments<FMCA::TotalDegreeInterpolator>", 1),
    ("fmca_amd::TensorProductMoments", 0),
    ("fmca_amd::TotalDegreeMoments", 1),
    ("FMCA::NystromMoments<FMCA::MonomialInterpolator>", None),
    ("int", None),
])
def test_facade_moments_type_is_checked(tmp_path, moments, interp):
    """The facade maps FMCA's Moments type to the interpolator it was
    instantiated with (NystromMoments.h:16-17) and refuses to compile any
    other: NystromMoments<MonomialInterpolator> (MonomialInterpolator.h:20)
    has no counterpart here and must not silently run another basis."""
    src = tmp_path / "m.cpp"
    src.write_text(_MOMENTS_SNIPPET.replace("MOMENTS", moments)
                   + (f"static_assert(fmca_amd::moments_interpolator<{moments}>::value == "
                      f"{interp}, \"id\");\n" if interp is not None else ""))
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    if interp is None:
        assert r.returncode != 0 and "Moments must be" in r.stderr
    else:
        assert r.returncode == 0, r.stderr
