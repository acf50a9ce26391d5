"""The host planner and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (g++ host build, tests/cpp/build_sanitize.sh): edge
point sets, both interpolators, the planner's permutation and block count
equal to the oracle's, and no sanitizer report."""
import os
import subprocess

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(HERE, "_build", "sanitize_check")
SOURCES = [os.path.join(HERE, "sanitize_check.cpp"),
           os.path.join(ROOT, "fmca_amd", "csrc", "plan.cpp"),
           os.path.join(ROOT, "fmca_amd", "csrc", "plan.hpp"),
           os.path.join(ROOT, "oracle", "h2_oracle.cpp")]


def test_planner_and_oracle_clean_under_asan_ubsan():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < max(map(os.path.getmtime, SOURCES)):
        subprocess.run(["bash", os.path.join(HERE, "build_sanitize.sh")], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "MISMATCH" not in r.stdout and "runtime error" not in r.stderr
    assert r.stdout.count(": ok") == 8
