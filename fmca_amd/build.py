"""Builds the in-tree shared libraries of fmca_amd for gfx950.

  libfmca_amd.so        C-ABI library (host planner + HIP kernels), hipcc
  _fmca_amd_py*.so      pybind11 module over the C ABI

Run ``python fmca_amd/build.py`` (or ``__graft_entry__.build()``).  Objects are
rebuilt every time; compilation of independent sources runs in parallel.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libfmca_amd.so")
ARCH = os.environ.get("FMCA_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# hipcc flags: gfx950 only; packed FP32 instructions are disabled
# (-packed-fp32-ops) so that float code is issued as plain VALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-mcode-object-version=5", "-pthread",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}",
              "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

SOURCES_HIP = ["kernels.hip", "kernels_mfma.hip", "kernels_bf16.hip", "kernels_grad.hip",
               "blas1.hip", "tree_device.hip", "assembled.hip", "pattern_device.hip",
               "samplets.hip"]
SOURCES_CPP = ["plan.cpp", "actor.cpp", "projector.cpp", "samplets.cpp", "samplets_api.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return r


def _objects(verbose):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in SOURCES_HIP:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o",
                     os.path.join(BUILD, s + ".o")])
    for s in SOURCES_CPP:
        jobs.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, s), "-o",
                     os.path.join(BUILD, s + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    return [j[-1] for j in jobs]


def build_library(verbose=False):
    objs = _objects(verbose)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-pthread", "-o", LIB, *objs,
          "-Wl,-rpath,/opt/rocm/lib"], verbose)
    return LIB


def build_python_module(verbose=False):
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_fmca_amd_py" + ext)
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", *HOST_FLAGS, "-shared", f"-I{pybind11.get_include()}", f"-I{py_inc}",
          os.path.join(CSRC, "pymodule.cpp"), "-o", out, f"-L{HERE}", "-lfmca_amd",
          "-Wl,-rpath,$ORIGIN"], verbose)
    return out


def build(verbose=False):
    build_library(verbose)
    build_python_module(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
