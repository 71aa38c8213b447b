"""Build recipe for the native libraries (in-tree, gfx950).

  fundem_amd/lib/libfundem_amd.so          HIP kernels + the C-ABI of include/fundem_amd.h
  fundem_amd/lib/libfundem_solver.so       C binding of the C++ host API (LSSolver)
  fundem_amd/lib/libfundem_amd_fp64.so     the same two with FP64 storage of the state a
  fundem_amd/lib/libfundem_solver_fp64.so  step carries over (FDM_STATE_FP64)
  tests/apps/build/tutorial_app            a tutorial-style application (test program)

Device code is compiled without floating-point contraction (each FP64
operation rounds once, in source order) and without packed FP32 arithmetic.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "lib")
CORE_LIB = os.path.join(LIB_DIR, "libfundem_amd.so")
SOLVER_LIB = os.path.join(LIB_DIR, "libfundem_solver.so")
CORE_LIB64 = os.path.join(LIB_DIR, "libfundem_amd_fp64.so")
SOLVER_LIB64 = os.path.join(LIB_DIR, "libfundem_solver_fp64.so")


def libs(precision: str = "fp32"):
    """(core, solver) library paths of a storage precision."""
    if precision == "fp32":
        return CORE_LIB, SOLVER_LIB
    if precision == "fp64":
        return CORE_LIB64, SOLVER_LIB64
    raise ValueError(f"unknown precision {precision!r}")

ARCH = os.environ.get("FDM_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def _defines(precision):
    return ["-DFDM_STATE_FP64"] if precision == "fp64" else []


def build_core(precision: str = "fp32") -> str:
    core, _ = libs(precision)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))
    tmp = core + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, *_defines(precision), "-shared", "-I", os.path.join(ROOT, "include"),
          *srcs, "-o", tmp])
    os.replace(tmp, core)
    return core


def build_solver(precision: str = "fp32") -> str:
    """The host API is plain C++ over the HIP runtime: built by the host
    compiler, linked against the core library of the same precision."""
    core, solver = libs(precision)
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = solver + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    core_name = os.path.basename(core)[3:-3]  # libX.so -> X
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared",
          "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-Wno-unused-result", *_defines(precision),
          "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
          os.path.join(PKG, "cpp", "solver_capi.cpp"), "-L", LIB_DIR, "-l" + core_name,
          "-L", os.path.join(rocm, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib"), "-o", tmp])
    os.replace(tmp, solver)
    return solver


APP_SRC = os.path.join(ROOT, "tests", "apps", "tutorial_app.cpp")
APP_BIN = os.path.join(ROOT, "tests", "apps", "build", "tutorial_app")


def build_test_app() -> str:
    """A tutorial-style application of the C++ host API (test program):
    built like the binding, against the headers and the core library."""
    os.makedirs(os.path.dirname(APP_BIN), exist_ok=True)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    tmp = APP_BIN + ".tmp"
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-ffp-contract=off",
          "-D__HIP_PLATFORM_AMD__", "-Wno-unused-result",
          "-I", os.path.join(PKG, "include"), "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(rocm, "include"), APP_SRC, "-L", LIB_DIR, "-lfundem_amd",
          "-L", os.path.join(rocm, "lib"), "-lamdhip64",
          "-Wl,-rpath," + LIB_DIR, "-Wl,-rpath,$ORIGIN/../../../fundem_amd/lib",
          "-Wl,-rpath," + os.path.join(rocm, "lib"), "-o", tmp])
    os.replace(tmp, APP_BIN)
    return APP_BIN


def build_all() -> None:
    for precision in ("fp32", "fp64"):
        build_core(precision)
        build_solver(precision)
    build_test_app()


if __name__ == "__main__":
    build_all()
    for path in (CORE_LIB, SOLVER_LIB, CORE_LIB64, SOLVER_LIB64, APP_BIN):
        print(path)
