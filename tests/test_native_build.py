"""The native libraries exist, load, and export every C-ABI symbol that
include/fundem_amd.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from fundem_amd import build

HEADER = os.path.join(build.ROOT, "include", "fundem_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdm_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_launchers():
    syms = declared_symbols()
    for s in ("fdm_launchNarrowPhase", "fdm_launchBroadPhase", "fdm_radixSortPairs",
              "fdm_launchReduceForceTorque", "fdm_launchBondForces", "fdm_launchPeriodicGhosts"):
        assert s in syms


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_core_library_exports_every_declared_symbol(precision):
    core, _ = build.libs(precision)
    if not os.path.exists(core):
        pytest.skip("core library not built")
    lib = ctypes.CDLL(core, mode=ctypes.RTLD_LOCAL | os.RTLD_DEEPBIND)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.parametrize("precision,nbytes", [("fp32", 4), ("fp64", 8)])
def test_solver_binding_stores_state_in_its_precision(precision, nbytes):
    _, solver_path = build.libs(precision)
    if not os.path.exists(solver_path):
        pytest.skip("solver library not built")
    from fundem_amd import solver
    lib = solver.solver_lib(precision)
    assert lib.fdmx_state_bytes() == nbytes


@pytest.mark.skipif(not os.path.exists(build.SOLVER_LIB), reason="solver library not built")
def test_solver_binding_loads():
    from fundem_amd import solver
    lib = solver.solver_lib()
    assert hasattr(lib, "fdmx_advance")
