"""The hand-written reverse-mode routines of the adjoint step
(fundem_amd/csrc/adjoint_math.hpp) against forward-mode dual numbers, on the
host: for random inputs, input directions and output adjoints, <yb, J dx>
from dual numbers through the primal routine must equal <xb, dx> from the
reverse routine, for every routine (rotations, tensor rotation, the angular
kick, quaternion integration, contact geometry, the contact law in its stick,
slip and frictionless branches, and a whole node-contact)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_vjp_routines_match_dual_numbers(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "adjoint_vjp_check"
    subprocess.run([cxx, "-O1", "-std=c++17", "-D__host__=", "-D__device__=",
                    os.path.join(HERE, "adjoint_vjp_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "failures 0" in r.stdout
    for branch in ("stick", "slip", "no spring"):
        assert f"branch {branch}" in r.stdout
