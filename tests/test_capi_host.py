"""Host-side behaviour of the C-ABI (include/fundem_amd.h) that needs no GPU:
argument validation returns FDM_EINVAL before anything is launched, the
process-wide launch tuning accepts only the documented shapes, and the sector
rotation constants equal the reference's host arithmetic (rotateVector,
myVec.h:143-181; rotateQuaternion, myQua.h:131-175) evaluated here with the
same C library."""
import ctypes as C
import math
import os

import pytest

from fundem_amd import build

pytestmark = pytest.mark.skipif(not os.path.exists(build.CORE_LIB), reason="core library not built")
EINVAL = -1


class Tuning(C.Structure):
    _fields_ = [("sweepWaves", C.c_int), ("sweepWavesPerCU", C.c_int), ("lawBlock", C.c_int),
                ("reduceBlock", C.c_int), ("lawMode", C.c_int), ("nodesPerLane", C.c_int),
                ("wallGroups", C.c_int)]


class Periodic(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("sector", C.c_int), ("pad_", C.c_int),
                ("angle", C.c_double * 3), ("A", C.c_double * 3), ("B", C.c_double * 3),
                ("halfCos", C.c_double * 3), ("halfSin", C.c_double * 3)]


@pytest.fixture(scope="module")
def lib():
    return C.CDLL(build.CORE_LIB)


def test_null_arguments_are_rejected(lib):
    for name, args in (("fdm_launchNarrowPhase", (None, None)),
                       ("fdm_launchNarrowPhaseSets", (None, 0, None, None)),
                       ("fdm_launchBroadPhase", (None, None, C.c_size_t(0), None)),
                       ("fdm_launchReduceForceTorque", (None, None)),
                       ("fdm_launchStepBegin", (None, None, None, C.c_size_t(0), None)),
                       ("fdm_launchAdjointStep", (None, None, C.c_size_t(0), None)),
                       ("fdm_setTuning", (None,))):
        assert getattr(lib, name)(*args) == EINVAL, name


def test_mesh_distance_arguments(lib):
    f = lib.fdm_launchMeshSignedDistanceGrid
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                  C.c_void_p]
    origin = (C.c_double * 3)(0, 0, 0)
    size = (C.c_int * 3)(4, 4, 4)
    dummy = C.c_void_p(16)  # never dereferenced: validation fails first
    assert f(None, 1, origin, 0.1, size, 0, dummy, None) == EINVAL
    assert f(dummy, 0, origin, 0.1, size, 0, dummy, None) == EINVAL
    assert f(dummy, 1, origin, 0.0, size, 0, dummy, None) == EINVAL
    bad = (C.c_int * 3)(4, 0, 4)
    assert f(dummy, 1, origin, 0.1, bad, 0, dummy, None) == EINVAL
    g = lib.fdm_launchMeshSignedDistancePoints
    g.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    assert g(dummy, 1, None, 0, 0, None, None, None) == 0  # nothing to do
    assert g(dummy, 1, None, 5, 0, None, None, None) == EINVAL
    assert g(None, 1, dummy, 5, 0, dummy, dummy, None) == EINVAL


def test_tuning_accepts_only_documented_shapes(lib):
    lib.fdm_setTuning.argtypes = [C.POINTER(Tuning)]
    lib.fdm_getTuning.argtypes = [C.POINTER(Tuning)]
    before = Tuning()
    lib.fdm_getTuning(C.byref(before))
    try:
        for bad in ((3, 0, 0, 0, 0), (0, 2, 0, 0, 0), (0, 200, 0, 0, 0), (0, 0, 96, 0, 0),
                    (0, 0, 0, 100, 0), (0, 0, 0, 2048, 0), (0, 0, 0, 0, 3), (0, 0, 0, 0, -1),
                    (0, 0, 0, 0, 0, 4), (0, 0, 0, 0, 0, -2), (0, 0, 0, 0, 0, 0, 3),
                    (0, 0, 0, 0, 0, 0, -1)):
            assert lib.fdm_setTuning(C.byref(Tuning(*bad))) == EINVAL, bad
        got = Tuning()
        lib.fdm_getTuning(C.byref(got))
        assert (got.sweepWaves, got.sweepWavesPerCU, got.lawBlock, got.reduceBlock) == \
            (before.sweepWaves, before.sweepWavesPerCU, before.lawBlock, before.reduceBlock)
        assert lib.fdm_setTuning(C.byref(Tuning(2, 0, 128, 0, 0))) == 0  # zero keeps a field
        lib.fdm_getTuning(C.byref(got))
        assert (got.sweepWaves, got.sweepWavesPerCU, got.lawBlock, got.reduceBlock, got.lawMode) == \
            (2, before.sweepWavesPerCU, 128, before.reduceBlock, before.lawMode)
        # a caller's own shapes: zero fields from the default, invalid ones refused
        lib.fdm_checkTuning.argtypes = [C.POINTER(Tuning)]
        mine = Tuning(0, 64, 0, 0, 2, 2, 1)
        assert lib.fdm_checkTuning(C.byref(mine)) == 0
        assert (mine.sweepWaves, mine.sweepWavesPerCU, mine.lawBlock, mine.lawMode,
                mine.nodesPerLane, mine.wallGroups) == (2, 64, 128, 2, 2, 1)
        assert lib.fdm_checkTuning(C.byref(Tuning(0, 0, 0, 0, 5))) == EINVAL
        lib.fdm_getTuning(C.byref(got))
        assert got.sweepWavesPerCU == before.sweepWavesPerCU and got.lawMode == before.lawMode
    finally:
        lib.fdm_setTuning(C.byref(before))


def test_sector_rotation_constants(lib):
    lib.fdm_sectorPeriodic.argtypes = [C.POINTER(Periodic)]
    p = Periodic()
    p.x = 1
    lib.fdm_sectorPeriodic(C.byref(p))
    assert (p.x, p.y, p.sector) == (0, 0, 1)
    for k, th in enumerate((0.5 * math.pi, math.pi, -0.5 * math.pi)):
        theta2 = 0.0 * 0.0 + 0.0 * 0.0 + th * th
        inv = 1.0 / math.sqrt(theta2)
        theta = theta2 * inv
        assert p.angle[k] == th
        assert p.A[k] == math.sin(theta) * inv
        assert p.B[k] == (1.0 - math.cos(theta)) * (inv * inv)
        half = 0.5 * math.sqrt(theta2)
        assert p.halfCos[k] == math.cos(half) and p.halfSin[k] == math.sin(half)
    # the rotations they make are the quarter and half turns
    assert abs(p.A[0] * (math.pi / 2) - 1.0) < 1e-15 and abs(p.B[1] * math.pi ** 2 - 2.0) < 1e-15
