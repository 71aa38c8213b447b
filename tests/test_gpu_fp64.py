"""The FP64-storage build (libfundem_amd_fp64.so: velocities, spins,
orientations, forces, torques, springs and bond constants stored in FP64)
against the oracle with FP64 storage.

With canonical sums (store_fp32=0, canonical_fold=1) the oracle performs the
same FP64 operations in the same order, so states, contact sets and springs
must agree bit for bit; with the reference's summation order
(canonical_fold=0) the two differ only in how the forces on a particle are
added up, within a 1e-5 relative bound."""
import numpy as np
import pytest

import pyoracle as po
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

SETS = (po.SET_PP, po.SET_PW, po.SET_GX, po.SET_GY, po.SET_GXY)


def fp64_pair(scn, canonical=True, skin=None):
    from fundem_amd import Solver
    g = Solver(precision="fp64")
    if skin is not None:
        g.set_skin(skin)
    W.apply(scn, g)
    o = po.OracleSim(False, canonical)
    W.apply(scn, o)
    return g, o


def assert_states_equal(a, b, where):
    for key in ("x", "v", "w", "q", "F", "T"):
        if not np.array_equal(a[key], b[key]):
            bad = np.argwhere(a[key] != b[key])[0]
            raise AssertionError(f"{where}: {key} differs first at {tuple(bad)}: "
                                 f"{a[key][tuple(bad)]!r} vs {b[key][tuple(bad)]!r}")


def assert_contacts_equal(g, o, where):
    for s in SETS:
        gc, oc = g.contacts(s), o.contacts(s)
        for key in ("A", "B", "k", "spring"):
            assert np.array_equal(gc[key], oc[key]), f"{where}: set {s} {key} differ"


CLUSTERS = {
    "mm-periodic-xy": dict(periodic=(True, True)),
    "cm-rotating": dict(periodic=(False, False), scale=1e-2, dt=1e-5, kn=1e5, spin=5.0),
    "cm-rotating-periodic": dict(periodic=(True, True), scale=1e-2, dt=1e-5, kn=1e5, spin=5.0,
                                 seed=3),
}


@pytest.mark.parametrize("name", sorted(CLUSTERS))
def test_fp64_build_bitexact_with_fp64_oracle(name):
    scn = W.cluster(n=27, speed=1.0, **CLUSTERS[name])
    g, o = fp64_pair(scn)
    assert g.precision == "fp64"
    assert_states_equal(g.state(), o.state(), "after initialise")
    total = 0
    for i in range(1, 401):
        g.step(1)
        o.step(1)
        if i % 25 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
            total += sum(o.contacts(s)["A"].size for s in SETS)
    assert total > 0 and g.status() == 0


def test_fp64_build_bonds_bitexact_with_fp64_oracle():
    scn = W.bonded_beam(counts=(8, 2, 2), tip_speed=5.0, strength_mean=2e5, strength_sd=5e4,
                        n_nodes=128, dt=1e-6, steps=600)
    g, o = fp64_pair(scn)
    nb = len(scn.bonds)
    for i in range(1, scn.steps + 1):
        g.step(1)
        o.step(1)
        if i % 50 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert np.array_equal(g.bonds(nb)["active"], o.bonds()[0]), f"step {i}"
    assert o.bonds()[0].sum() < nb, "no bond broke"


def test_fp64_build_cluster_in_reference_order_within_tolerance():
    """A rotating cluster with particle, wall and ghost contacts for 2000
    steps against the oracle summing in the reference's order: positions,
    orientations and forces within 1e-5 relative."""
    scn = W.cluster(n=27, speed=1.0, **CLUSTERS["cm-rotating-periodic"])
    g, o = fp64_pair(scn, canonical=False)
    worst = 0.0
    for i in range(1, 2001):
        g.step(1)
        o.step(1)
        if i % 250 == 0:
            gs, os_ = g.state(), o.state()
            ex = np.abs(gs["x"] - os_["x"]).max() / np.abs(os_["x"]).max()
            eq = np.abs(gs["q"] - os_["q"]).max()
            fs = np.linalg.norm(os_["F"], axis=1).max()
            ef = np.linalg.norm(gs["F"] - os_["F"], axis=1).max() / fs if fs > 0 else 0.0
            worst = max(worst, ex, eq, ef)
            print(f"step {i}: x {ex:.2e} q {eq:.2e} F {ef:.2e}", flush=True)
    assert worst <= 1e-5, worst
