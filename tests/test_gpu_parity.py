"""Parity of the HIP path against the oracle.

The device stores state in FP32 and computes in FP64 without contraction,
summing forces in a fixed canonical order. The oracle run with
store_fp32=1, canonical_fold=1 performs the same operations, so states,
contact sets (A, target, node) and tangential springs must agree bit for
bit at every step. Against the FP64 restatement of the reference
(store_fp32=0, canonical_fold=0) states agree within FP32 tolerances.
"""
import numpy as np
import pytest

import pyoracle as po
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

SETS = (po.SET_PP, po.SET_PW, po.SET_GX, po.SET_GY, po.SET_GXY)


def make_pair(scn, mirror=True, skin=None):
    from fundem_amd import Solver
    g = Solver()
    if skin is not None:
        g.set_skin(skin)
    W.apply(scn, g)
    o = po.OracleSim(mirror, mirror)
    W.apply(scn, o)
    return g, o


def assert_states_equal(gs, os_, where):
    for key in ("x", "v", "w", "q", "F", "T"):
        a, b = gs[key], os_[key]
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)[0]
            raise AssertionError(f"{where}: {key} differs first at {tuple(bad)}: "
                                 f"gpu={a[tuple(bad)]!r} oracle={b[tuple(bad)]!r}")


def assert_contacts_equal(g, o, where, sets=SETS):
    for s in sets:
        gc, oc = g.contacts(s), o.contacts(s)
        for key in ("A", "B", "k"):
            assert np.array_equal(gc[key], oc[key]), \
                f"{where}: set {s} {key} differ (gpu {gc[key].size} contacts, oracle {oc[key].size})"
        assert np.array_equal(gc["spring"], oc["spring"]), f"{where}: set {s} springs differ"


def run_bitexact(scn, steps, every, skin=None):
    g, o = make_pair(scn, skin=skin)
    assert_states_equal(g.state(), o.state(), "after initialise")
    total = 0
    for i in range(1, steps + 1):
        g.step(1)
        o.step(1)
        if i % every == 0 or i == steps:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
            total += sum(o.contacts(s)["A"].size for s in SETS)
    assert g.status() == 0
    builds, ghost_builds, used = g.list_builds()
    print(f"skin {used:.3e}: {builds} list builds, {ghost_builds} ghost-list builds in {steps} steps")
    if used == 0.0:
        assert builds == steps + 1
    return total


CLUSTERS = {
    "mm": dict(periodic=(False, False)),
    "mm-periodic-xy": dict(periodic=(True, True)),
    "mm-periodic-x": dict(periodic=(True, False)),
    "mm-periodic-y": dict(periodic=(False, True), seed=5),
    "cm-rotating": dict(periodic=(False, False), scale=1e-2, dt=1e-5, kn=1e5, spin=5.0),
    "cm-rotating-periodic": dict(periodic=(True, True), scale=1e-2, dt=1e-5, kn=1e5, spin=5.0, seed=3),
}


@pytest.mark.parametrize("name", sorted(CLUSTERS))
def test_cluster_bitexact(name):
    scn = W.cluster(n=27, speed=1.0, **CLUSTERS[name])
    contacts = run_bitexact(scn, steps=400, every=20)
    assert contacts > 0, "scenario produced no contacts"


@pytest.mark.parametrize("name,skin", [("mm-periodic-xy", 0.0), ("mm-periodic-xy", 4e-4),
                                       ("cm-rotating-periodic", 0.0), ("cm-rotating-periodic", 3e-3),
                                       ("mm", 1e-3)])
def test_cluster_bitexact_any_skin(name, skin):
    """Candidate lists rebuilt every step (skin 0) or kept across steps (a
    skin of 0.4 to 1 particle radius) give the oracle's bits."""
    scn = W.cluster(n=27, speed=1.0, **CLUSTERS[name])
    contacts = run_bitexact(scn, steps=400, every=25, skin=skin)
    assert contacts > 0, "scenario produced no contacts"


@pytest.mark.parametrize("skin", [0.0, None])
def test_moving_wall_bitexact(skin):
    """A floor with a prescribed velocity and spin (setFixedVelocityToWall,
    setFixedAngularVelocityToWall: the wall drifts, is never kicked, and its
    velocity enters the contact law) under a cluster, bit for bit."""
    import dataclasses
    scn = W.cluster(n=27, speed=1.0, scale=1e-2, dt=1e-5, kn=1e5, spin=5.0)
    x, v, w, q, m, g = scn.walls[0]
    # bring the bottom layer down to the floor
    px = scn.x.copy()
    px[:, 2] -= px[:, 2].min() - 0.011
    scn = dataclasses.replace(scn, x=px, walls=[(x, np.array([0.02, -0.01, 0.0]),
                                                 np.array([0.0, 0.0, 0.8]), q, m, g)])
    contacts = run_bitexact(scn, steps=400, every=20, skin=skin)
    assert contacts > 0


def test_list_reuse_slow_pour_bitexact():
    """A slow, settling pour (periodic, wall, many steps): lists are reused
    for many steps, particles wrap and ghosts come and go, and every checked
    step still equals the oracle."""
    scn = W.cluster(n=64, speed=0.3, spin=2.0, periodic=(True, True), seed=11, dt=2e-6)
    g, o = make_pair(scn, skin=3e-4)
    for i in range(1, 1201):
        g.step(1)
        o.step(1)
        if i % 100 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
    builds, ghost_builds, _ = g.list_builds()
    print(f"{builds} list builds, {ghost_builds} ghost-list builds in 1200 steps")
    assert builds < 600, "lists were not reused"
    assert g.status() == 0


def test_hash_table_matches_reference_sort():
    scn = W.pour((16, 16, 8), n_templates=2, n_nodes=64, n_grid=16)
    scn.periodic = (False, False)  # no wrap: stored positions are the hashed ones
    # a few particles outside the domain on every side hash to -1
    scn.x[:6] = scn.domain_min - 1e-3 + np.array(
        [[0, 0, 0], [0.0, 5e-3, 5e-3], [20e-3, 0.0, 5e-3], [5e-3, 5e-3, 0.0], [0.05, 0.05, 0.05],
         [0.06, 0.01, 0.02]])
    from fundem_amd import Solver
    g = Solver()
    W.apply(scn, g)
    x = g.state()["x"]
    grid = g.grid()
    h = po.hash_(x, grid["min"], grid["max"], grid["inv"], grid["size"])
    ng = int(np.prod(grid["size"].astype(np.int64)))
    hs, idx, st, en = po.sort_start_end(h, ng)
    t = g.hash_table(len(x))
    assert np.array_equal(t["hash"], h)
    assert np.array_equal(t["sorted_hash"], hs)
    assert np.array_equal(t["index"], idx)
    assert np.array_equal(t["start"], st)
    assert np.array_equal(t["end"], en)
    assert (h == -1).sum() >= 5


def test_single_particle_roll_bitexact_and_fp64():
    """The rolling superellipsoid: bit-exact against the FP32-storage oracle,
    and within FP32 tolerances of the FP64 restatement, over 3000 steps."""
    scn = W.single_particle_on_wall(n_nodes=2000, steps=3000)
    g, o = make_pair(scn)
    ref = po.OracleSim(False, False)
    W.apply(scn, ref)
    for i in range(1, scn.steps + 1):
        g.step(1)
        o.step(1)
        ref.step(1)
        if i % 500 == 0:
            gs = g.state()
            assert_states_equal(gs, o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}", sets=(po.SET_PW,))
            rs = ref.state()
            np.testing.assert_allclose(gs["x"], rs["x"], rtol=1e-5, atol=1e-7)
            np.testing.assert_allclose(gs["q"], rs["q"], rtol=0, atol=1e-5)
            # velocities within 1e-4 of the velocity scale
            assert np.abs(gs["v"] - rs["v"]).max() <= 1e-4 * np.abs(rs["v"]).max()
    assert o.contacts(po.SET_PW)["A"].size > 0


def test_deterministic_and_graph_replay():
    """Two runs of the same input give identical bits, and replaying the
    captured two-step graph gives the same bits as launching each stage."""
    scn = W.cluster(n=27, periodic=(True, True), speed=1.0)
    from fundem_amd import Solver
    runs = []
    for graph in (False, False, True):
        g = Solver()
        W.apply(scn, g)
        g.use_graph(graph)
        g.step(1)      # first step runs eagerly (hook detection)
        g.step(200)
        runs.append(g.state())
        assert g.status() == 0
    for r in runs[1:]:
        assert_states_equal(r, runs[0], "repeat")


def test_bonded_beam_bitexact_and_break_events():
    scn = W.bonded_beam(counts=(8, 2, 2), tip_speed=5.0, strength_mean=2e5, strength_sd=5e4,
                        n_nodes=128, dt=1e-6, steps=600)
    g, o = make_pair(scn)
    nb = len(scn.bonds)
    for i in range(1, scn.steps + 1):
        g.step(1)
        o.step(1)
        if i % 50 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            ga = g.bonds(nb)["active"]
            oa = o.bonds()[0]
            assert np.array_equal(ga, oa), f"step {i}: bond flags differ"
    assert o.bonds()[0].sum() < nb, "no bond broke"


def wrap_contact_scenario(periodic=(True, True)):
    """Two centimetre spheres in contact moving +x together; the leading one
    crosses the periodic face x = max within a few steps while the contact
    lasts (the wrap step: detection before the wrap, forces after it)."""
    from fundem_amd import generators as gen
    r = 1e-2
    t = gen.sphere_template(r, n_grid=16, n_nodes=400)
    L = 8 * r
    x = np.array([[L - 0.02 * r, 4 * r, 4 * r], [L - 1.96 * r, 4.0 * r, 4.0 * r]])
    v = np.array([[5.0, 0.0, 0.0], [5.0, 0.3, 0.0]])
    w = np.array([[0.0, 0.0, 3.0], [0.0, 2.0, 0.0]])
    q = np.array([[1.0, 0.0, 0.0, 0.0]] * 2)
    return W.Scenario(
        name="wrap-contact", templates=[t], density=2500.0,
        materials=[(1e4, 4e3, 0.4, 0.5)], x=x, v=v, w=w, q=q,
        mat=np.zeros(2, np.int32), geo=np.zeros(2, np.int32),
        domain_min=np.zeros(3), domain_max=np.array([L, L, L]), gravity=np.array([0.0, 0.0, -9.81]),
        dt=1e-5, periodic=tuple(periodic), steps=30)


@pytest.mark.parametrize("skin", [0.0, None])
def test_contact_across_periodic_wrap_bitexact(skin):
    """A particle that wraps while in contact: the contact of the wrap step is
    detected at the pre-wrap position and its force uses the post-wrap arm, as
    in the reference (LSSolver.h:1166-1178); states, contact sets and springs
    equal the oracle's every step."""
    scn = wrap_contact_scenario()
    g, o = make_pair(scn, skin=skin)
    seen = 0
    for i in range(1, 21):
        g.step(1)
        o.step(1)
        assert_states_equal(g.state(), o.state(), f"step {i}")
        assert_contacts_equal(g, o, f"step {i}")
        seen += o.contacts(po.SET_PP)["A"].size
    assert seen > 0


def sector_scenario():
    """A centimetre cluster in the sector periodic boundary (the quarter plane
    about the domain's min corner) whose first column is driven through the
    x = min face: particles turn by -pi/2 into the y = min face, and R90, R180
    and R270 images all carry contacts."""
    import dataclasses
    scn = W.cluster(n=27, speed=1.0, scale=1e-2, dt=1e-5, kn=1e5, spin=5.0)
    v = scn.v.copy()
    v[scn.x[:, 0] < scn.domain_min[0] + 0.02, 0] -= 5.0
    return dataclasses.replace(scn, sector=True, v=v, steps=600)


def test_sector_periodic_bitexact():
    """PeriodicBoundarySector (PeriodicBoundary.h:319-426): wraps that turn
    position, velocity and orientation, and the three image sets with turned
    poses and velocities; states, contact sets and springs equal the
    oracle's every step."""
    scn = sector_scenario()
    g, o = make_pair(scn)
    assert_states_equal(g.state(), o.state(), "after initialise")
    seen = np.zeros(5, int)
    wraps = 0
    for i in range(1, scn.steps + 1):
        xb = o.state()["x"]
        g.step(1)
        o.step(1)
        os_ = o.state()
        assert_states_equal(g.state(), os_, f"step {i}")
        wraps += int(np.abs(os_["x"] - xb).max() > 1e-3)
        if i % 10 == 0:
            assert_contacts_equal(g, o, f"step {i}")
            seen += [o.contacts(s)["A"].size > 0 for s in SETS]
    assert g.status() == 0
    assert wraps >= 3 and seen[0] > 0 and seen[2] > 0 and seen[3] > 0 and seen[4] > 0, (wraps, seen)


def test_sector_graph_replay_and_guards():
    """The sector step replayed from captured graphs gives the eager bits;
    the adjoint refuses the sector."""
    from fundem_amd import Adjoint, Solver
    scn = sector_scenario()
    runs = []
    for graph in (False, True):
        g = Solver()
        W.apply(scn, g)
        g.use_graph(graph)
        g.step(1)
        g.step(499)
        runs.append(g.state())
        assert g.status() == 0
    assert_states_equal(runs[1], runs[0], "graph replay")
    g = Solver()
    W.apply(scn, g)
    with pytest.raises(RuntimeError, match="not differentiated"):
        Adjoint(g).forward(5, 5)


def test_external_force_hooks_bitexact():
    """The tutorials' addExternalForceTorque overrides (global damping and the
    constant-force kernel, apps/tutorial/globalDamping.cu) run between the
    force reduction and the second half-kick; states equal the oracle's."""
    scn = W.cluster(n=27, periodic=(True, True), speed=1.0, scale=1e-2, dt=1e-5, kn=1e5, spin=5.0)
    from fundem_amd import Solver
    g = Solver()
    g.set_global_damping(0.2)
    g.set_constant_force(3, [1e-3, 0.0, 2e-3], [0.0, -1e-3, 0.0])
    W.apply(scn, g)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    o.set_global_damping(0.2)
    o.set_constant_force(3, [1e-3, 0.0, 2e-3], [0.0, -1e-3, 0.0])
    for i in range(1, 301):
        g.step(1)
        o.step(1)
        if i % 10 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
    g.set_global_damping(None)
    g.set_constant_force(None, [0, 0, 0], [0, 0, 0])
    o.set_global_damping(None)
    o.set_constant_force(None, [0, 0, 0], [0, 0, 0])
    g.step(20)
    o.step(20)
    assert_states_equal(g.state(), o.state(), "hooks off")
