"""Parity on the configurations' own shapes and sizes.

* The pour's own workload (16 superellipsoid templates, 32^3 FP32 SDF, 1000
  nodes, kn 1e4, mu 0.5, periodic x/y with a floor), on a 6^3 slice whose
  layers were thrown together on the device: 200 steps of their collisions
  bit for bit against the mirror oracle, with particle-particle, wall and
  ghost contacts present.
* Pairs holding more than 64 node contacts (a boxy superellipsoid pressed
  flat onto the floor and a second one pressed onto its top face), which
  the narrow phase handles past its LDS staging: bit for bit.
* The single rolling particle to 1e4 steps, FP64-storage build against the
  FP64 restatement (the reference's summation order): positions,
  orientations and contact forces within 1e-5 relative.
* One step from identical states of a pour slice and a packing of the
  batch (FP32-storage build) and a bonded FCC beam (FP64-storage build)
  against the FP64 restatement: forces and torques within 1e-5 relative.
* The bonded cantilever at its full size (197,676 FCC spheres, 16^3 SDF,
  400 nodes, 1,137,642 bonds): kept candidate lists equal rebuilding every
  step bit for bit, bond failures included; a 2016-sphere FCC beam's
  failures are bit-exact against the oracle.
* The 512-packing batch at its full size: sampled packings evolve bit for
  bit as in solvers of their own.
"""
import dataclasses

import numpy as np
import pytest

import pyoracle as po
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

SETS = (po.SET_PP, po.SET_PW, po.SET_GX, po.SET_GY, po.SET_GXY)
KEYS = ("x", "v", "w", "q", "F", "T")


def solver(skin=None, precision="fp32"):
    from fundem_amd import Solver
    g = Solver(precision=precision)
    if skin is not None:
        g.set_skin(skin)
    return g


def assert_states_equal(a, b, where):
    for key in KEYS:
        if not np.array_equal(a[key], b[key]):
            bad = np.argwhere(a[key] != b[key])[0]
            raise AssertionError(f"{where}: {key} differs first at {tuple(bad)}: "
                                 f"{a[key][tuple(bad)]!r} vs {b[key][tuple(bad)]!r}")


def assert_contacts_equal(g, o, where):
    for s in SETS:
        gc, oc = g.contacts(s), o.contacts(s)
        for key in ("A", "B", "k", "spring"):
            assert np.array_equal(gc[key], oc[key]), \
                f"{where}: set {s} {key} differ ({gc['A'].size} vs {oc['A'].size} contacts)"


def rel_err(a, b):
    """Largest per-row error |a_i - b_i| over the largest row norm of b."""
    scale = np.linalg.norm(b, axis=1).max()
    return float(np.linalg.norm(a - b, axis=1).max() / scale) if scale > 0 else 0.0


# ------------------------------------------------------------------ pour
def compress(scn, steps, closing_time=1.5e-3):
    """The scenario driven into a contact-dense state on the device: each
    particle starts falling at a speed proportional to its height above the
    lowest, so that the layers close in on one another within about
    `closing_time` and collide at about (lattice spacing) / closing_time,
    with seeded random horizontal velocities of about half a lattice spacing
    per closing time (particles meet sideways and across the periodic faces
    too); after `steps` steps of the scenario's dt the state (x, v, w,
    q) is taken as a new initial state (springs start from zero)."""
    v = scn.v.copy()
    v[:, 2] = -(scn.x[:, 2] - scn.x[:, 2].min()) / closing_time
    # lateral motion as well, so that particles also meet across the periodic faces
    spacing = 2.1 * max(float(np.linalg.norm(t.nodes, axis=1).max()) for t in scn.templates)
    v[:, :2] = np.random.default_rng(7).normal(0.0, 0.5 * spacing / closing_time, size=(scn.n, 2))
    g = solver()
    W.apply(dataclasses.replace(scn, v=v), g)
    g.step(steps)
    assert g.status() == 0
    st = g.state()
    g.close()
    return dataclasses.replace(scn, x=st["x"], v=st["v"], w=st["w"], q=st["q"])


_SETTLED = {}


def settled_pour_slice(side=6, steps=1100):
    """A side^3 slice of the pour (its templates, materials, periodic box and
    floor) in the middle of its layers' collisions (`compress`)."""
    key = (side, steps)
    if key not in _SETTLED:
        _SETTLED[key] = compress(W.pour((side, side, side)), steps)
    return _SETTLED[key]


def test_pour_slice_contact_dense_bitexact():
    scn = settled_pour_slice()
    print(f"settled pour slice: {scn.n} particles, 16 templates, 1000 nodes", flush=True)
    g = solver()
    W.apply(scn, g)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    assert_states_equal(g.state(), o.state(), "after initialise")
    seen = np.zeros(5, int)
    for i in range(1, 201):
        g.step(1)
        o.step(1)
        if i % 25 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
            seen = np.maximum(seen, [o.contacts(s)["A"].size for s in SETS])
            print(f"step {i}: equal", flush=True)
    print(f"most node contacts per set (pp, wall, ghost x, y, xy): {seen.tolist()}", flush=True)
    assert g.status() == 0
    assert seen[0] >= 50 and seen[1] > 0 and seen[2] + seen[3] + seen[4] > 0, seen


# ------------------------------------------------- pairs past LDS staging
def pressed_boxes():
    """Two boxy superellipsoids (semi-axes 10 mm, powers 8, 1000 nodes) with
    identity orientation: A pressed 0.3 mm into the floor, B pressed 0.3 mm
    onto A's top face and sliding; kn 1e3 keeps them pressed for the run."""
    from fundem_amd import generators as gen
    t = gen.superellipsoid_template(0.01, 0.01, 0.01, 8.0, 8.0, n_grid=32, n_nodes=1000)
    low = float(t.nodes[:, 2].astype(np.float64).min())
    high = float(t.nodes[:, 2].astype(np.float64).max())
    d = 3e-4
    xa = np.array([0.05, 0.05, -low - d])
    xb = xa + np.array([0.001, -0.0005, high - low - d])
    wt, wall = W.floor_wall([0.0, 0.0], [0.1, 0.1])
    return W.Scenario(
        name="pressed-boxes", templates=[t], density=2500.0, materials=[(1e3, 4e2, 0.4, 0.5)],
        x=np.array([xa, xb]), v=np.array([[0.0, 0.0, 0.0], [0.05, 0.02, 0.0]]),
        w=np.zeros((2, 3)), q=np.tile(W.IDENT, (2, 1)), mat=np.zeros(2, np.int32),
        geo=np.zeros(2, np.int32), domain_min=np.array([0.0, 0.0, -0.02]),
        domain_max=np.array([0.1, 0.1, 0.08]), gravity=np.array([0.0, 0.0, -9.81]), dt=1e-6,
        wall_templates=[wt], wall_materials=[(0.3, 0.5)], walls=[wall])


@pytest.mark.parametrize("skin", [0.0, None])
def test_pairs_past_lds_staging_bitexact(skin):
    scn = pressed_boxes()
    g = solver(skin)
    W.apply(scn, g)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    most = [0, 0]
    for i in range(1, 201):
        g.step(1)
        o.step(1)
        if i % 20 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
            for k, s in enumerate((po.SET_PP, po.SET_PW)):
                c = o.contacts(s)
                if c["A"].size:
                    most[k] = max(most[k], int(np.unique(c["A"] * 8 + c["B"],
                                                         return_counts=True)[1].max()))
    print(f"node contacts of the fullest pair: particle {most[0]}, wall {most[1]}", flush=True)
    assert g.status() == 0
    assert most[0] > 64 and most[1] > 64, most


# ------------------------------------------------ single particle, 1e4 steps
def test_single_particle_roll_1e4_steps_against_fp64():
    """The FP64-storage build against the FP64 restatement in the
    reference's summation order. (The FP32-storage build follows it to FP32
    tolerances over 3000 steps, test_gpu_parity.py; over the bounces that
    come later, the contact's stiffness amplifies FP32 rounding of the state
    beyond 1e-5.)"""
    scn = W.single_particle_on_wall(n_nodes=2000, steps=10000)
    g = solver(precision="fp64")
    W.apply(scn, g)
    ref = po.OracleSim(False, False)
    W.apply(scn, ref)
    worst = dict(x=0.0, q=0.0, F=0.0)
    touching = 0
    for i in range(1, scn.steps + 1, 1000):
        g.step(1000)
        ref.step(1000)
        gs, rs = g.state(), ref.state()
        touching += ref.contacts(po.SET_PW)["A"].size
        ex = float(np.abs(gs["x"] - rs["x"]).max() / np.abs(rs["x"]).max())
        eq = float(np.abs(gs["q"] - rs["q"]).max())
        ef = rel_err(gs["F"], rs["F"])
        worst = dict(x=max(worst["x"], ex), q=max(worst["q"], eq), F=max(worst["F"], ef))
        print(f"step {i + 999}: x {ex:.2e} q {eq:.2e} F {ef:.2e} (|F| {np.linalg.norm(rs['F']):.3e} N)",
              flush=True)
    assert touching > 0
    assert worst["x"] <= 1e-5 and worst["q"] <= 1e-5 and worst["F"] <= 1e-5, worst


# --------------------------------------------- one step from identical states
def one_step_errors(scn, steps=(1,), precision="fp32"):
    g = solver(precision=precision)
    W.apply(scn, g)
    ref = po.OracleSim(False, False)
    W.apply(scn, ref)
    out = []
    done = 0
    for s in steps:
        g.step(s - done)
        ref.step(s - done)
        done = s
        gs, rs = g.state(), ref.state()
        out.append((rel_err(gs["F"], rs["F"]), rel_err(gs["T"], rs["T"]),
                    float(np.linalg.norm(rs["F"], axis=1).max())))
    g.close()
    return out


def test_one_step_forces_pour_slice_against_fp64():
    (ef, et, fmax), = one_step_errors(settled_pour_slice())
    print(f"pour slice: F {ef:.2e} T {et:.2e} (largest |F| {fmax:.3e} N)", flush=True)
    assert fmax > 0 and ef <= 1e-5 and et <= 1e-5


def test_forces_bonded_fcc_beam_against_fp64():
    """The FP64-storage build (bond constants and frames in FP64): at the
    beam's small strains FP32 bond geometry alone moves the bond forces by
    about 1e-4 relative (an FP32 ulp of the 1 mm rest length against
    extensions of 1e-7 m)."""
    scn = W.bonded_beam(counts=(14, 6, 6), lattice="fcc", strength_mean=1e4, strength_sd=2e3,
                        tip_speed=5.0)
    for s, (ef, et, fmax) in zip((1, 10, 50), one_step_errors(scn, (1, 10, 50), "fp64")):
        print(f"FCC beam step {s}: F {ef:.2e} T {et:.2e} (largest |F| {fmax:.3e} N)", flush=True)
        assert fmax > 0 and ef <= 1e-5 and et <= 1e-5, s


def test_one_step_forces_batch_packing_against_fp64():
    """The largest packing among eight of the batch's seeded draws, in the
    middle of its layers' collisions (`compress`)."""
    scns = W.batch_packings(count=8)
    k = int(np.argmax([s.n for s in scns]))
    scn = compress(scns[k], 1100)
    (ef, et, fmax), = one_step_errors(scn)
    print(f"packing {k} ({scn.n} particles): F {ef:.2e} T {et:.2e} (largest |F| {fmax:.3e} N)",
          flush=True)
    assert fmax > 0 and ef <= 1e-5 and et <= 1e-5


# ------------------------------------------------------------------ beams
def test_fcc_beam_break_events_bitexact():
    scn = W.bonded_beam(counts=(14, 6, 6), lattice="fcc", strength_mean=1e4, strength_sd=2e3,
                        tip_speed=5.0, steps=100)
    print(f"FCC beam: {scn.n} spheres, {len(scn.bonds)} bonds", flush=True)
    g = solver()
    W.apply(scn, g)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    nb = len(scn.bonds)
    for i in range(1, scn.steps + 1):
        g.step(1)
        o.step(1)
        if i % 25 == 0:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            gb, ob = g.bonds(nb), o.bonds()
            assert np.array_equal(gb["active"], ob[0]), f"step {i}: bond flags differ"
            print(f"step {i}: equal, {int(nb - ob[0].sum())} bonds broken", flush=True)
    broken = int(nb - o.bonds()[0].sum())
    print(f"{broken} of {nb} bonds broke", flush=True)
    assert broken > 0


def test_full_size_fcc_beam_kept_lists_equal_rebuilds():
    scn = W.bonded_beam(counts=(171, 17, 17), lattice="fcc", strength_mean=1e4, strength_sd=2e3,
                        tip_speed=1.0)
    nb = len(scn.bonds)
    print(f"full beam: {scn.n} spheres, {nb} bonds", flush=True)
    runs = []
    for skin in (None, 0.0):
        g = solver(skin)
        W.apply(scn, g)
        g.step(1)
        g.step(399)
        assert g.status() == 0
        runs.append((g.state(), g.bonds(nb), g.list_builds()))
        g.close()
    (a, ba, (builds, _, s)), (b, bb, (builds0, _, _)) = runs
    print(f"skin {s:.3g} m: {builds} list builds in 400 steps; skin 0: {builds0}; "
          f"{int(nb - ba['active'].sum())} bonds broke", flush=True)
    assert builds < 100 and builds0 >= 400
    assert_states_equal(a, b, "kept lists vs rebuilds")
    assert np.array_equal(ba["active"], bb["active"])
    assert np.array_equal(ba["break_step"], bb["break_step"])
    assert ba["active"].sum() < nb


# ------------------------------------------------------------------ batch
def test_full_batch_sampled_packings_equal_plain_runs():
    scns = W.batch_packings()
    total = sum(s.n for s in scns)
    print(f"batch of {len(scns)} packings, {total} particles", flush=True)
    g = solver()
    g.set_capacity(8, 24)
    W.apply_batch(scns, g)
    g.step(1)
    g.step(99)
    assert g.status() == 0
    gs = g.state()
    g.close()
    off = np.cumsum([0] + [s.n for s in scns])
    sizes = np.array([s.n for s in scns])
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 300})
    for k in pick:
        p = solver()
        p.set_capacity(8, 24)
        W.apply(scns[k], p)
        p.step(1)
        p.step(99)
        ps = p.state()
        p.close()
        part = {key: gs[key][off[k]:off[k + 1]] for key in KEYS}
        assert_states_equal(part, ps, f"packing {k} ({scns[k].n} particles)")
