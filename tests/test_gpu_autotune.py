"""The autotuner changes speed, not results: a run after autotune() (which
times candidate skins, cell sizes and launch shapes from a saved point and
returns the solver to it) equals a plain run bit for bit, for a
non-periodic system and a periodic one; cells coarser than the grid rule's
keep periodic contact sets and states equal to the oracle's; every launch
shape gives the same bits."""
import numpy as np
import pytest

import pyoracle as po
from fundem_amd import workloads as W

SETS = (po.SET_PP, po.SET_PW, po.SET_GX, po.SET_GY, po.SET_GXY)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [(False, False), (True, True)])
def test_autotune_keeps_results(periodic):
    from fundem_amd import Solver
    scn = W.cluster(n=27, periodic=periodic, speed=1.0)
    plain = Solver()
    W.apply(scn, plain)
    plain.step(150)
    ref = plain.state()
    tuned = Solver()
    W.apply(scn, tuned)
    choice = tuned.autotune(trial_steps=6)
    for line in choice["trials"]:
        print(line)
    print({k: v for k, v in choice.items() if k != "trials"})
    assert len(choice["trials"]) >= 16
    assert sum(line.startswith("cell") for line in choice["trials"]) >= 3, "cell size not tuned"
    assert choice["law_block"] in (64, 128, 256) and choice["sweep_waves"] in (1, 2, 4)
    tuned.step(150)
    got = tuned.state()
    assert tuned.status() == 0
    for key in ("x", "v", "w", "q", "F", "T"):
        assert np.array_equal(got[key], ref[key]), key


SHAPES = [dict(sweep_waves=w, nodes_per_lane=npl, law_mode=1) for w in (1, 4) for npl in (1, 2, 3)] + \
         [dict(sweep_waves=w, law_mode=2) for w in (1, 2)]


def _shapes_case(name):
    import dataclasses
    if name == "pressed":
        from test_gpu_configs import pressed_boxes
        return pressed_boxes(), 150
    scn = W.pour((5, 5, 5))
    v = scn.v.copy()
    v[:, 2] = -(scn.x[:, 2] - scn.x[:, 2].min()) / 1.5e-3
    spacing = 2.1 * max(float(np.linalg.norm(t.nodes, axis=1).max()) for t in scn.templates)
    v[:, :2] = np.random.default_rng(7).normal(0.0, 0.5 * spacing / 1.5e-3, size=(scn.n, 2))
    return dataclasses.replace(scn, v=v), 1200


@pytest.mark.parametrize("name", ["pour", "pressed"])
def test_launch_shapes_keep_results(name):
    """Every launch shape of one solver (sweep waves per workgroup, how a
    sweep wave walks a pair's chunks, where the contact law runs) gives the
    same bits: a 5^3 slice of the pour whose layers collide (particle, wall
    and ghost contacts) and a pair held at more than 64 node contacts."""
    from fundem_amd import Solver
    scn, steps = _shapes_case(name)
    runs = []
    for shape in SHAPES:
        g = Solver()
        g.set_launch_tuning(**shape)
        W.apply(scn, g)
        g.step(steps)
        assert g.status() == 0
        assert g.launch_tuning()["nodes_per_lane"] == shape.get("nodes_per_lane", 1)
        runs.append((shape, g.state(), [g.contacts(k)["A"].size for k in range(5)]))
        g.close()
    print(f"{name}: contacts per set {runs[0][2]}", flush=True)
    assert runs[0][2][0] > 0
    for shape, st, _ in runs[1:]:
        for key in ("x", "v", "w", "q", "F", "T"):
            assert np.array_equal(st[key], runs[0][1][key]), (shape, key)


@pytest.mark.parametrize("periodic,factor", [((True, True), 1.25), ((True, True), 2.0),
                                             ((True, False), 1.5), ((False, True), 1.5)])
def test_coarser_cells_keep_periodic_results(periodic, factor):
    """A periodic cluster gridded with cells `factor` times the rule's (2 x
    the largest radius), which the autotuner may pick, against the oracle
    gridded by the rule: states, contact sets and springs bit for bit."""
    from fundem_amd import Solver
    scn = W.cluster(n=125, periodic=periodic, speed=1.0)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    rule = 2.0 * max(o.geometry(k)["radius"] for k in range(len(scn.templates)))
    g = Solver()
    g.set_cell_size(factor * rule)
    W.apply(scn, g)
    assert g.grid()["size"][0] < o.grid()["size"][0], (g.grid()["size"], o.grid()["size"])
    seen = np.zeros(len(SETS), int)
    for i in range(1, 301):
        g.step(1)
        o.step(1)
        if i % 50 == 0:
            gs, os_ = g.state(), o.state()
            for key in ("x", "v", "w", "q", "F", "T"):
                assert np.array_equal(gs[key], os_[key]), (i, key)
            for j, s in enumerate(SETS):
                gc, oc = g.contacts(s), o.contacts(s)
                for key in ("A", "B", "k", "spring"):
                    assert np.array_equal(gc[key], oc[key]), (i, s, key)
                seen[j] += oc["A"].size
    print(f"{periodic} x{factor}: contacts seen per set {seen.tolist()}", flush=True)
    assert g.status() == 0 and seen[0] > 0 and seen[2:].sum() > 0


def test_periodic_cell_below_twice_the_radius_is_kept():
    """A periodic system configured with cells below twice the largest radius
    (as the pour's 2.5 mm cell) leaves some bodies near a face without an
    image; a coarser cell would add their contacts, so the autotuner keeps
    the configured cell and the tuned run equals a plain one."""
    import dataclasses
    from fundem_amd import Solver
    scn = W.cluster(n=125, periodic=(True, True), speed=1.0)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    R = max(o.geometry(k)["radius"] for k in range(len(scn.templates)))
    scn = dataclasses.replace(scn, cell_size=1.3 * R)
    plain = Solver()
    W.apply(scn, plain)
    plain.step(100)
    ref = plain.state()
    tuned = Solver()
    W.apply(scn, tuned)
    choice = tuned.autotune(trial_steps=6)
    assert not any(line.startswith("cell") for line in choice["trials"])
    assert choice["cell_size"] == 1.3 * R
    tuned.step(100)
    got = tuned.state()
    for key in ("x", "v", "w", "q", "F", "T"):
        assert np.array_equal(got[key], ref[key]), key
