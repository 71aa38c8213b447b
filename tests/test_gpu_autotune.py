"""The autotuner changes speed, not results: a run after autotune() (which
times candidate skins, cell sizes and launch shapes from a saved point and
returns the solver to it) equals a plain run bit for bit, for a
non-periodic system (where the cell size is tuned too) and a periodic one."""
import numpy as np
import pytest

from fundem_amd import workloads as W

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
