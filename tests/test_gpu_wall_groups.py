"""Wall pairs four per wave (kWallGroups, sixteen lanes per pair) against the
oracle, on systems too small to select it by default: forced with the launch
tuning's wallGroups = 1, states, contact sets (A, target, node) and springs
equal the mirror oracle's bit for bit.

* slices of the pour resting on their floor: 6 x 6 x 4 bodies of two
  64-node templates, and 4 x 4 x 2 bodies of the pour's own templates (1000
  nodes, sixteen chunks each);
* two boxy superellipsoids pressed onto the floor: the wall pair holds more
  than 64 node contacts, past the 16 a group stages per pair, so its
  contacts are recomputed into the workspace;
* the first slice with its floor twice (two walls at the same place): each
  body on the floor has a pair with both, so groups of four pairs have
  different targets and are swept one pair at a time."""
import dataclasses

import pytest

import pyoracle as po
from fundem_amd import workloads as W
from test_gpu_configs import pressed_boxes
from test_gpu_parity import assert_contacts_equal, assert_states_equal

pytestmark = pytest.mark.gpu

SLICES = {
    "small-templates": lambda: W.pour((6, 6, 4), n_templates=2, n_nodes=64, n_grid=16),
    "pour-templates": lambda: W.pour((4, 4, 2)),
}


def run_grouped(scn, steps, every):
    from fundem_amd import Solver
    g = Solver()
    g.set_launch_tuning(wall_groups=1)
    W.apply(scn, g)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    walls = 0
    for i in range(1, steps + 1):
        g.step(1)
        o.step(1)
        walls += o.contacts(po.SET_PW)["A"].size
        if i % every == 0 or i == steps:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
    assert g.status() == 0 and g.launch_tuning()["wall_groups"] == 1
    g.close()
    return walls


@pytest.mark.parametrize("name", sorted(SLICES))
def test_grouped_wall_pairs_bitexact(name):
    walls = run_grouped(SLICES[name](), 100, 10)
    print(f"{name}: {walls} wall node contacts summed over the steps", flush=True)
    assert walls > 0


def test_grouped_wall_pairs_past_the_staged_contacts_bitexact():
    assert run_grouped(pressed_boxes(), 200, 20) > 64


def test_grouped_wall_pairs_with_different_targets_bitexact():
    scn = SLICES["small-templates"]()
    scn = dataclasses.replace(scn, walls=list(scn.walls) + [scn.walls[0]])
    walls = run_grouped(scn, 100, 10)
    print(f"two floors: {walls} wall node contacts summed over the steps", flush=True)
    assert walls > 0
