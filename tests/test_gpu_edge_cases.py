"""Edge cases of the step against the oracle, bit for bit:
  * a system with walls and no particles, and one with particles and no
    walls, step without error;
  * particles that start or end up outside the domain (Q8: their cell id is
    -1, so they are never a neighbour, while their own nodes still search
    clamped cells) and a particle exactly on a cell face;
  * a particle that leaves the domain through the top while periodic in x, y."""
import dataclasses

import numpy as np
import pytest

import pyoracle as po
from fundem_amd import workloads as W
from test_gpu_parity import SETS, assert_contacts_equal, assert_states_equal, make_pair

pytestmark = pytest.mark.gpu


def run_pair(scn, steps, every=10, skin=None):
    g, o = make_pair(scn, skin=skin)
    assert_states_equal(g.state(), o.state(), "after initialise")
    for i in range(1, steps + 1):
        g.step(1)
        o.step(1)
        if i % every == 0 or i == steps:
            assert_states_equal(g.state(), o.state(), f"step {i}")
            assert_contacts_equal(g, o, f"step {i}")
    assert g.status() == 0
    return g, o


def test_walls_without_particles_and_particles_without_walls():
    from fundem_amd import Solver
    scn = W.cluster(n=8, speed=1.0)
    empty = dataclasses.replace(scn, x=scn.x[:0], v=scn.v[:0], w=scn.w[:0], q=scn.q[:0],
                                mat=scn.mat[:0], geo=scn.geo[:0])
    g = Solver()
    W.apply(empty, g)
    g.step(5)
    assert g.status() == 0 and g.state()["x"].shape == (0, 3)
    nowalls = dataclasses.replace(scn, wall_templates=[], wall_materials=[], walls=[])
    run_pair(nowalls, 60)


@pytest.mark.parametrize("skin", [0.0, None])
def test_particles_outside_the_domain(skin):
    scn = W.cluster(n=27, speed=1.0)
    x = scn.x.copy()
    x[0, 2] = scn.domain_max[2] + 0.4e-3                 # above the top: cell id -1
    x[1, 0] = scn.domain_max[0] + 0.2e-3                 # beyond x = max (not periodic here)
    cs = (scn.domain_max[2] - scn.domain_min[2]) / 4
    x[2, 2] = scn.domain_min[2] + cs                      # on a cell face
    v = scn.v.copy()
    v[3] = [0.0, 0.0, 3.0]                               # leaves through the top
    run_pair(dataclasses.replace(scn, x=x, v=v), 200, skin=skin)


def test_leaving_through_the_top_with_periodic_sides():
    scn = W.cluster(n=27, periodic=(True, True), speed=1.0)
    v = scn.v.copy()
    v[:4, 2] += 6.0
    run_pair(dataclasses.replace(scn, v=v), 300)
