"""The differentiated column collapse at its full size (workloads.column:
4096 superellipsoids of 8 templates with 32^3 SDF and 800 nodes, dropped on
a floor layer of 256 fixed bodies, dt 1e-6 s), 2000 steps from a state
packed by 20000 settling steps:

* the forward run under the adjoint gives the plain run's states;
* the adjoint with a checkpoint every 100 steps equals, bit for bit, the
  adjoint recomputed from a single checkpoint at the start, for
  J = sum |v|^2 / 2 over the moving bodies and J = runout."""
import dataclasses

import numpy as np
import pytest

from fundem_amd import objectives as O
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 2000
KEYS = ("x", "q", "v", "w", "kn", "ks", "mu", "zeta", "wall_mu", "wall_zeta")


@pytest.fixture(scope="module")
def column():
    from fundem_amd import Solver
    sc = W.column()
    s0 = Solver()
    W.apply(sc, s0)
    s0.step(20000)
    st = s0.state()
    assert s0.status() == 0
    s0.close()
    return dataclasses.replace(sc, x=st["x"], q=st["q"], v=st["v"], w=st["w"])


def test_full_column_checkpoints_equal_one_segment(column):
    from fundem_amd import Adjoint, Solver
    sc = column
    moving = np.array([int(g) not in sc.fixed_templates for g in sc.geo])
    assert moving.sum() == 4096 and (~moving).sum() == 256
    plain = Solver()
    W.apply(sc, plain)
    plain.step(STEPS)
    ref = plain.state()
    contacts = [int(plain.contacts(k)["A"].size) for k in (0, 1)]
    plain.close()
    print(f"contacts at step {STEPS}: particle {contacts[0]}, wall {contacts[1]}", flush=True)
    assert contacts[0] > 0
    runs = []
    for every in (100, STEPS):
        g = Solver()
        W.apply(sc, g)
        adj = Adjoint(g, contacts_per_particle=4)
        adj.forward(STEPS, every)
        fin = g.state()
        assert g.status() == 0
        for key in ("x", "q", "v", "w", "F", "T"):
            assert np.array_equal(fin[key], ref[key]), (every, key)
        out = []
        for J, bar in (O.kinetic_energy(fin, np.ones(len(moving)), moving),
                       O.runout(fin, fin["x"][moving].mean(axis=0), moving, tau=1e-4)):
            out.append(adj.backward(bar["x"], bar["q"], bar["v"], bar["w"]))
        runs.append(out)
        adj.close()
        g.close()
    for k, name in enumerate(("kinetic energy", "runout")):
        a, b = runs[0][k], runs[1][k]
        for key in KEYS:
            assert np.array_equal(a[key], b[key]), (name, key)
        assert np.any(a["v"] != 0.0) and a["kn"][0] != 0.0, name
        print(f"{name}: dJ/dkn {a['kn'][0]:.6e} dJ/dzeta {a['zeta'][0]:.6e} dJ/dmu {a['mu'][0]:.6e}",
              flush=True)
