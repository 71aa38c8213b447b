"""The oracle's external-force hooks, the tutorials' addExternalForceTorque
overrides (apps/tutorial/globalDamping.cu), pinned by what the dynamics fix:
  * a constant force F on a free particle, applied after every step's forces,
    gives v = (n - 1/2) dt F / m after n kick-drift-kick steps (the first
    half-kick of step 1 sees no force); the tutorial kernel adds its torque
    argument to the force as well, so F + T is what acts;
  * global damping c on a block sliding on the floor (normal force m g,
    friction at its cap mu m g, so |F| = m g sqrt(1 + mu^2) along a velocity
    in +x) adds c |F| / m to the deceleration: mu g + c g sqrt(1 + mu^2)."""
import math

import numpy as np

import pyoracle as po
from fundem_amd import generators as gen
from test_oracle import _sphere_probe


def test_constant_force_on_a_free_particle():
    r, rho = 0.01, 2000.0
    m = 4 / 3 * math.pi * r ** 3 * rho
    s = po.OracleSim(False, False)
    s.add_material(1e5, 0, 0, 0.5)
    s.add_geometry(m, [0.4 * m * r * r] * 3 + [0, 0, 0], gen.sphere_template(r, n_grid=13, n_nodes=64))
    s.add_particle([0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.add_particle([0.3, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.initialize([-1, -1, -1], [1, 1, 1], [0, 0, 0], 1e-4)
    F, T = np.array([1e-3, -2e-3, 5e-4]), np.array([0.0, 1e-3, 0.0])
    s.set_constant_force(1, F, T)
    n = 50
    s.step(n)
    st = s.state()
    np.testing.assert_allclose(st["v"][1], (n - 0.5) * 1e-4 * (F + T) / m, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(st["v"][0], 0.0)
    np.testing.assert_array_equal(st["w"][1], 0.0)


def test_global_damping_adds_to_sliding_deceleration():
    kn, mu, c, g = 1e5, 0.3, 0.2, 9.81
    s, m = _sphere_probe(0.05, kn, mu=mu, ks=0.4 * kn, v0=(0.5, 0, 0), gravity=(0, 0, -g), dt=2e-6,
                         rotate=False)
    s.step(5000)  # settle vertically while sliding
    s.set_global_damping(c)
    s.step(200)
    v1 = s.state()["v"][0, 0]
    s.step(2000)
    v2 = s.state()["v"][0, 0]
    assert v2 > 0.1
    expect = mu * g + c * g * math.sqrt(1 + mu * mu)
    assert abs((v1 - v2) / (2000 * 2e-6) - expect) / expect < 2e-3
