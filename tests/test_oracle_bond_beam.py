"""Pins of the oracle's VBond shear, bending and torsion terms (the B2-B4
terms of contactKernel.cuh:138-199 with the stiffnesses of
VBondedInteraction.h:74-195) by beam theory.

Two massless spheres are bonded along x and held in place; the slave is
turned by theta, once about the bond axis and once about an axis
perpendicular to it. A clamped-clamped Euler-Bernoulli beam whose end turns
by theta carries the end moments 4EI/L theta (turned end) and 2EI/L theta
(far end) and the shear 6EI/L^2 theta; a shaft twisted by theta carries
GJ/L theta. The V-model reproduces these with sin(theta) in place of theta
(a x R(theta) a = sin(theta) times the axis for a unit vector a normal to
the axis), so they are checked at a finite angle to rounding and against
the small-angle textbook values."""
import math

import numpy as np
import pytest

from fundem_amd import generators as gen
from oracle import pyoracle as po

R = 0.5e-3
L = 2 * R
E = 1e9 * L
NU = 0.3
RB = L / math.sqrt(6)


def bonded_pair(q_slave):
    s = po.OracleSim(False, False)
    s.add_material(1e4, 4e3, 0.0, 0.5)
    t = gen.sphere_template(R, n_grid=16, n_nodes=64)
    s.add_geometry(0.0, [0] * 6, t)  # zero mass: poses stay as set
    s.add_particle([0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.add_particle([L, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.add_bond(0, 1, [L / 2, 0, 0], [-1, 0, 0], RB, L, E, NU, 0.0, 0.0, 0.0)
    s.initialize([-0.01, -0.01, -0.01], [0.01, 0.01, 0.01], [0, 0, 0], 1e-6)
    q = s.state()["q"].copy()
    q[1] = q_slave
    s.set_state(q=q)
    s.step(1)
    return s


def section():
    A = math.pi * RB * RB
    I = A * RB * RB / 4
    return A, I, 2 * I, E / (2 * (1 + NU))


def turn(axis, theta):
    a = np.asarray(axis, float)
    return np.concatenate([[math.cos(theta / 2)], math.sin(theta / 2) * a / np.linalg.norm(a)])


@pytest.mark.parametrize("theta", [1e-4, 0.3])
def test_bending_gives_the_clamped_beam_end_moments_and_shear(theta):
    st = bonded_pair(turn([0, 0, 1], theta)).state()
    A, I, J, G = section()
    s = math.sin(theta)
    scale = E * I / L
    # turned end: restoring 4EI/L; far end: carry-over 2EI/L, same sense
    np.testing.assert_allclose(st["T"][1], [0, 0, -4 * scale * s], rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(st["T"][0], [0, 0, -2 * scale * s], rtol=0, atol=1e-9 * scale)
    # shear 6EI/L^2 pulls the master toward the turned end's new axis (-y)
    V = 6 * E * I / L**2 * s
    np.testing.assert_allclose(st["F"][0], [0, -V, 0], rtol=0, atol=1e-9 * V)
    np.testing.assert_allclose(st["F"][1], [0, V, 0], rtol=0, atol=1e-9 * V)
    # angular momentum: the torques and the force couple balance about the origin
    total = st["T"][0] + st["T"][1] + np.cross([L, 0, 0], st["F"][1])
    assert np.abs(total).max() <= 1e-9 * scale
    if theta < 1e-3:  # the textbook small-angle values
        assert abs(-st["T"][1][2] - 4 * scale * theta) <= 1e-7 * 4 * scale * theta
    # stresses: bending moment 2EI/L at the bond, shear over the area
    _, sig, tau = po.OracleSim.bonds(bonded_pair(turn([0, 0, 1], theta)))
    assert sig[0] == pytest.approx(2 * scale * s / I * RB, rel=1e-9)
    assert tau[0] == pytest.approx(V / A, rel=1e-9)


@pytest.mark.parametrize("theta", [1e-4, 0.3])
def test_torsion_gives_the_shaft_torque(theta):
    st = bonded_pair(turn([1, 0, 0], theta)).state()
    A, I, J, G = section()
    Tq = G * J / L * math.sin(theta)
    np.testing.assert_allclose(st["T"][1], [-Tq, 0, 0], rtol=0, atol=1e-9 * Tq)
    np.testing.assert_allclose(st["T"][0], [Tq, 0, 0], rtol=0, atol=1e-9 * Tq)
    assert np.abs(st["F"]).max() <= 1e-9 * Tq / L
    _, sig, tau = po.OracleSim.bonds(bonded_pair(turn([1, 0, 0], theta)))
    assert abs(sig[0]) <= 1e-9 * Tq / I * RB
    assert tau[0] == pytest.approx(Tq / J * RB, rel=1e-9)
