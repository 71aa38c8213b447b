"""The oracle's sector periodic boundary (PeriodicBoundarySector,
PeriodicBoundary.h:319-426; mySpatialGrid.cu:59-85,154-212) pinned by
geometry, in FP64:
  * a particle leaving through the x = min face comes back through the
    y = min face turned by -pi/2 about the z axis through min (position,
    velocity and orientation), and the reverse through y = min;
  * an R90 or R270 image interacts exactly as a real particle placed at the
    image's pose would: the node contacts of the ghost set (node, overlap,
    normal, point, spring after one step) equal those of an explicit pair in
    an open domain."""
import math

import numpy as np
import pytest

import pyoracle as po
from fundem_amd import generators as gen
from fundem_amd import workloads as W

R = 1e-2


def sphere_scene(x, v, w, q, lo, hi, sector):
    t = gen.sphere_template(R, n_grid=16, n_nodes=400)
    n = len(x)
    return W.Scenario(
        name="sector", templates=[t], density=2500.0, materials=[(1e4, 4e3, 0.4, 0.5)],
        x=np.asarray(x, float), v=np.asarray(v, float), w=np.asarray(w, float),
        q=np.asarray(q, float), mat=np.zeros(n, np.int32), geo=np.zeros(n, np.int32),
        domain_min=np.asarray(lo, float), domain_max=np.asarray(hi, float),
        gravity=np.zeros(3), dt=1e-5, sector=sector, steps=1)


def rot_z(v, th):
    c, s = math.cos(th), math.sin(th)
    v = np.asarray(v, float)
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1], v[2]])


def qmul(a, b):
    a0, a1, a2, a3 = a
    b0, b1, b2, b3 = b
    return np.array([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3, a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                     a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1, a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0])


def qz(th):
    return np.array([math.cos(th / 2), 0.0, 0.0, math.sin(th / 2)])


@pytest.mark.parametrize("face", ["x", "y"])
def test_sector_wrap_turns_the_particle(face):
    L = 8 * R
    q0 = np.array([0.9, 0.1, -0.3, 0.2])
    q0 /= np.linalg.norm(q0)
    if face == "x":  # leaves through x = min: turned by -pi/2
        x0, v0, th = [1e-6, 3 * R, 4 * R], [-1.0, 0.2, 0.1], -math.pi / 2
    else:  # leaves through y = min: turned by +pi/2
        x0, v0, th = [3 * R, 1e-6, 4 * R], [0.2, -1.0, 0.1], math.pi / 2
    scn = sphere_scene([x0], [v0], [[0.0, 0.0, 0.0]], [q0], [0, 0, 0], [L, L, L], True)
    o = po.OracleSim(False, False)
    W.apply(scn, o)
    o.step(1)
    st = o.state()
    pre = np.asarray(x0) + scn.dt * np.asarray(v0)
    np.testing.assert_allclose(st["x"][0], rot_z(pre, th), rtol=0, atol=1e-15)
    np.testing.assert_allclose(st["v"][0], rot_z(v0, th), rtol=0, atol=1e-14)
    np.testing.assert_allclose(st["q"][0], qmul(qz(th), q0), rtol=0, atol=1e-14)
    assert st["x"][0][0] >= 0 and st["x"][0][1] >= 0


def test_ghost_contacts_equal_explicit_pairs():
    L = 8 * R
    qa = np.array([0.8, 0.2, 0.1, -0.3]); qa /= np.linalg.norm(qa)
    qb = np.array([0.6, -0.4, 0.5, 0.2]); qb /= np.linalg.norm(qb)
    xa, xb = np.array([1.55 * R, 3 * R, 4 * R]), np.array([3 * R, 0.4 * R, 4 * R])
    va, vb = np.array([0.1, 0.0, 0.0]), np.array([0.0, 0.2, 0.05])
    zero = np.zeros(3)
    scn = sphere_scene([xa, xb], [va, vb], [zero, zero], [qa, qb], [0, 0, 0], [L, L, L], True)
    o = po.OracleSim(False, False)
    W.apply(scn, o)
    o.step(1)
    r90, r180, r270 = (o.contacts(s) for s in (po.SET_GX, po.SET_GY, po.SET_GXY))
    assert r180["A"].size == 0
    assert r90["A"].size > 0 and np.all(r90["A"] == 0) and np.all(r90["B"] == 1)
    assert r270["A"].size > 0 and np.all(r270["A"] == 1) and np.all(r270["B"] == 0)
    # the same pairs with the images as real particles, in an open domain
    for got, (i, j, th) in ((r90, (0, 1, math.pi / 2)), (r270, (1, 0, -math.pi / 2))):
        xs, vs, qs = [xa, xb], [va, vb], [qa, qb]
        xj = rot_z(xs[j], th)  # min is the origin
        e = sphere_scene([xs[i], xj], [vs[i], rot_z(vs[j], th)], [zero, zero],
                         [qs[i], qmul(qz(th), qs[j])], [-4 * R, -4 * R, 0], [L, L, L], False)
        oe = po.OracleSim(False, False)
        W.apply(e, oe)
        oe.step(1)
        ref = oe.contacts(po.SET_PP)
        assert np.array_equal(ref["k"], got["k"])
        for key in ("overlap", "normal", "point", "spring"):
            np.testing.assert_allclose(got[key], ref[key], rtol=0, atol=1e-12 * R, err_msg=key)
