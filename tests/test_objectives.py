"""The outcomes the adjoint is taken of (fundem_amd.objectives): values in
closed form and their adjoints against central differences."""
import numpy as np

from fundem_amd import objectives as O


def state(n, seed=0):
    rng = np.random.default_rng(seed)
    return dict(x=rng.normal(0.0, 0.01, size=(n, 3)) + [0.0, 0.0, 0.02],
                q=rng.normal(size=(n, 4)), v=rng.normal(0.0, 0.1, size=(n, 3)),
                w=rng.normal(size=(n, 3)))


def central(f, st, key, i, c, h):
    a, b = {k: v.copy() for k, v in st.items()}, {k: v.copy() for k, v in st.items()}
    a[key][i, c] += h
    b[key][i, c] -= h
    return (f(a) - f(b)) / (2 * h)


def check_adjoint(f, st, h):
    J, bar = f(st)
    for key in ("x", "v"):
        for i in range(len(st["x"])):
            for c in range(3):
                fd = central(lambda s: f(s)[0], st, key, i, c, h)
                assert abs(bar[key][i, c] - fd) <= 1e-6 * max(1.0, abs(fd)), (key, i, c, bar[key][i, c], fd)


def test_kinetic_energy_value_and_adjoint():
    st = state(6)
    m = np.linspace(1.0, 2.0, 6)
    moving = np.array([1, 1, 0, 1, 1, 1], bool)
    J, bar = O.kinetic_energy(st, m, moving)
    assert np.isclose(J, 0.5 * sum(m[i] * st["v"][i] @ st["v"][i] for i in range(6) if moving[i]))
    assert np.all(bar["v"][2] == 0.0)
    check_adjoint(lambda s: O.kinetic_energy(s, m, moving), st, 1e-6)


def test_runout_max_and_smooth_max():
    st = state(7, seed=3)
    c = np.array([0.001, -0.002, 0.0])
    r = np.linalg.norm(st["x"][:, :2] - c[:2], axis=1)
    J, bar = O.runout(st, c)
    k = int(np.argmax(r))
    assert J == r[k]
    assert np.allclose(bar["x"][k, :2], (st["x"][k, :2] - c[:2]) / r[k])
    assert np.count_nonzero(bar["x"]) == 2
    tau = 1e-3
    Js, _ = O.runout(st, c, tau=tau)
    assert r.max() <= Js <= r.max() + tau * np.log(len(r)) + 1e-15
    check_adjoint(lambda s: O.runout(s, c, tau=tau), st, 1e-7)
    # a fixed particle farther out is ignored
    moving = np.ones(7, bool)
    moving[k] = False
    J2, bar2 = O.runout(st, c, moving)
    assert J2 == np.sort(r)[-2] and np.all(bar2["x"][k] == 0.0)


def test_packing_fraction_of_a_uniform_bed():
    # a simple cubic lattice of spheres of radius a at spacing 2a, nz layers
    # resting on the floor: mean centre height is nz a, H = 2 nz a, and the
    # fraction is the lattice's pi / 6
    a, nx, nz = 1e-3, 5, 4
    pts = np.array([[(i + 0.5) * 2 * a, (j + 0.5) * 2 * a, (k + 0.5) * 2 * a]
                    for k in range(nz) for j in range(nx) for i in range(nx)])
    st = dict(x=pts, q=np.zeros((len(pts), 4)), v=np.zeros_like(pts), w=np.zeros_like(pts))
    vol = np.full(len(pts), 4.0 / 3.0 * np.pi * a ** 3)
    J, _ = O.packing_fraction(st, vol, (2 * a * nx) ** 2, 0.0)
    assert np.isclose(J, np.pi / 6.0, rtol=1e-14)
    st2 = state(8, seed=5)
    v2 = np.linspace(1e-9, 2e-9, 8)
    check_adjoint(lambda s: O.packing_fraction(s, v2, 4e-4, -0.001), st2, 1e-7)
