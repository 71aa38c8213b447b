"""The oracle, pinned without the reference (which has no tests, golden
vectors or CPU path): definitions checked by brute force, closed-form
mechanics, conservation laws and cross-checks between the oracle's modes.
CPU only."""
import math

import numpy as np
import pytest

import pyoracle as po
from fundem_amd import generators as gen
from fundem_amd import workloads as W


# ------------------------------------------------------------------- grid
def test_grid_set_rules():
    mn, mx, size, inv = po.grid_set([0, 0, 0], [1.0, 0.3, 0.1], 0.1)
    # size = int(L / cs) when L > 2 cs, else 2; inv = size / L
    assert list(size) == [10, 2, 2]
    np.testing.assert_array_equal(inv, [10 / 1.0, 2 / 0.3, 2 / 0.1])
    mn, mx, size, inv = po.grid_set([0, 0, 0], [1, 1, 1], 0.0)
    assert list(size) == [2, 2, 2] and np.all(inv == 0)
    mn, mx, size, inv = po.grid_set([0, 0, 0], [0, 1, 1], 0.25)  # degenerate axis grows by cs
    assert mx[0] == 0.25


def _brute_hash(pos, mn, mx, inv, size):
    out = np.full(len(pos), -1, np.int64)
    inside = np.all((pos >= mn) & (pos < mx), axis=1)
    c = np.trunc((pos - mn) * inv).astype(np.int64)
    h = c[:, 2] * size[1] * size[0] + c[:, 1] * size[0] + c[:, 0]
    out[inside] = h[inside]
    return out


def test_hash_sort_start_end_by_definition():
    rng = np.random.default_rng(0)
    mn, mx, size, inv = po.grid_set([-1, 0, 2], [3, 1, 4], 0.05)
    pos = rng.uniform([-1.2, -0.1, 1.9], [3.2, 1.1, 4.1], size=(50000, 3))
    pos[:100] = np.round(pos[:100] / 0.05) * 0.05  # exact cell faces
    h = po.hash_(pos, mn, mx, inv, size)
    np.testing.assert_array_equal(h, _brute_hash(pos, mn, mx, inv, size))
    ng = int(np.prod(size))
    hs, idx, st, en = po.sort_start_end(h, ng)
    order = np.argsort(h, kind="stable")  # stable by (hash, index)
    np.testing.assert_array_equal(idx, order)
    np.testing.assert_array_equal(hs, h[order])
    exp_st, exp_en = np.full(ng, -1), np.full(ng, -1)
    for i, v in enumerate(hs):
        if 0 <= v < ng:
            if exp_st[v] < 0:
                exp_st[v] = i
            exp_en[v] = i + 1
    np.testing.assert_array_equal(st, exp_st)
    np.testing.assert_array_equal(en, exp_en)


def test_ghost_hash_clamps_and_uses_y_bound_for_z():
    mn, mx, size, inv = po.grid_set([0, 0, 0], [1, 1, 2], 0.25)
    pts = np.array([[1.1, 0.5, 0.5],    # beyond max.x -> clamped to the last x cell
                    [0.5, 0.5, 0.5],    # inside -> -1
                    [0.5, 0.5, 1.5],    # inside the z range but z >= max.y: out of bounds
                    [-0.1, -0.1, 0.1]])
    h = po.hash_(pts, mn, mx, inv, size, ghost=True)
    nx, ny = size[0], size[1]
    assert h[0] == 2 * nx * ny + 2 * nx + (nx - 1)
    assert h[1] == -1
    assert h[2] == (size[2] - 1) * nx * ny + 2 * nx + 2
    assert h[3] == 0


def test_inclusive_scan():
    c = np.random.default_rng(1).integers(0, 9, size=10001).astype(np.int32)
    np.testing.assert_array_equal(po.inclusive_scan(c), np.cumsum(c))


# -------------------------------------------------------- contact mechanics
def _sphere_probe(e, kn, mu=0.0, ks=0.0, v0=(0, 0, -1.0), dt=1e-6, node=(0, 0, -0.01), steps=200000,
                  gravity=(0, 0, 0), rotate=True, zeta=None):
    r, rho = 0.01, 2000.0
    m = 4 / 3 * math.pi * r ** 3 * rho
    I = 0.4 * m * r * r if rotate else 0.0
    s = po.OracleSim(False, False)
    if zeta is None:
        s.add_material(kn, ks, mu, e)
    else:
        s.add_material_damping(kn, ks, mu, zeta)
    t = gen.sphere_template(r, n_grid=13, n_nodes=400)
    t.nodes = np.array([node], np.float32)
    s.add_geometry(m, [I, I, I, 0, 0, 0], t)
    if zeta is None:
        s.add_wall_material(mu, e)
    else:
        s.add_wall_material_damping(mu, zeta)
    s.add_wall_geometry(gen.plane_wall([-0.5, -0.5], [0.5, 0.5]))
    s.add_particle([0, 0, r + 1e-5], v0, [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.add_wall([0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.initialize([-0.5, -0.5, -1.0], [0.5, 0.5, 1.0], gravity, dt)
    return s, m


@pytest.mark.parametrize("e,damping", [(0.3, False), (0.6, False), (0.9, False), (0.3, True), (0.6, True)])
def test_single_contact_restitution_and_duration(e, damping):
    """Linear spring-dashpot: rebound ratio e and contact time pi/omega_d,
    for a restitution material and for a damping-ratio material with the
    matching zeta = -ln e / sqrt(ln^2 e + pi^2) (whose dashpot then gives
    back e)."""
    kn = 1e5
    z = -math.log(e) / math.sqrt(math.log(e) ** 2 + math.pi ** 2)
    s, m = _sphere_probe(e, kn, zeta=z if damping else None)
    started, n_in = False, 0
    for _ in range(20000):
        s.step(1)
        c = s.contacts(po.SET_PW)
        inc = c["A"].size > 0 and c["overlap"][0] > 0
        n_in += inc
        started |= inc
        if started and not inc:
            break
    vz = s.state()["v"][0, 2]
    zeta = -math.log(e) / math.sqrt(math.log(e) ** 2 + math.pi ** 2)
    wd = math.sqrt(kn / m) * math.sqrt(1 - zeta * zeta)
    # discretisation error is O(omega_n dt) ~ 4e-4 at these settings
    assert abs(vz - e) < 1e-3
    assert abs(n_in * 1e-6 - math.pi / wd) < 2e-6


def test_static_rest_balances_gravity():
    """A resting single-node sphere: kn * delta = m g at equilibrium."""
    kn = 1e5
    s, m = _sphere_probe(0.05, kn, v0=(0, 0, 0), gravity=(0, 0, -9.81), dt=2e-6)
    s.step(30000)
    c = s.contacts(po.SET_PW)
    assert c["A"].size == 1
    assert abs(kn * c["overlap"][0] - m * 9.81) / (m * 9.81) < 1e-3
    assert abs(s.state()["F"][0, 2] - m * 9.81) / (m * 9.81) < 1e-3


def test_sliding_friction_decelerates_at_mu_g_and_respects_cap():
    """A single-node block sliding on the floor (zero inertia, so it cannot
    roll): the friction force sits at its cap mu |F_n| and the block
    decelerates at mu g."""
    kn, mu = 1e5, 0.3
    s, m = _sphere_probe(0.05, kn, mu=mu, ks=0.4 * kn, v0=(0.5, 0, 0), gravity=(0, 0, -9.81), dt=2e-6,
                         rotate=False)
    s.step(5000)  # settle vertically while sliding
    v1 = s.state()["v"][0, 0]
    s.step(5000)
    st = s.state()
    v2 = st["v"][0, 0]
    F = st["F"][0]
    Fn = m * 9.81
    assert v2 > 0.1  # still sliding
    assert abs(F[0]) <= mu * math.hypot(F[2], 0.0) * (1 + 1e-9)
    assert abs(abs(F[0]) - mu * Fn) / (mu * Fn) < 1e-3
    assert abs((v1 - v2) / (5000 * 2e-6) - mu * 9.81) / (mu * 9.81) < 1e-3


def test_free_fall_is_exact_under_kick_drift_kick():
    s, m = _sphere_probe(0.5, 1e5, v0=(0.1, 0, 0.2), gravity=(0, 0, -9.81), dt=1e-4)
    x0 = s.state()["x"][0].copy()
    n = 100
    s.step(n)
    t = n * 1e-4
    x = s.state()["x"][0]
    exp = x0 + np.array([0.1, 0, 0.2]) * t + 0.5 * np.array([0, 0, -9.81]) * t * t
    np.testing.assert_allclose(x, exp, rtol=0, atol=1e-12)


def test_quaternion_spin_about_principal_axis():
    r, rho = 0.01, 2000.0
    m = 4 / 3 * math.pi * r ** 3 * rho
    I = 0.4 * m * r * r
    s = po.OracleSim(False, False)
    s.add_material(1e5, 0, 0, 0.5)
    t = gen.sphere_template(r, n_grid=13, n_nodes=64)
    s.add_geometry(m, [I, I, I, 0, 0, 0], t)
    w = 3.0
    s.add_particle([0, 0, 0.5], [0, 0, 0], [0, 0, w], [1, 0, 0, 0], 0, 0)
    dt, n = 1e-4, 5000
    s.initialize([-1, -1, -1], [1, 1, 1], [0, 0, 0], dt)
    s.step(n)
    q = s.state()["q"][0]
    angle = 2 * math.atan2(q[3], q[0])
    assert abs(angle - w * n * dt) < 5 * w * w * dt * n * dt
    np.testing.assert_allclose(s.state()["w"][0], [0, 0, w], atol=1e-12)


def test_sphere_mass_properties_converge():
    r, rho = 0.01, 2000.0
    s = po.OracleSim(False, False)
    t = gen.sphere_template(r, n_grid=48, n_nodes=64)
    g, off = s.add_geometry_density(rho, t)
    info = s.geometry(g)
    m = 1.0 / info["inv_mass"]
    m_exact = 4 / 3 * math.pi * r ** 3 * rho
    assert abs(m - m_exact) / m_exact < 0.01
    I = info["inertia"]
    assert abs(I[0] - 0.4 * m * r * r) / (0.4 * m * r * r) < 0.02
    Im = np.array([[I[0], I[3], I[4]], [I[3], I[1], I[5]], [I[4], I[5], I[2]]])
    Ii = info["inv_inertia"]
    Iim = np.array([[Ii[0], Ii[3], Ii[4]], [Ii[3], Ii[1], Ii[5]], [Ii[4], Ii[5], Ii[2]]])
    np.testing.assert_allclose(Im @ Iim, np.eye(3), atol=1e-9)
    assert np.all(np.abs(off) < 1e-9)


def test_small_inertia_inverts_to_zero():
    """The reference's inverse returns zero below |det| 1e-30 (SI units), so
    millimetre particles do not rotate."""
    s = po.OracleSim(False, False)
    t = gen.sphere_template(1e-3, n_grid=16, n_nodes=64)
    g, _ = s.add_geometry_density(2500.0, t)
    assert np.all(s.geometry(g)["inv_inertia"] == 0.0)


# ------------------------------------------------------- detection & modes
def test_grid_detection_equals_bruteforce():
    scn = W.cluster(n=27, periodic=(False, False), speed=1.0)
    o = po.OracleSim(True, True)
    W.apply(scn, o)
    for _ in range(8):
        o.step(25)
        c = o.contacts(po.SET_PP)
        a, b, k = o.bruteforce_pp()
        np.testing.assert_array_equal(c["A"], a)
        np.testing.assert_array_equal(c["B"], b)
        np.testing.assert_array_equal(c["k"], k)


def test_pair_forces_cancel():
    """Without walls and gravity, contact forces cancel in pairs: the total
    force is zero to rounding."""
    scn = W.cluster(n=27, periodic=(False, False), walls=False, speed=1.0)
    scn.gravity = np.zeros(3)
    o = po.OracleSim(False, True)
    W.apply(scn, o)
    seen = 0
    for _ in range(40):
        o.step(10)
        F = o.state()["F"]
        scale = np.abs(F).max()
        if scale > 0:
            seen += 1
            assert np.abs(F.sum(axis=0)).max() <= 1e-9 * scale * len(F)
    assert seen > 0


@pytest.mark.parametrize("periodic", [(False, False), (True, True)])
def test_fp32_storage_and_canonical_fold_stay_close_to_fp64(periodic):
    scn = W.cluster(n=27, periodic=periodic, speed=1.0)
    a = po.OracleSim(False, False)
    b = po.OracleSim(True, True)
    W.apply(scn, a)
    W.apply(scn, b)
    for _ in range(4):
        a.step(25)
        b.step(25)
        sa, sb = a.state(), b.state()
        np.testing.assert_allclose(sb["x"], sa["x"], rtol=1e-5, atol=1e-9)
        # FP32 positions move contact overlaps by ~1 ulp; velocities agree to
        # a small fraction of the velocity scale
        assert np.abs(sb["v"] - sa["v"]).max() <= 1e-4 * np.abs(sa["v"]).max()


# ------------------------------------------------------------------- bonds
def _two_spheres_bond(gap_extra, st=0.0, coh=0.0):
    R = 0.5e-3
    L = 2 * R
    s = po.OracleSim(False, False)
    s.add_material(1e4, 4e3, 0.0, 0.5)
    t = gen.sphere_template(R, n_grid=16, n_nodes=64)
    s.add_geometry(0.0, [0] * 6, t)  # zero mass: positions stay as set
    s.add_particle([0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    s.add_particle([L, 0, 0], [gap_extra, 0, 0], [0, 0, 0], [1, 0, 0, 0], 0, 0)
    E = 1e9 * L
    Rb = L / math.sqrt(6)
    s.add_bond(0, 1, [L / 2, 0, 0], [-1, 0, 0], Rb, L, E, 0.3, st, coh, 0.0)
    s.initialize([-0.01, -0.01, -0.01], [0.01, 0.01, 0.01], [0, 0, 0], 1e-6)
    return s, Rb, L, E


def test_bond_axial_force_is_b1_times_extension():
    s, Rb, L, E = _two_spheres_bond(0.1)  # slave drifts away at 0.1 m/s
    s.step(1)
    st = s.state()
    D = st["x"][1, 0] - st["x"][0, 0]
    B1 = E * math.pi * Rb * Rb / L
    # force on the master points toward the slave: F = B1 (D - L) along +x
    assert abs(st["F"][0, 0] - B1 * (D - L)) / (B1 * abs(D - L)) < 1e-9
    assert abs(st["F"][1, 0] + st["F"][0, 0]) < 1e-12
    act, sig, tau = s.bonds()
    A = math.pi * Rb * Rb
    assert abs(sig[0] - B1 * (D - L) / A) / abs(B1 * (D - L) / A) < 1e-9
    assert act[0] == 1


def test_bond_breaks_when_tensile_strength_exceeded_and_stays_broken():
    s, Rb, L, E = _two_spheres_bond(1.0, st=1e5)
    steps = 0
    while s.bonds()[0][0] == 1 and steps < 1000:
        s.step(1)
        steps += 1
    D = s.state()["x"][1, 0] - s.state()["x"][0, 0]
    B1 = E * math.pi * Rb * Rb / L
    sig_now = B1 * (D - L) / (math.pi * Rb * Rb)
    assert s.bonds()[0][0] == 0
    assert sig_now > 1e5  # failed at the first step past the strength
    assert B1 * (D - 1e-6 - L) / (math.pi * Rb * Rb) <= 1e5 * (1 + 1e-9)
    F_break = s.state()["F"][0, 0]
    assert F_break != 0.0  # force still applied on the failing step
    s.step(5)
    assert s.bonds()[0][0] == 0
    assert s.state()["F"][0, 0] == 0.0
