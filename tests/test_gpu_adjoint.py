"""The adjoint timestep against central finite differences.

Centimetre superellipsoids in resting contact (one on the floor carrying a
second one, which touches it at a single node, and a third sliding on the
floor), with small random velocities, are run on the GPU solver, whose
adjoint (fundem_amd.Adjoint, checkpointed) gives the gradient of an outcome
of the final state with respect to the contact parameters (kn, ks, mu and
damping ratio of the particle material, mu and damping ratio of the wall
material) and to initial velocities, angular velocities and positions.
Central differences with a relative step of 1e-4 are taken on the FP64
oracle (store_fp32=0), which has no FP32 rounding noise; the scenario keeps
every contact closed and its stick/slip state fixed over the window, so the
trajectory is smooth in the parameters. Sensitivities theta * dJ/dtheta must
agree within 1e-3 relative (with a floor of 1e-4 of the largest one).
"""
import numpy as np
import pytest

import pyoracle as po
from fundem_amd import generators as gen
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

DT = 1e-5
PARAMS = dict(kn=1e3, ks=4e2, mu=0.4, zeta=0.3, wall_mu=0.3, wall_zeta=0.35)


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


_TM = {}


def _templates():
    """Two templates whose SDF values and nodes are FP32 numbers, with their
    mass properties (density 2500) given explicitly, so that the device
    (which stores SDF and nodes in FP32) and the FP64 oracle see the same
    geometry."""
    if not _TM:
        tm = [gen.superellipsoid_template(0.006, 0.005, 0.004, 3.0, 3.0, n_grid=24, n_nodes=400),
              gen.superellipsoid_template(0.005, 0.005, 0.0045, 2.5, 3.5, n_grid=24, n_nodes=400)]
        o = po.OracleSim(False, True)
        props = []
        for t in tm:
            t.sdf = _f32(t.sdf)
            t.nodes = _f32(t.nodes)
            g, _ = o.add_geometry_density(2500.0, t)
            info = o.geometry(g)
            props.append((1.0 / info["inv_mass"], info["inertia"].copy()))
        _TM["v"] = (tm, props)
    return _TM["v"]


L = 0.08


def _scenario(tm, x, q, v, w, p, periodic=False):
    tm, props = tm
    n = len(x)
    # the floor reaches past the box, under bodies that straddle a periodic face
    wt, wall = W.floor_wall([-L - 0.02, -L - 0.02], [L + 0.02, L + 0.02])
    return W.Scenario(
        name="adjoint-pile", templates=tm, density=2500.0,
        materials=[(p["kn"], p["ks"], p["mu"], p["zeta"])],
        x=np.array(x, dtype=np.float64), v=np.array(v, dtype=np.float64),
        w=np.array(w, dtype=np.float64), q=np.array(q, dtype=np.float64),
        mat=np.zeros(n, np.int32), geo=(np.arange(n) % len(tm)).astype(np.int32),
        domain_min=np.array([-L, -L, -L]), domain_max=np.array([L, L, L]),
        gravity=np.array([0.0, 0.0, -9.81]), dt=DT,
        wall_templates=[wt], wall_materials=[(p["wall_mu"], p["wall_zeta"])], walls=[wall],
        periodic=(bool(periodic), False), damping_ratio=True, mass_properties=props)


def _rot(q):
    q0, q1, q2, q3 = q / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
        [2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1)],
        [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)]])


def _pair_contacts(tm, x, q, periodic):
    o = po.OracleSim(False, True)
    W.apply(_scenario(tm, x, q, np.zeros((len(x), 3)), np.zeros((len(x), 3)), PARAMS, periodic), o)
    return o.contacts(po.SET_GX if periodic else po.SET_PP)["A"].size


_PILE = {}


def resting_pile(periodic=False):
    """Particle 0 rests on the floor, particle 1 rests on top of particle 0,
    particle 2 sits on the floor apart from them. Each floor overlap is a
    third of the one that carries the weight above it on a single node, and
    particle 1 is lowered onto particle 0 by a third of its own such
    overlap from the height where the two first touch (found by bisection
    on the oracle's detection), so every contact stays closed while the
    bodies oscillate about equilibrium over the test window. Periodic in x:
    particle 0 sits just inside the -x face and particle 1 just inside the
    +x face, on particle 0's periodic image (a ghost-set contact)."""
    if periodic in _PILE:
        return _PILE[periodic]
    tmp = _templates()
    tm, props = tmp
    rng = np.random.default_rng(11)
    q = gen.shoemake_quaternions(rng, 3)
    geo = np.arange(3) % len(tm)
    m = np.array([props[k][0] for k in geo])
    lowest = np.array([np.min((tm[geo[i]].nodes.astype(np.float64) @ _rot(q[i]).T)[:, 2])
                       for i in range(3)])
    gz = 9.81
    d0 = (m[0] + m[1]) * gz / PARAMS["kn"] / 3.0
    d2 = m[2] * gz / PARAMS["kn"] / 3.0
    x = np.array([[0.0, 0.0, -lowest[0] - d0], [0.0005, -0.0003, 0.05], [0.02, 0.01, -lowest[2] - d2]])
    if periodic:
        x[0, 0] = -L + 0.0002
        x[1, 0] = L - 0.0003
    lo, hi = x[0, 2], x[0, 2] + 0.03  # particle 1 height: touching at lo, apart at hi
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        x[1, 2] = mid
        if _pair_contacts(tmp, x, q, periodic) > 0:
            lo = mid
        else:
            hi = mid
    d1 = m[1] * gz / (0.5 * PARAMS["kn"]) / 3.0
    x[1, 2] = hi - d1
    _PILE[periodic] = (tmp, x, q)
    return _PILE[periodic]


def objectives(state, masses, coef):
    """Two outcomes of a final state: kinetic energy of translation, and a
    fixed linear functional of positions, angular velocities and orientations."""
    ke = 0.5 * float(np.sum(masses[:, None] * state["v"] ** 2))
    lin = float(np.sum(coef["x"] * state["x"]) + np.sum(coef["w"] * state["w"]) +
                np.sum(coef["q"] * state["q"]))
    return ke, lin


def objective_adjoints(state, masses, coef):
    return [dict(x=np.zeros_like(state["x"]), q=np.zeros_like(state["q"]),
                 v=masses[:, None] * state["v"], w=np.zeros_like(state["w"])),
            dict(x=coef["x"], q=coef["q"], v=np.zeros_like(state["v"]), w=coef["w"])]


def run_oracle(tm, x, q, v, w, p, steps, periodic=False):
    o = po.OracleSim(False, True)
    W.apply(_scenario(tm, x, q, v, w, p, periodic), o)
    o.step(steps)
    return o.state(), o


@pytest.mark.parametrize("steps,every,periodic", [(60, 20, False), (60, 20, True)])
def test_adjoint_matches_finite_differences(steps, every, periodic):
    from fundem_amd import Adjoint, Solver
    tm, x, q = resting_pile(periodic)
    n = len(x)
    rng = np.random.default_rng(5)
    v0 = rng.normal(0.0, 1e-5, size=(n, 3))
    w0 = rng.normal(0.0, 1e-3, size=(n, 3))
    v0[2] = [0.05, 0.01, 0.0]  # particle 2 slides
    coef = dict(x=rng.normal(0, 1e-4, size=(n, 3)), w=rng.normal(0, 1e-8, size=(n, 3)),
                q=rng.normal(0, 1e-8, size=(n, 4)))
    # GPU run with its adjoint
    g = Solver()
    W.apply(_scenario(tm, x, q, v0, w0, PARAMS, periodic), g)
    adj = Adjoint(g)
    adj.forward(steps, every)
    assert g.status() == 0
    pair_set = po.SET_GX if periodic else po.SET_PP
    assert g.contacts(pair_set)["A"].size > 0 and g.contacts(po.SET_PW)["A"].size > 0
    fin = g.state()
    ref, o = run_oracle(tm, x, q, v0, w0, PARAMS, steps, periodic)
    masses = np.array([tm[1][k][0] for k in (np.arange(n) % len(tm[0]))])
    grads = [adj.backward(a["x"], a["q"], a["v"], a["w"]) for a in objective_adjoints(fin, masses, coef)]
    # the same adjoint without intermediate checkpoints, bit for bit
    g2 = Solver()
    W.apply(_scenario(tm, x, q, v0, w0, PARAMS, periodic), g2)
    adj2 = Adjoint(g2)
    adj2.forward(steps, steps)
    a0 = objective_adjoints(g2.state(), masses, coef)[0]
    one = adj2.backward(a0["x"], a0["q"], a0["v"], a0["w"])
    for key in ("v", "w", "x", "q", "kn", "ks", "mu", "zeta", "wall_mu", "wall_zeta"):
        assert np.array_equal(one[key], grads[0][key]), key

    checks = []  # (name, theta scale, adjoint per objective, perturbation function)
    for name in ("kn", "ks", "mu", "zeta", "wall_mu", "wall_zeta"):
        gk = name if name in grads[0] else name
        checks.append((name, PARAMS[name], [gr[gk][0] for gr in grads], ("param", name)))
    for (i, c) in ((0, 0), (1, 2), (2, 0)):
        checks.append((f"v0[{i},{c}]", 1e-2, [gr["v"][i, c] for gr in grads], ("v", i, c)))
    checks.append(("w0[1,0]", 1.0, [gr["w"][1, 0] for gr in grads], ("w", 1, 0)))
    checks.append(("x0[1,2]", 1e-6, [gr["x"][1, 2] for gr in grads], ("x", 1, 2)))

    def perturbed(spec, delta):
        p, xx, vv, ww = dict(PARAMS), x.copy(), v0.copy(), w0.copy()
        if spec[0] == "param":
            p[spec[1]] += delta
        elif spec[0] == "v":
            vv[spec[1], spec[2]] += delta
        elif spec[0] == "w":
            ww[spec[1], spec[2]] += delta
        else:
            xx[spec[1], spec[2]] += delta
        st, _ = run_oracle(tm, xx, q, vv, ww, p, steps, periodic)
        return np.array(objectives(st, masses, coef))

    base = np.array(objectives(ref, masses, coef))
    rows = []
    for name, scale, adjv, spec in checks:
        h = 1e-4 * scale
        jp, jm = perturbed(spec, h), perturbed(spec, -h)
        fd = (jp - jm) / (2 * h)
        one_sided = ((jp - base) / h, (base - jm) / h)
        rows.append((name, scale, np.array(adjv), fd, one_sided))
    for k in range(2):
        top = max(abs(r[1] * r[3][k]) for r in rows)
        for name, scale, adjv, fd, (fp, fm) in rows:
            sa, sf = scale * adjv[k], scale * fd[k]
            print(f"J{k} {name:10s} adjoint {sa: .6e} central {sf: .6e} one-sided {scale * fp[k]: .4e} "
                  f"{scale * fm[k]: .4e}")
            assert abs(sa - sf) <= 1e-3 * max(abs(sf), 1e-4 * top), (k, name, sa, sf)
