"""The forward-mode derivative tool (oracle/tangent.py) against what fixes
it independently:

* free flight under gravity: after n steps, dx/dv0 = n dt, dv/dv0 = 1,
  dx/dx0 = 1 exactly, and the derivatives along parameters vanish;
* the resting pile of test_gpu_adjoint.py (every contact closed, stick/slip
  fixed over the window, so the trajectory is smooth): directional
  derivatives of kinetic energy and runout agree with central differences
  of the FP64 oracle. Along kn and ks the pile is run without damping,
  because the damping coefficients take FP32 square roots of kn m_eff and
  ks m_eff, whose rounding makes central differences along kn and ks
  unreliable below about 1e-3.

The tool's restated values also track the oracle's state to rounding."""
import dataclasses

import numpy as np

import pyoracle as po
import tangent as tg
from fundem_amd import objectives as O
from fundem_amd import workloads as W
from test_gpu_adjoint import PARAMS, _scenario, resting_pile

STEPS = 60


def test_free_flight_closed_form():
    scn = W.cluster(n=1, walls=False, speed=0.3, spin=0.0)
    dirs = [("v", 0, 0), ("v", 0, 2), ("x", 0, 1), ("kn", 0), ("zeta", 0)]
    t = tg.Tangent(scn, dirs)
    n = 40
    t.step(n)
    st = t.state()
    dx, dv = st["dx"][0], st["dv"][0]
    assert np.allclose(dx[:, 0], [n * scn.dt, 0, 0], rtol=1e-12, atol=1e-18)
    assert np.allclose(dx[:, 1], [0, 0, n * scn.dt], rtol=1e-12, atol=1e-18)
    assert np.allclose(dx[:, 2], [0, 1, 0], rtol=0, atol=1e-15)
    assert np.allclose(dv[:, 0], [1, 0, 0], atol=1e-15) and np.allclose(dv[:, 1], [0, 0, 1], atol=1e-15)
    assert np.all(st["dx"][:, :, 3:] == 0.0) and np.all(st["dv"][:, :, 3:] == 0.0)
    assert t.max_mismatch <= 1e-15, t.mismatch


def _pile(damped):
    tm, x, q = resting_pile(False)
    n = len(x)
    rng = np.random.default_rng(8)
    v0 = rng.normal(0.0, 1e-5, size=(n, 3))
    w0 = rng.normal(0.0, 1e-3, size=(n, 3))
    v0[2] = [0.05, 0.01, 0.0]
    p = dict(PARAMS) if damped else dict(PARAMS, zeta=0.0, wall_zeta=0.0)
    masses = np.array([tm[1][k][0] for k in (np.arange(n) % len(tm[0]))])
    return tm, x, q, v0, w0, p, masses


def _outcomes(st, masses, centre):
    return [O.kinetic_energy(st, masses), O.runout(st, centre, tau=1e-3)]


def _check(damped, dirs, rtol):
    tm, x, q, v0, w0, p, masses = _pile(damped)
    scn = _scenario(tm, x, q, v0, w0, p)
    centre = x.mean(axis=0)
    t = tg.Tangent(scn, dirs)
    t.step(STEPS)
    fin = t.state()
    got = [t.directional(bar) for _, bar in _outcomes(fin, masses, centre)]

    def J(spec, delta):
        pp, xx, vv, ww = dict(p), x.copy(), v0.copy(), w0.copy()
        if spec[0] in ("x", "v", "w"):
            {"x": xx, "v": vv, "w": ww}[spec[0]][spec[1], spec[2]] += delta
        else:
            pp[spec[0]] += delta
        o = po.OracleSim(False, True)
        W.apply(_scenario(tm, xx, q, vv, ww, pp), o)
        o.step(STEPS)
        return np.array([j for j, _ in _outcomes(o.state(), masses, centre)])

    rows = []
    for d, spec in enumerate(dirs):
        scale = p[spec[0]] if spec[0] in p else (1e-2 if spec[0] in ("v", "x") else 1.0)
        h = 1e-5 * scale
        fd = (J(spec, h) - J(spec, -h)) / (2 * h)
        rows.append((spec, scale * np.array([g[d] for g in got]), scale * fd))
    for k, name in enumerate(("kinetic energy", "runout")):
        # sensitivities theta dJ/dtheta; the central differences' rounding
        # noise is below 1e-6 of the largest
        top = max(abs(r[2][k]) for r in rows)
        assert top > 0.0
        for spec, a, f in rows:
            print(f"{name:15s} {str(spec):15s} forward mode {a[k]: .9e} central {f[k]: .9e}", flush=True)
            assert abs(a[k] - f[k]) <= rtol * abs(f[k]) + 1e-6 * top, (name, spec, a[k], f[k])
    assert t.max_mismatch <= 1e-12, t.mismatch


def test_pile_along_damping_friction_and_initial_state():
    _check(True, [("zeta", 0), ("wall_zeta", 0), ("mu", 0), ("wall_mu", 0), ("v", 2, 0), ("v", 1, 2),
                  ("w", 1, 0), ("x", 1, 2)], 1e-5)


def test_pile_along_stiffness_without_damping():
    _check(False, [("kn", 0), ("ks", 0), ("v", 2, 0)], 1e-5)
