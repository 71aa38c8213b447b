"""The adjoint for the configurations' own outcomes, and for a batch.

* Runout (the smooth maximum of the horizontal distances from the pile's
  centre) and packing fraction (particle volume over floor area times twice
  the mean centre height) of the resting pile of test_gpu_adjoint.py: their
  gradients with respect to the contact parameters and initial velocities
  agree with central differences of the FP64 oracle within 1e-3.
* A batch of three piles with different contact parameters in one solver:
  the batched adjoint gives each system's parameter and initial-state
  gradients equal to those of the system's own plain solver."""
import numpy as np
import pytest

import pyoracle as po
from fundem_amd import objectives as O
from fundem_amd import workloads as W
from test_gpu_adjoint import PARAMS, _scenario, resting_pile, run_oracle

pytestmark = pytest.mark.gpu

STEPS, EVERY = 60, 20


def outcomes(state, masses, centre):
    vol = masses / 2500.0
    run = O.runout(state, centre, tau=1e-3)
    pf = O.packing_fraction(state, vol, 0.02 * 0.02, 0.0)
    return [run, pf]


def test_runout_and_packing_fraction_gradients_match_finite_differences():
    from fundem_amd import Adjoint, Solver
    tm, x, q = resting_pile(False)
    n = len(x)
    rng = np.random.default_rng(8)
    v0 = rng.normal(0.0, 1e-5, size=(n, 3))
    w0 = rng.normal(0.0, 1e-3, size=(n, 3))
    v0[2] = [0.05, 0.01, 0.0]
    masses = np.array([tm[1][k][0] for k in (np.arange(n) % len(tm[0]))])
    centre = x.mean(axis=0)
    g = Solver()
    W.apply(_scenario(tm, x, q, v0, w0, PARAMS), g)
    adj = Adjoint(g)
    adj.forward(STEPS, EVERY)
    fin = g.state()
    grads = []
    for J, bar in outcomes(fin, masses, centre):
        grads.append(adj.backward(bar["x"], bar["q"], bar["v"], bar["w"]))

    def J_of(p=None, vv=None):
        st, _ = run_oracle(tm, x, q, v0 if vv is None else vv, w0, p or PARAMS, STEPS)
        return np.array([o[0] for o in outcomes(st, masses, centre)])

    rows = []
    for name in ("kn", "mu", "zeta", "wall_mu", "wall_zeta"):
        h = 1e-4 * PARAMS[name]
        pp, pm = dict(PARAMS), dict(PARAMS)
        pp[name] += h
        pm[name] -= h
        fd = (J_of(pp) - J_of(pm)) / (2 * h)
        rows.append((name, PARAMS[name], np.array([gr[name][0] for gr in grads]), fd))
    for (i, c) in ((2, 0), (1, 2)):
        h = 1e-4 * 1e-2
        vp, vm = v0.copy(), v0.copy()
        vp[i, c] += h
        vm[i, c] -= h
        fd = (J_of(vv=vp) - J_of(vv=vm)) / (2 * h)
        rows.append((f"v0[{i},{c}]", 1e-2, np.array([gr["v"][i, c] for gr in grads]), fd))
    for k, label in enumerate(("runout", "packing fraction")):
        top = max(abs(r[1] * r[3][k]) for r in rows)
        assert top > 0, label
        for name, scale, a, fd in rows:
            sa, sf = scale * a[k], scale * fd[k]
            print(f"{label:16s} {name:10s} adjoint {sa: .6e} central {sf: .6e}", flush=True)
            assert abs(sa - sf) <= 1e-3 * max(abs(sf), 1e-4 * top), (label, name, sa, sf)


def test_batched_adjoint_equals_plain_adjoints():
    from fundem_amd import Adjoint, Solver
    tm, x, q = resting_pile(False)
    n = len(x)
    rng = np.random.default_rng(21)
    variants = [dict(PARAMS), dict(PARAMS, kn=1.3e3, ks=5e2, mu=0.35),
                dict(PARAMS, zeta=0.2, wall_mu=0.4, wall_zeta=0.25)]
    scns, starts = [], []
    for p in variants:
        v0 = rng.normal(0.0, 1e-5, size=(n, 3))
        v0[2] = [0.05, 0.01, 0.0]
        w0 = rng.normal(0.0, 1e-3, size=(n, 3))
        scns.append(_scenario(tm, x, q, v0, w0, p))
    masses = np.array([tm[1][k][0] for k in (np.arange(n) % len(tm[0]))])
    # plain runs
    plain = []
    for s in scns:
        g = Solver()
        W.apply(s, g)
        adj = Adjoint(g)
        adj.forward(STEPS, EVERY)
        J, bar = O.kinetic_energy(g.state(), masses)
        plain.append(adj.backward(bar["x"], bar["q"], bar["v"], bar["w"]))
    # one solver, three systems
    b = Solver()
    W.apply_batch(scns, b)
    adj = Adjoint(b)
    adj.forward(STEPS, EVERY)
    st = b.state()
    bar = {k: np.zeros_like(v) for k, v in st.items() if k in ("x", "q", "v", "w")}
    for k in range(3):
        part = {key: st[key][k * n:(k + 1) * n] for key in ("x", "q", "v", "w")}
        _, pb = O.kinetic_energy(part, masses)
        for key in bar:
            bar[key][k * n:(k + 1) * n] = pb[key]
    gb = adj.backward(bar["x"], bar["q"], bar["v"], bar["w"])
    for k in range(3):
        for key in ("x", "q", "v", "w"):
            assert np.array_equal(gb[key][k * n:(k + 1) * n], plain[k][key]), (k, key)
        for key in ("kn", "ks", "mu", "zeta", "wall_mu", "wall_zeta"):
            assert gb[key][k] == plain[k][key][0], (k, key, gb[key][k], plain[k][key][0])
    print("batched adjoint equals the plain adjoints bit for bit", flush=True)
