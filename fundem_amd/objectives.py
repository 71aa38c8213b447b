"""Scalar outcomes of a final state and their adjoints (the cotangents
Adjoint.backward takes), for the differentiated configurations:

* kinetic_energy: J = sum_i m_i |v_i|^2 / 2 over the moving particles.
* runout: how far the moving particles spread horizontally from a centre,
  r_i = |(x_i - c)_xy|. With tau > 0 the smooth maximum
  J = tau log sum_i exp(r_i / tau) (within tau log N of the largest r_i and
  differentiable everywhere); with tau = 0 the largest r_i itself with the
  subgradient of the particle that attains it (ties to the lowest index).
* packing_fraction: J = (sum_i V_i) / (A H) with the bed height taken as
  twice the mean height of the particles' centres above the floor,
  H = 2 (mean_i z_i - z_floor) (exact for a bed of uniform density, smooth in
  every position).

Each returns (J, dict(x=..., q=..., v=..., w=...)) with the adjoint of J
with respect to the final positions, orientations, velocities and angular
velocities (zero for fixed particles)."""
from __future__ import annotations

import numpy as np


def _zeros(state):
    return dict(x=np.zeros_like(state["x"]), q=np.zeros_like(state["q"]),
                v=np.zeros_like(state["v"]), w=np.zeros_like(state["w"]))


def _mask(state, moving):
    n = len(state["x"])
    return np.ones(n, bool) if moving is None else np.asarray(moving, bool)


def kinetic_energy(state, masses, moving=None):
    m = np.where(_mask(state, moving), np.asarray(masses, np.float64), 0.0)
    v = state["v"]
    J = 0.5 * float(np.sum(m[:, None] * v * v))
    bar = _zeros(state)
    bar["v"] = m[:, None] * v
    return J, bar


def runout(state, centre, moving=None, tau=0.0):
    sel = _mask(state, moving)
    d = state["x"][:, :2] - np.asarray(centre, np.float64)[:2]
    r = np.sqrt(np.sum(d * d, axis=1))
    bar = _zeros(state)
    idx = np.flatnonzero(sel)
    if idx.size == 0:
        return 0.0, bar
    dr = np.zeros_like(d)
    ok = r > 0.0
    dr[ok] = d[ok] / r[ok, None]
    if tau > 0.0:
        z = r[idx] / tau
        zmax = z.max()
        e = np.exp(z - zmax)
        J = tau * (zmax + np.log(e.sum()))
        wts = e / e.sum()
        bar["x"][idx, :2] = wts[:, None] * dr[idx]
        return float(J), bar
    k = idx[int(np.argmax(r[idx]))]
    bar["x"][k, :2] = dr[k]
    return float(r[k]), bar


def packing_fraction(state, volumes, area, floor_z=0.0, moving=None):
    sel = _mask(state, moving)
    idx = np.flatnonzero(sel)
    bar = _zeros(state)
    if idx.size == 0:
        return 0.0, bar
    vol = float(np.sum(np.asarray(volumes, np.float64)[idx]))
    hmean = float(np.mean(state["x"][idx, 2])) - floor_z
    H = 2.0 * hmean
    J = vol / (area * H)
    # dJ/dz_i = -vol / (A H^2) * dH/dz_i, dH/dz_i = 2 / N
    bar["x"][idx, 2] = -vol / (area * H * H) * (2.0 / idx.size)
    return J, bar
