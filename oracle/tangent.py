"""TEST INFRASTRUCTURE ONLY. Forward-mode (dual-number) derivatives of the
FP64 oracle's timestep, for checking the adjoint without finite differences.

The step is restated with numpy on dual numbers: each quantity carries its
value and its derivatives along a fixed set of directions (material
parameters, initial velocities, angular velocities or positions). The
contact path of every step (which node of which body touches which target,
in the oracle's list order) is taken from the FP64 oracle stepped in
lockstep; the tangents follow that path, so they are the derivative of the
oracle's trajectory with the contact set held fixed - what the adjoint
differentiates - and need no finite-difference step. After each step the
values are compared with the oracle's state (they must agree to rounding)
and then replaced by it, so the linearisation point is the oracle's own.

Scope: particles, fixed particles (zero mass), static walls, particle and
wall contacts with the linear spring-dashpot law and Coulomb sliding spring,
gravity; no periodic images, bonds or external-force hooks. The FP32 square
roots of the damping coefficients are differentiated as square roots.

Restated from (reference files):
  half-kick        particleIntegrationKernel.cu:3-32 (rotateTensor myMat.h:107-158)
  drift            particleIntegrationKernel.cu:34-48 (myQua.h:77-93, 51-62)
  contact geometry LSParticleContactDetectionKernel.cu:72-119 and
                   LSParticleContactDetectionKernel.cuh:4-77 (the gradient's x
                   weight is an int there: wx0 = 1, wx1 = 0)
  contact law      contactKernel.cuh:4-77, material mixing contactKernel.cu:79-110,167-200
  force fold       contactKernel.cu:120-123,203-204
"""
from __future__ import annotations

import numpy as np

import pyoracle as po


class D:
    """Values v (shape s) with derivatives t (shape s + (nd,))."""
    __slots__ = ("v", "t")
    __array_priority__ = 1000

    def __init__(self, v, t):
        self.v = np.asarray(v, np.float64)
        self.t = np.asarray(t, np.float64)

    def _w(self):
        return self.v[..., None]

    def __add__(a, b):
        if isinstance(b, D):
            return D(a.v + b.v, a.t + b.t)
        b = np.asarray(b, np.float64)
        return D(a.v + b, np.broadcast_to(a.t, np.broadcast(a.v, b).shape + a.t.shape[-1:]))

    __radd__ = __add__

    def __neg__(a):
        return D(-a.v, -a.t)

    def __sub__(a, b):
        return a + (-b)

    def __rsub__(a, b):
        return (-a) + b

    def __mul__(a, b):
        if isinstance(b, D):
            return D(a.v * b.v, a.t * b._w() + b.t * a._w())
        b = np.asarray(b, np.float64)
        return D(a.v * b, a.t * b[..., None])

    __rmul__ = __mul__

    def recip(a):
        return D(1.0 / a.v, -a.t / (a.v * a.v)[..., None])

    def __truediv__(a, b):
        if isinstance(b, D):
            return a * b.recip()
        return a * (1.0 / np.asarray(b, np.float64))

    def __getitem__(self, idx):
        return D(self.v[idx], self.t[idx])

    def sqrt(a):
        s = np.sqrt(a.v)
        scale = np.where(s > 0.0, 0.5 / np.where(s > 0.0, s, 1.0), 0.0)
        return D(s, a.t * scale[..., None])

    def sqrtf(a):
        """The value of sqrtf((float)x) (correctly rounded FP32), the
        derivative of the square root."""
        r = a.sqrt()
        return D(np.sqrt(a.v.astype(np.float32)).astype(np.float64), r.t)


def const(v, nd):
    v = np.asarray(v, np.float64)
    return D(v, np.zeros(v.shape + (nd,)))


def where(m, a, b):
    m = np.asarray(m, bool)
    return D(np.where(m, a.v, b.v), np.where(m[..., None], a.t, b.t))


def vwhere(m, a, b):
    return [where(m, x, y) for x, y in zip(a, b)]


def vadd(a, b):
    return [x + y for x, y in zip(a, b)]


def vsub(a, b):
    return [x - y for x, y in zip(a, b)]


def vscl(a, s):
    return [x * s for x in a]


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def length(a):
    return dot(a, a).sqrt()


def normalize(a):
    # myVec.h:121-137: zero for |a|^2 <= 1e-20
    l2 = dot(a, a)
    ok = np.abs(l2.v) > 1e-20
    inv = where(ok, D(np.where(ok, l2.v, 1.0), l2.t).sqrt().recip(), const(np.zeros_like(l2.v), l2.t.shape[-1]))
    return [x * inv for x in a]


def rotate(q, v):
    u = q[1:]
    t = vscl(cross(u, v), 2.0)
    return vadd(vadd(v, vscl(t, q[0])), cross(u, t))


def rotate_t(v, q):
    u = [-c for c in q[1:]]
    t = vscl(cross(u, v), 2.0)
    return vadd(vadd(v, vscl(t, q[0])), cross(u, t))


def qnormalize(q):
    l2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]
    ok = np.abs(l2.v) > 1e-20
    inv = D(np.where(ok, l2.v, 1.0), l2.t).sqrt().recip()
    return [where(ok, c * inv, c) for c in q]


def qintegrate(q, w, dt):
    v = vscl(w, 0.5 * dt)
    d0 = -q[1] * v[0] - q[2] * v[1] - q[3] * v[2]
    d1 = q[0] * v[0] + q[3] * v[1] - q[2] * v[2]
    d2 = -q[3] * v[0] + q[0] * v[1] + q[1] * v[2]
    d3 = q[2] * v[0] - q[1] * v[1] + q[0] * v[2]
    return qnormalize([q[0] + d0, q[1] + d1, q[2] + d2, q[3] + d3])


def rotate_tensor(q, I):
    """R I R^T for a symmetric I given as columns (xx, yy, zz, xy, xz, yz)."""
    q0, q1, q2, q3 = q
    q11, q22, q33 = 2.0 * q1 * q1, 2.0 * q2 * q2, 2.0 * q3 * q3
    a, b, c = 1.0 - q22 - q33, 2.0 * (q1 * q2 - q0 * q3), 2.0 * (q1 * q3 + q0 * q2)
    d, e, f = 2.0 * (q1 * q2 + q0 * q3), 1.0 - q11 - q33, 2.0 * (q2 * q3 - q0 * q1)
    g, h, i = 2.0 * (q1 * q3 - q0 * q2), 2.0 * (q2 * q3 + q0 * q1), 1.0 - q11 - q22
    xx, yy, zz, xy, xz, yz = (I[:, k] for k in range(6))
    r0 = [a * xx + b * xy + c * xz, a * xy + b * yy + c * yz, a * xz + b * yz + c * zz]
    r1 = [d * xx + e * xy + f * xz, d * xy + e * yy + f * yz, d * xz + e * yz + f * zz]
    r2 = [g * xx + h * xy + i * xz, g * xy + h * yy + i * yz, g * xz + h * yz + i * zz]
    return [r0[0] * a + r0[1] * b + r0[2] * c, r1[0] * d + r1[1] * e + r1[2] * f,
            r2[0] * g + r2[1] * h + r2[2] * i, r0[0] * d + r0[1] * e + r0[2] * f,
            r0[0] * g + r0[1] * h + r0[2] * i, r1[0] * g + r1[1] * h + r1[2] * i]


def sym_mul(m, v):
    xx, yy, zz, xy, xz, yz = m
    return [xx * v[0] + xy * v[1] + xz * v[2], xy * v[0] + yy * v[1] + yz * v[2],
            xz * v[0] + yz * v[1] + zz * v[2]]


def restitution_of_damping(zeta):
    if zeta <= 0.0:
        return 1.0
    if zeta >= 1.0:
        return 0.0
    return float(np.exp(-np.pi * zeta / np.sqrt(1.0 - zeta * zeta)))


def dissipation_of(e):
    le = float(np.log(e))
    return -le / float(np.sqrt(np.float32(le * le + np.pi * np.pi)))


class _Recorder:
    """Feeds a scenario to the oracle (workloads.apply) and keeps what the
    restatement needs: materials, the templates as the oracle stores them
    (nodes and SDF origin shifted by the centroid of a density template) and
    the walls."""

    def __init__(self, o):
        self.o = o
        self.pmat, self.wmat, self.ptm, self.wtm, self.walls = [], [], [], [], []

    # materials as (kn, ks, mu, damping ratio, restitution); a material given
    # by restitution e has the ratio -ln e / sqrtf(ln^2 e + pi^2)
    # (contactKernel.cuh:36-44, FP32 square root)
    def add_material_damping(self, kn, ks, mu, zeta):
        z = max(zeta, 0.0)
        self.pmat.append((kn, ks, mu, z, restitution_of_damping(z)))
        return self.o.add_material_damping(kn, ks, mu, zeta)

    def add_wall_material_damping(self, mu, zeta):
        z = max(zeta, 0.0)
        self.wmat.append((mu, z, restitution_of_damping(z)))
        return self.o.add_wall_material_damping(mu, zeta)

    def add_material(self, kn, ks, mu, e):
        e = e if 0.0 < e < 1.0 else 1.0
        self.pmat.append((kn, ks, mu, dissipation_of(e), e))
        return self.o.add_material(kn, ks, mu, e)

    def add_wall_material(self, mu, e):
        e = e if 0.0 < e < 1.0 else 1.0
        self.wmat.append((mu, dissipation_of(e), e))
        return self.o.add_wall_material(mu, e)

    def add_geometry(self, mass, inertia6, t):
        self.ptm.append((t, np.zeros(3)))
        return self.o.add_geometry(mass, inertia6, t)

    def add_geometry_density(self, density, t):
        g, off = self.o.add_geometry_density(density, t)
        self.ptm.append((t, np.asarray(off, np.float64).copy()))
        return g, off

    def add_wall_geometry(self, t):
        self.wtm.append(t)
        return self.o.add_wall_geometry(t)

    def add_particles(self, *a):
        return self.o.add_particles(*a)

    def add_wall(self, x, v, w, q, mat, geo):
        if np.any(np.asarray(v) != 0.0) or np.any(np.asarray(w) != 0.0):
            raise NotImplementedError("moving walls")
        self.walls.append((np.asarray(x, np.float64), np.asarray(q, np.float64), int(mat), int(geo)))
        return self.o.add_wall(x, v, w, q, mat, geo)

    def add_bond(self, *a):
        raise NotImplementedError("bonds")

    def set_periodic(self, *a):
        raise NotImplementedError("periodic images")

    def set_sector(self, *a):
        raise NotImplementedError("sector images")

    def set_cell_size(self, cs):
        return self.o.set_cell_size(cs)

    def initialize(self, *a):
        return self.o.initialize(*a)


class _Grids:
    """SDF grids of a template set, concatenated, with per-template origin,
    inverse spacing, size and offset."""

    def __init__(self, entries):
        self.origin = np.array([e[0] for e in entries], np.float64)
        self.ih = np.array([e[1] for e in entries], np.float64)
        self.size = np.array([e[2] for e in entries], np.int64)
        offs, data = [], []
        n = 0
        for e in entries:
            offs.append(n)
            data.append(np.asarray(e[3], np.float64))
            n += len(e[3])
        self.off = np.array(offs, np.int64)
        self.sdf = np.concatenate(data) if data else np.zeros(0)


class Tangent:
    """Forward-mode derivatives of the oracle's trajectory of scenario `scn`
    along `directions`, each one of
      ("kn" | "ks" | "mu" | "zeta", material), ("wall_mu" | "wall_zeta", wall material),
    where zeta is the damping ratio the contact law uses (the dissipation),
      ("x" | "v" | "w", particle, component).
    With store_fp32 the oracle stores its state in FP32 as the FP32 build
    does, and the tangents are taken along that trajectory (the roundings
    count as the identity, as they do for the adjoint); the restated values
    then differ from the oracle's by those roundings."""

    def __init__(self, scn, directions, store_fp32=False):
        self.o = po.OracleSim(bool(store_fp32), True)
        rec = _Recorder(self.o)
        from fundem_amd import workloads as W
        W.apply(scn, rec)
        self.dirs = list(directions)
        nd = self.nd = len(self.dirs)
        self.dt = float(scn.dt)
        self.g = np.asarray(scn.gravity, np.float64)
        st = self.o.state()
        if np.any(st["F"] != 0.0) or np.any(st["T"] != 0.0):
            raise NotImplementedError("a state that starts with forces")
        n = self.n = len(st["x"])
        self.geo = np.asarray(scn.geo, np.int64)
        self.mat = np.asarray(scn.mat, np.int64)
        # particle templates as the oracle holds them
        infos = [self.o.geometry(k) for k in range(len(rec.ptm))]
        self.inv_mass = np.array([i["inv_mass"] for i in infos])[self.geo]
        self.I = np.array([i["inertia"] for i in infos])[self.geo]
        self.Iinv = np.array([i["inv_inertia"] for i in infos])[self.geo]
        self.f32 = bool(store_fp32)
        self.nodes = [self._r(np.asarray(t.nodes, np.float64) - off) for t, off in rec.ptm]
        self.pgrid = _Grids([(np.asarray(t.origin, np.float64) - off, 1.0 / t.spacing, tuple(t.size), t.sdf)
                             for t, off in rec.ptm])
        self.wgrid = _Grids([(np.asarray(t.origin, np.float64), 1.0 / t.spacing, tuple(t.size), t.sdf)
                             for t in rec.wtm])
        self.wx = np.array([w[0] for w in rec.walls]).reshape(-1, 3)
        self.wq = np.array([w[1] / np.linalg.norm(w[1]) for w in rec.walls]).reshape(-1, 4)
        self.wmat_of = np.array([w[2] for w in rec.walls], np.int64)
        self.wgeo = np.array([w[3] for w in rec.walls], np.int64)
        # material parameters with their seeds
        def seeded(vals, names):
            out = {}
            for name, col in names:
                v = np.array([m[col] for m in vals], np.float64)
                t = np.zeros(v.shape + (nd,))
                for d, spec in enumerate(self.dirs):
                    if spec[0] == name:
                        t[spec[1], d] = 1.0
                out[name] = D(v, t)
            return out
        pm = seeded(rec.pmat, (("kn", 0), ("ks", 1), ("mu", 2), ("zeta", 3)))
        wm = seeded(rec.wmat, (("wall_mu", 0), ("wall_zeta", 1)))
        self.kn = D(np.where(pm["kn"].v > 0, pm["kn"].v, 0.0), pm["kn"].t)
        self.ks = D(np.where(pm["ks"].v > 0, pm["ks"].v, 0.0), pm["ks"].t)
        self.mu = D(np.where(pm["mu"].v > 0, pm["mu"].v, 0.0), pm["mu"].t)
        self.diss = D(np.where(pm["zeta"].v > 0, pm["zeta"].v, 0.0), pm["zeta"].t)
        self.e = np.array([m[4] for m in rec.pmat])
        self.wmu = D(np.where(wm["wall_mu"].v > 0, wm["wall_mu"].v, 0.0), wm["wall_mu"].t) if rec.wmat else None
        self.wdiss = D(np.where(wm["wall_zeta"].v > 0, wm["wall_zeta"].v, 0.0), wm["wall_zeta"].t) if rec.wmat else None
        self.we = np.array([m[2] for m in rec.wmat])
        # state with its seeds
        def state_of(key, k):
            comps = []
            for c in range(k):
                t = np.zeros((n, nd))
                for d, spec in enumerate(self.dirs):
                    if spec[0] == key and spec[2] == c:
                        t[spec[1], d] = 1.0
                comps.append(D(st[key][:, c].copy(), t))
            return comps
        self.x, self.v, self.w = state_of("x", 3), state_of("v", 3), state_of("w", 3)
        self.q = [const(st["q"][:, c].copy(), nd) for c in range(4)]
        self.F = [const(np.zeros(n), nd) for _ in range(3)]
        self.T = [const(np.zeros(n), nd) for _ in range(3)]
        self.springs = [{}, {}]  # (A, node, target) -> (value (3,), tangent (3, nd)) per set
        # largest difference of the restated values from the oracle's, relative
        # to the largest magnitude of the quantity, per quantity
        self.mismatch = {}
        self.max_mismatch = 0.0

    # values stored at state precision (FP32 with store_fp32), tangents kept
    def _r(self, a):
        return np.asarray(a, np.float32).astype(np.float64) if self.f32 else a

    def _rv(self, comps):
        return [D(self._r(c.v), c.t) for c in comps] if self.f32 else comps

    # particleVelocityAngularVelocityIntegrationKernel
    def _kick(self, h):
        moving = np.abs(self.inv_mass) > 1e-20
        v = [self.v[c] + (self.F[c] * self.inv_mass + self.g[c]) * h for c in range(3)]
        Iw = sym_mul(rotate_tensor(self.q, self.I), self.w)
        dw = sym_mul(rotate_tensor(self.q, self.Iinv), vsub(self.T, cross(self.w, Iw)))
        w = [self.w[c] + dw[c] * h for c in range(3)]
        self.v = self._rv(vwhere(moving, v, self.v))
        self.w = self._rv(vwhere(moving, w, self.w))

    def _drift(self):
        self.x = vadd(self.x, vscl(self.v, self.dt))
        self.q = self._rv(qintegrate(self.q, self.w, self.dt))

    def _sample(self, grids, tg, local):
        """Overlap and local gradient at `local` in template tg's SDF."""
        g = [(local[c] - grids.origin[tg, c]) * grids.ih[tg] for c in range(3)]
        i0 = [np.floor(gc.v).astype(np.int64) for gc in g]
        f = [g[c] - i0[c] for c in range(3)]
        sx, sy = grids.size[tg, 0], grids.size[tg, 0] * grids.size[tg, 1]
        base = grids.off[tg] + i0[2] * sy + i0[1] * sx + i0[0]
        c = [grids.sdf[base + o] for o in (0, 1, sx, sx + 1, sy, sy + 1, sy + sx, sy + sx + 1)]
        x, y, z = f
        wx0, wx1, wy0, wy1, wz0, wz1 = 1.0 - x, x, 1.0 - y, y, 1.0 - z, z
        phi = wx0 * c[0] * wy0 * wz0 + wx1 * c[1] * wy0 * wz0 + wx0 * c[2] * wy1 * wz0 + \
            wx1 * c[3] * wy1 * wz0 + wx0 * c[4] * wy0 * wz1 + wx1 * c[5] * wy0 * wz1 + \
            wx0 * c[6] * wy1 * wz1 + wx1 * c[7] * wy1 * wz1
        # the gradient's x weights are wx0 = 1, wx1 = 0
        dx = wy0 * (c[1] - c[0]) * wz0 + wy1 * (c[3] - c[2]) * wz0 + wy0 * (c[5] - c[4]) * wz1 + \
            wy1 * (c[7] - c[6]) * wz1
        dy = wz0 * (c[2] - c[0]) + wz1 * (c[6] - c[4])
        dz = wy0 * (c[4] - c[0]) + wy1 * (c[6] - c[2])
        return -phi, [dx, dy, dz]

    def _forces(self, contacts):
        nd = self.nd
        F = [const(np.zeros(self.n), nd) for _ in range(3)]
        T = [const(np.zeros(self.n), nd) for _ in range(3)]
        new_springs = [{}, {}]
        for si, (set_id, c) in enumerate(contacts):
            m = c["A"].size
            if m == 0:
                continue
            A, B, K = c["A"].astype(np.int64), c["B"].astype(np.int64), c["k"].astype(np.int64)
            xA, qA = [x[A] for x in self.x], [q[A] for q in self.q]
            s = np.stack([self.nodes[self.geo[a]][k] for a, k in zip(A, K)])
            p = vadd(rotate(qA, [s[:, 0], s[:, 1], s[:, 2]]), xA)
            pp = set_id == po.SET_PP
            if pp:
                xB, qB = [x[B] for x in self.x], [q[B] for q in self.q]
                vB, wB = [v[B] for v in self.v], [w[B] for w in self.w]
                grids, tg = self.pgrid, self.geo[B]
            else:
                xB = [const(self.wx[B, k], nd) for k in range(3)]
                qB = [const(self.wq[B, k], nd) for k in range(4)]
                vB = wB = [const(np.zeros(m), nd) for _ in range(3)]
                grids, tg = self.wgrid, self.wgeo[B]
            ov, grad = self._sample(grids, tg, rotate_t(vsub(p, xB), qB))
            n = normalize(rotate(qB, grad))
            pc = vadd(p, vscl(n, ov * 0.5))
            ovs = max(float(np.max(np.abs(c["overlap"]))), 1e-300)
            self.mismatch["overlap"] = max(self.mismatch.get("overlap", 0.0),
                                           float(np.max(np.abs(ov.v - c["overlap"]))) / ovs)
            pcs = max(float(np.max(np.abs(c["point"]))), 1e-300)
            self.mismatch["point"] = max(self.mismatch.get("point", 0.0),
                                         float(np.max(np.abs(np.stack([z.v for z in pc], 1) - c["point"]))) / pcs)
            # materials
            mi = self.mat[A]
            invMi = self.inv_mass[A]
            kni, ksi, mui, di = self.kn[mi], self.ks[mi], self.mu[mi], self.diss[mi]
            if pp:
                mj = self.mat[B]
                invMj = self.inv_mass[B]
                knj, ksj, muj, dj = self.kn[mj], self.ks[mj], self.mu[mj], self.diss[mj]
                kn = where((kni.v > 0) & (knj.v > 0), kni * knj / where(kni.v + knj.v != 0, kni + knj, const(np.ones(m), nd)),
                           const(np.zeros(m), nd))
                ks = where((ksi.v > 0) & (ksj.v > 0), ksi * ksj / where(ksi.v + ksj.v != 0, ksi + ksj, const(np.ones(m), nd)),
                           const(np.zeros(m), nd))
                tot = invMi + invMj
                effM = np.where((invMi > 0) | (invMj > 0), 1.0 / np.where(tot != 0, tot, 1.0), 0.0)
                diss = where(self.e[mi] <= self.e[mj], di, dj)
                mu = where(np.fmin(mui.v, muj.v) == mui.v, mui, muj)
            else:
                wmi = self.wmat_of[B]
                effM = np.where(invMi > 0, 1.0 / np.where(invMi > 0, invMi, 1.0), 0.0)
                kn, ks = kni, ksi
                muw, dw = self.wmu[wmi], self.wdiss[wmi]
                mu = where(np.fmin(mui.v, muw.v) == mui.v, mui, muw)
                diss = where(self.e[mi] <= self.we[wmi], di, dw)
            vA, wA = [v[A] for v in self.v], [w[A] for w in self.w]
            vc = vsub(vadd(vA, cross(wA, vsub(pc, xA))), vadd(vB, cross(wB, vsub(pc, xB))))
            prev_v = np.zeros((m, 3))
            prev_t = np.zeros((m, 3, nd))
            old = self.springs[si]
            for r, key in enumerate(zip(A.tolist(), K.tolist(), B.tolist())):
                if key in old:
                    prev_v[r], prev_t[r] = old[key]
            prev = [D(prev_v[:, k], prev_t[:, k]) for k in range(3)]
            Fc, spring = self._law(vc, n, ov, kn, ks, mu, diss, effM, prev)
            for r, key in enumerate(zip(A.tolist(), K.tolist(), B.tolist())):
                new_springs[si][key] = (self._r(np.array([spring[k].v[r] for k in range(3)])),
                                        np.array([spring[k].t[r] for k in range(3)]))
            TA = cross(vsub(pc, xA), Fc)
            for k in range(3):
                np.add.at(F[k].v, A, Fc[k].v)
                np.add.at(F[k].t, A, Fc[k].t)
                np.add.at(T[k].v, A, TA[k].v)
                np.add.at(T[k].t, A, TA[k].t)
            if pp:
                TB = cross(vsub(pc, xB), [-f for f in Fc])
                for k in range(3):
                    np.add.at(F[k].v, B, -Fc[k].v)
                    np.add.at(F[k].t, B, -Fc[k].t)
                    np.add.at(T[k].v, B, TB[k].v)
                    np.add.at(T[k].t, B, TB[k].t)
        self.springs = new_springs
        self.F, self.T = self._rv(F), self._rv(T)

    def _law(self, v, n, delta, kn, ks, mu, diss, effM, prev):
        """linearContact with integrateSlidingSpring (contactKernel.cuh:4-77)."""
        nd = self.nd
        m = delta.v.shape[0]
        zero = const(np.zeros(m), nd)
        zero3 = [zero, zero, zero]
        cn = diss * 2.0 * (kn * effM).sqrtf()
        vn = vscl(n, dot(v, n))
        Fn = vsub(vscl(n, kn * delta), vscl(vn, cn))
        ct = diss * 2.0 * (ks * effM).sqrtf()
        vt = vsub(v, vn)
        # sliding spring
        p1 = vsub(prev, vscl(n, dot(prev, n)))
        a1 = length(p1)
        nz = a1.v > 1e-20
        a1s = where(nz, a1, const(np.ones(m), nd))
        p1 = vwhere(nz, vscl(p1, length(prev) / a1s), p1)
        spring = vadd(p1, vscl(vt, self.dt))
        sf = vsub(vscl(spring, -ks), vscl(vt, ct))
        asf, afn = length(sf), length(Fn)
        slip = (asf.v > 1e-20) & (asf.v > mu.v * afn.v)
        asfs = where(asf.v > 1e-20, asf, const(np.ones(m), nd))
        ratio = mu * afn / asfs
        sf2 = vscl(sf, ratio)
        kss = where(ks.v != 0, ks, const(np.ones(m), nd))
        spring2 = vscl(vadd(sf2, vscl(vt, ct)), -kss.recip())
        spring = vwhere(slip, spring2, spring)
        spring = vwhere((mu.v > 0) & (ks.v > 0), spring, zero3)
        Fs = vsub(vscl(spring, -ks), vscl(vt, ct))
        F = vadd(Fn, Fs)
        on = delta.v > 0
        return vwhere(on, F, zero3), vwhere(on, spring, zero3)

    def step(self, steps=1):
        for _ in range(steps):
            h = 0.5 * self.dt
            self._kick(h)
            self._drift()
            self.o.step(1)
            contacts = [(po.SET_PP, self.o.contacts(po.SET_PP)), (po.SET_PW, self.o.contacts(po.SET_PW))]
            self._forces(contacts)
            self._kick(h)
            self._resync(contacts)

    def _resync(self, contacts):
        st = self.o.state()
        for key, k in (("x", 3), ("v", 3), ("w", 3), ("q", 4), ("F", 3), ("T", 3)):
            comps = getattr(self, key)
            scale = max(float(np.max(np.abs(st[key]))), 1e-300)
            for c in range(k):
                err = float(np.max(np.abs(comps[c].v - st[key][:, c]))) / scale
                self.mismatch[key] = max(self.mismatch.get(key, 0.0), err)
                comps[c] = D(st[key][:, c].copy(), comps[c].t)
        self.max_mismatch = max(self.mismatch.values())
        for si, (_, c) in enumerate(contacts):
            sp = self.springs[si]
            for r, key in enumerate(zip(c["A"].tolist(), c["k"].tolist(), c["B"].tolist())):
                val, tan = sp[key]
                sp[key] = (c["spring"][r].copy(), tan)

    def state(self):
        """Values (the oracle's) and tangents (n, k, nd) of x, q, v, w."""
        out = {}
        for key in ("x", "q", "v", "w"):
            comps = getattr(self, key)
            out[key] = np.stack([c.v for c in comps], 1)
            out["d" + key] = np.stack([c.t for c in comps], 1)
        return out

    def directional(self, bar):
        """dJ along each direction, from the adjoint of J with respect to the
        final x, q, v, w (the objectives' cotangents)."""
        st = self.state()
        return sum(np.einsum("ik,ikd->d", np.asarray(bar[k], np.float64), st["d" + k])
                   for k in ("x", "q", "v", "w"))
