"""TEST INFRASTRUCTURE ONLY: ctypes binding of the serial FP64 oracle.

Used by tests/, the smoke check and bench.py's CPU baseline as the checker.
The library is built by oracle/Makefile into oracle/_build/liboracle.so.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

SET_PP, SET_PW, SET_GX, SET_GY, SET_GXY = range(5)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    V, I, LL, D, P = None, C.c_int, C.c_longlong, C.c_double, C.c_void_p
    sig = {
        "oracle_grid_set": (V, [_dp, _dp, D, _ip, _dp]),
        "oracle_hash": (V, [I, _dp, _dp, _dp, _dp, _ip, _ip]),
        "oracle_ghost_hash": (V, [I, _dp, _dp, _dp, _dp, _ip, _ip]),
        "oracle_sort_start_end": (V, [I, _ip, _ip, LL, _ip, _ip]),
        "oracle_inclusive_scan": (V, [LL, _ip, _ip]),
        "osim_create": (P, [I, I]),
        "osim_destroy": (V, [P]),
        "osim_add_material": (I, [P, D, D, D, D]),
        "osim_add_wall_material": (I, [P, D, D]),
        "osim_add_material_damping": (I, [P, D, D, D, D]),
        "osim_set_sector": (V, [P, I]),
        "osim_set_global_damping": (V, [P, D]),
        "osim_set_constant_force": (V, [P, I, _dp, _dp]),
        "oracle_mesh_points": (V, [_dp, I, _ip, I, _dp, I, I, _dp, _dp]),
        "oracle_mesh_grid": (LL, [_dp, I, _ip, I, D, I, I, _dp, _ip, _dp]),
        "oracle_sample_level_set": (V, [_dp, _ip, _dp, D, _dp, I, _dp, _dp, _ip]),
        "osim_add_wall_material_damping": (I, [P, D, D]),
        "osim_add_geometry": (I, [P, D, _dp, _dp, D, _ip, _dp, _dp, I]),
        "osim_add_geometry_density": (I, [P, D, _dp, D, _ip, _dp, _dp, I, _dp]),
        "osim_add_wall_geometry": (I, [P, _dp, D, _ip, _dp, _dp, I]),
        "osim_add_particle": (I, [P, _dp, _dp, _dp, _dp, I, I]),
        "osim_add_wall": (I, [P, _dp, _dp, _dp, _dp, I, I]),
        "osim_add_bond": (I, [P, I, I, _dp, _dp, D, D, D, D, D, D, D]),
        "osim_set_periodic": (V, [P, I, I]),
        "osim_set_cell_size": (V, [P, D]),
        "osim_initialize": (I, [P, _dp, _dp, _dp, D]),
        "osim_step": (I, [P, I]),
        "osim_bruteforce_pp": (LL, [P, _ip, _ip, _ip, LL]),
        "osim_num_particles": (I, [P]),
        "osim_num_walls": (I, [P]),
        "osim_num_bonds": (I, [P]),
        "osim_get_state": (V, [P, _dp, _dp, _dp, _dp, _dp, _dp]),
        "osim_set_state": (V, [P, _dp, _dp, _dp, _dp, _dp, _dp]),
        "osim_get_walls": (V, [P, _dp, _dp]),
        "osim_get_geometry": (V, [P, I, _dp, _dp, _dp, _dp, _dp, _dp]),
        "osim_get_grid": (V, [P, I, _dp, _dp, _dp, _ip]),
        "osim_get_contacts": (LL, [P, I, _ip, _ip, _ip, _dp, _dp, _dp, _dp, LL]),
        "osim_get_hash": (V, [P, I, _ip, _ip, _ip, _ip]),
        "osim_get_bonds": (V, [P, _ip, _dp, _dp]),
        "osim_set_spring": (I, [P, I, I, I, I, _dp]),
        "osim_get_ghosts": (I, [P, I, _dp, _ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def d3(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))


def grid_set(mn, mx, cell):
    mn, mx = d3(mn), d3(mx)
    size = np.zeros(3, np.int32)
    inv = np.zeros(3)
    lib().oracle_grid_set(dptr(mn), dptr(mx), float(cell), iptr(size), dptr(inv))
    return mn, mx, size, inv


def hash_(pos, mn, mx, inv, size, ghost=False):
    pos = d3(pos)
    n = pos.size // 3
    out = np.zeros(n, np.int32)
    fn = lib().oracle_ghost_hash if ghost else lib().oracle_hash
    fn(n, dptr(pos), dptr(d3(mn)), dptr(d3(mx)), dptr(d3(inv)), iptr(np.ascontiguousarray(size, np.int32)),
       iptr(out))
    return out


def sort_start_end(hash_values, num_grid):
    h = np.ascontiguousarray(hash_values, dtype=np.int32).copy()
    n = h.size
    idx = np.zeros(n, np.int32)
    st = np.zeros(num_grid, np.int32)
    en = np.zeros(num_grid, np.int32)
    lib().oracle_sort_start_end(n, iptr(h), iptr(idx), int(num_grid), iptr(st), iptr(en))
    return h, idx, st, en


def inclusive_scan(count):
    c = np.ascontiguousarray(count, dtype=np.int32)
    out = np.zeros_like(c)
    lib().oracle_inclusive_scan(c.size, iptr(c), iptr(out))
    return out


class OracleSim:
    """The oracle solver (restates LSSolver's step loop)."""

    def __init__(self, store_fp32: bool, canonical_fold: bool):
        self.h = lib().osim_create(int(store_fp32), int(canonical_fold))

    def __del__(self):
        if getattr(self, "h", None):
            lib().osim_destroy(self.h)
            self.h = None

    def add_material(self, kn, ks, mu, e):
        return lib().osim_add_material(self.h, kn, ks, mu, e)

    def add_wall_material(self, mu, e):
        return lib().osim_add_wall_material(self.h, mu, e)

    def add_material_damping(self, kn, ks, mu, zeta):
        return lib().osim_add_material_damping(self.h, kn, ks, mu, zeta)

    def add_wall_material_damping(self, mu, zeta):
        return lib().osim_add_wall_material_damping(self.h, mu, zeta)

    def add_geometry(self, mass, inertia6, t):
        return lib().osim_add_geometry(self.h, mass, dptr(d3(inertia6)), dptr(d3(t.origin)), t.spacing,
                                       iptr(np.asarray(t.size, np.int32)), dptr(d3(t.sdf)), dptr(d3(t.nodes)),
                                       len(t.nodes))

    def add_geometry_density(self, density, t):
        off = np.zeros(3)
        g = lib().osim_add_geometry_density(self.h, density, dptr(d3(t.origin)), t.spacing,
                                            iptr(np.asarray(t.size, np.int32)), dptr(d3(t.sdf)),
                                            dptr(d3(t.nodes)), len(t.nodes), dptr(off))
        return g, off

    def add_wall_geometry(self, t):
        return lib().osim_add_wall_geometry(self.h, dptr(d3(t.origin)), t.spacing,
                                            iptr(np.asarray(t.size, np.int32)), dptr(d3(t.sdf)),
                                            dptr(d3(t.nodes)), len(t.nodes))

    def add_particle(self, x, v, w, q, mat, geo):
        return lib().osim_add_particle(self.h, dptr(d3(x)), dptr(d3(v)), dptr(d3(w)), dptr(d3(q)), mat, geo)

    def add_particles(self, x, v, w, q, mat, geo):
        x, v, w, q = (np.asarray(a, np.float64) for a in (x, v, w, q))
        n = len(x)
        mat = np.broadcast_to(np.asarray(mat), (n,))
        geo = np.broadcast_to(np.asarray(geo), (n,))
        for i in range(n):
            if self.add_particle(x[i], v[i], w[i], q[i], int(mat[i]), int(geo[i])) < 0:
                raise ValueError(f"oracle rejected particle {i}")

    def add_wall(self, x, v, w, q, mat, geo):
        return lib().osim_add_wall(self.h, dptr(d3(x)), dptr(d3(v)), dptr(d3(w)), dptr(d3(q)), mat, geo)

    def add_bond(self, i, j, point, normal, radius, length, youngs, poisson, tensile=0.0, cohesion=0.0,
                 friction=0.0):
        return lib().osim_add_bond(self.h, i, j, dptr(d3(point)), dptr(d3(normal)), radius, length, youngs,
                                   poisson, tensile, cohesion, friction)

    def set_periodic(self, x, y):
        lib().osim_set_periodic(self.h, int(x), int(y))

    def set_sector(self, on=True):
        lib().osim_set_sector(self.h, int(on))

    def set_global_damping(self, c):
        lib().osim_set_global_damping(self.h, -1.0 if c is None else float(c))

    def set_constant_force(self, index, force, torque):
        f, t = d3(force), d3(torque)
        lib().osim_set_constant_force(self.h, -1 if index is None else int(index), dptr(f), dptr(t))

    def set_cell_size(self, cs):
        lib().osim_set_cell_size(self.h, float(cs))

    def initialize(self, mn, mx, gravity, dt):
        r = lib().osim_initialize(self.h, dptr(d3(mn)), dptr(d3(mx)), dptr(d3(gravity)), float(dt))
        if r != 0:
            raise ValueError("oracle initialize rejected its input")

    def step(self, n=1):
        if lib().osim_step(self.h, int(n)) != 0:
            raise RuntimeError("oracle step failed")

    @property
    def n(self):
        return lib().osim_num_particles(self.h)

    def state(self):
        n = self.n
        x, v, w, f, t = (np.zeros((n, 3)) for _ in range(5))
        q = np.zeros((n, 4))
        lib().osim_get_state(self.h, dptr(x), dptr(v), dptr(w), dptr(q), dptr(f), dptr(t))
        return dict(x=x, v=v, w=w, q=q, F=f, T=t)

    def set_state(self, x=None, v=None, w=None, q=None, F=None, T=None):
        def p(a):
            return None if a is None else dptr(d3(a))
        arrs = [None if a is None else d3(a) for a in (x, v, w, q, F, T)]
        lib().osim_set_state(self.h, *[None if a is None else dptr(a) for a in arrs])

    def walls(self):
        nw = lib().osim_num_walls(self.h)
        x, q = np.zeros((nw, 3)), np.zeros((nw, 4))
        lib().osim_get_walls(self.h, dptr(x), dptr(q))
        return x, q

    def geometry(self, g):
        r, im, ih = (np.zeros(1) for _ in range(3))
        I, Ii, o = np.zeros(6), np.zeros(6), np.zeros(3)
        lib().osim_get_geometry(self.h, g, dptr(r), dptr(im), dptr(I), dptr(Ii), dptr(o), dptr(ih))
        return dict(radius=r[0], inv_mass=im[0], inertia=I, inv_inertia=Ii, origin=o, inv_h=ih[0])

    def grid(self, walls=False):
        mn, mx, inv = np.zeros(3), np.zeros(3), np.zeros(3)
        size = np.zeros(3, np.int32)
        lib().osim_get_grid(self.h, int(walls), dptr(mn), dptr(mx), dptr(inv), iptr(size))
        return dict(min=mn, max=mx, inv=inv, size=size)

    def contacts(self, set_id):
        n = lib().osim_get_contacts(self.h, set_id, None, None, None, None, None, None, None, 0)
        a, b, k = (np.zeros(n, np.int32) for _ in range(3))
        sp, pt, nm = (np.zeros((n, 3)) for _ in range(3))
        ov = np.zeros(n)
        lib().osim_get_contacts(self.h, set_id, iptr(a), iptr(b), iptr(k), dptr(sp), dptr(pt), dptr(nm),
                                dptr(ov), n)
        return dict(A=a, B=b, k=k, spring=sp, point=pt, normal=nm, overlap=ov)

    def bruteforce_pp(self, cap=10_000_000):
        a, b, k = (np.zeros(cap, np.int32) for _ in range(3))
        n = lib().osim_bruteforce_pp(self.h, iptr(a), iptr(b), iptr(k), cap)
        n = min(n, cap)
        return a[:n], b[:n], k[:n]

    def hash_table(self, walls=False):
        n = lib().osim_num_walls(self.h) if walls else self.n
        g = self.grid(walls)
        ng = int(np.prod(g["size"].astype(np.int64)))
        h, idx = np.zeros(n, np.int32), np.zeros(n, np.int32)
        st, en = np.zeros(ng, np.int32), np.zeros(ng, np.int32)
        lib().osim_get_hash(self.h, int(walls), iptr(h), iptr(idx), iptr(st), iptr(en))
        return h, idx, st, en

    def bonds(self):
        nb = lib().osim_num_bonds(self.h)
        act = np.zeros(nb, np.int32)
        s, t = np.zeros(nb), np.zeros(nb)
        lib().osim_get_bonds(self.h, iptr(act), dptr(s), dptr(t))
        return act, s, t

    def set_spring(self, set_id, a, target, k, spring):
        return lib().osim_set_spring(self.h, set_id, a, target, k, dptr(d3(spring)))

    def ghosts(self, set_id):
        """Ghost set 0 (X), 1 (Y) or 2 (XY) of the latest step, per particle:
        positions (n, 3) and ghost cell ids (-1: invisible)."""
        n = lib().osim_get_ghosts(self.h, set_id, None, None)
        pos, h = np.zeros((n, 3)), np.zeros(n, np.int32)
        lib().osim_get_ghosts(self.h, set_id, dptr(pos), iptr(h))
        return pos, h


def mesh_points(vertices, triangles, points, positive_inside=False):
    """Oracle signed distance and gradient of a closed mesh at points (n, 3)."""
    v = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1))
    f = np.ascontiguousarray(np.asarray(triangles, np.int32).reshape(-1))
    p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 3))
    val, grad = np.zeros(len(p)), np.zeros((len(p), 3))
    lib().oracle_mesh_points(v.ctypes.data_as(_dp), len(v) // 3, f.ctypes.data_as(_ip), len(f) // 3,
                             p.ctypes.data_as(_dp), len(p), int(positive_inside),
                             val.ctypes.data_as(_dp), grad.ctypes.data_as(_dp))
    return val, grad


def mesh_grid(vertices, triangles, spacing, padding=2, positive_inside=False):
    """Oracle Base::buildGrid of a closed mesh: (origin, size, values) or None."""
    v = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1))
    f = np.ascontiguousarray(np.asarray(triangles, np.int32).reshape(-1))
    o, sz = np.zeros(3), np.zeros(3, np.int32)
    args = (v.ctypes.data_as(_dp), len(v) // 3, f.ctypes.data_as(_ip), len(f) // 3, float(spacing),
            int(padding), int(positive_inside), o.ctypes.data_as(_dp), sz.ctypes.data_as(_ip))
    n = lib().oracle_mesh_grid(*args, None)
    if n == 0:
        return None
    vals = np.zeros(n)
    lib().oracle_mesh_grid(*args, vals.ctypes.data_as(_dp))
    return o, tuple(int(x) for x in sz), vals


def sample_level_set(sdf, size, origin, spacing, points):
    """Oracle trilinear value, Q1 gradient (per cell) and inside flag at local points (n, 3);
    sdf holds size[0]*size[1]*size[2] node values, x fastest."""
    g = np.ascontiguousarray(np.asarray(sdf, np.float64).reshape(-1))
    sz = np.ascontiguousarray(np.asarray(size, np.int32))
    o = np.ascontiguousarray(np.asarray(origin, np.float64))
    p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 3))
    if g.size != int(np.prod(sz)):
        raise ValueError("sdf size does not match the grid")
    val, grad, ins = np.zeros(len(p)), np.zeros((len(p), 3)), np.zeros(len(p), np.int32)
    lib().oracle_sample_level_set(g.ctypes.data_as(_dp), sz.ctypes.data_as(_ip), o.ctypes.data_as(_dp),
                                  float(spacing), p.ctypes.data_as(_dp), len(p),
                                  val.ctypes.data_as(_dp), grad.ctypes.data_as(_dp),
                                  ins.ctypes.data_as(_ip))
    return val, grad, ins.astype(bool)
