"""Input generators for level-set particles, walls and packings.

These prepare inputs: they are the counterpart of FunDEM's host-side
helpers (kernel/helper/LevelSetObjectGenerator.h), written for this library.
Every generator is seeded and returns float32-representable data, so the
same arrays can be given to the device path and to the FP64 oracle.

Geometry conventions follow the solver API (LSSolver.h:61-84):
  * SDF values on an (nx, ny, nz) grid, x fastest, at origin + h*(i, j, k);
  * surface nodes in the template frame.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

CHUNK = 64  # nodes per culling chunk (one wavefront)


@dataclass
class Template:
    """One level-set particle template."""

    origin: np.ndarray  # (3,) float64
    spacing: float
    size: tuple  # (nx, ny, nz)
    sdf: np.ndarray  # (nx*ny*nz,) float32, x fastest
    nodes: np.ndarray  # (M, 3) float32
    tris: np.ndarray | None = None  # (T, 3) int32 node ids of the surface triangles, outward


def _sphere_triangles(dirs: np.ndarray, order: np.ndarray) -> np.ndarray:
    """Triangles of unit directions (their convex hull), wound outward, with
    node ids renumbered for the node order `order` (new position of old node
    i is where i appears in `order`)."""
    from scipy.spatial import ConvexHull

    tri = ConvexHull(dirs).simplices.astype(np.int64)
    a, b, c = dirs[tri[:, 0]], dirs[tri[:, 1]], dirs[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0.0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    return inv[tri].astype(np.int32)


def _safe_pow(base: np.ndarray, exponent: float) -> np.ndarray:
    base = np.abs(base)
    out = np.zeros_like(base)
    nz = base > 0.0
    out[nz] = np.power(base[nz], exponent)
    return out


def superellipsoid_value_grad(p: np.ndarray, a: float, b: float, c: float, ee: float, en: float):
    """Barr superellipsoid implicit function and gradient.

    f = (|x/a|^(2/ee) + |y/b|^(2/ee))^(ee/en) + |z/c|^(2/en) - 1, the form
    used by LevelSetObjectGenerator.h:988-1010 (powers 2/ee and 2/en).
    """
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    X, Y, Z = np.abs(x / a), np.abs(y / b), np.abs(z / c)
    exy, ez, eo = 2.0 / ee, 2.0 / en, ee / en
    u = _safe_pow(X, exy) + _safe_pow(Y, exy)
    f = _safe_pow(u, eo) + _safe_pow(Z, ez) - 1.0
    # d/du u^eo = eo u^(eo-1); guard u = 0 where the chain factor vanishes
    du = np.zeros_like(u)
    nz = u > 0.0
    du[nz] = eo * np.power(u[nz], eo - 1.0)
    gx = du * exy * _safe_pow(X, exy - 1.0) * np.sign(x) / a
    gy = du * exy * _safe_pow(Y, exy - 1.0) * np.sign(y) / b
    gz = ez * _safe_pow(Z, ez - 1.0) * np.sign(z) / c
    gx = np.where(X > 0.0, gx, 0.0)
    gy = np.where(Y > 0.0, gy, 0.0)
    gz = np.where(Z > 0.0, gz, 0.0)
    return f, np.stack([gx, gy, gz], axis=-1)


def _radial_surface_points(dirs: np.ndarray, a, b, c, ee, en) -> np.ndarray:
    """Scale unit directions onto the superellipsoid surface (f(t*d) = 0)."""
    exy, ez, eo = 2.0 / ee, 2.0 / en, ee / en
    m = _safe_pow(_safe_pow(dirs[:, 0] / a, exy) + _safe_pow(dirs[:, 1] / b, exy), eo)
    m = m + _safe_pow(dirs[:, 2] / c, ez)
    t = np.power(m, -0.5 * en)
    return dirs * t[:, None]


def fibonacci_directions(m: int) -> np.ndarray:
    i = np.arange(m, dtype=np.float64)
    z = 1.0 - 2.0 * (i + 0.5) / m
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * (math.pi * (3.0 - math.sqrt(5.0)))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def chunk_order(points: np.ndarray, chunk: int = CHUNK) -> np.ndarray:
    """Permutation grouping points into spatially compact runs of `chunk`.

    Recursive bisection along the longest extent, with every split placed on
    a multiple of `chunk`, so consecutive runs of `chunk` nodes are compact
    patches. The device sweep culls whole runs, so compact runs cull well.
    """
    idx = np.arange(len(points))

    def rec(ids: np.ndarray) -> list:
        if len(ids) <= chunk:
            return [ids]
        pts = points[ids]
        axis = int(np.argmax(pts.max(axis=0) - pts.min(axis=0)))
        ids = ids[np.argsort(pts[:, axis], kind="stable")]
        nchunks = (len(ids) + chunk - 1) // chunk
        cut = (nchunks // 2) * chunk
        return rec(ids[:cut]) + rec(ids[cut:])

    return np.concatenate(rec(idx))


def superellipsoid_template(a: float, b: float, c: float, p1: float, p2: float,
                            n_grid: int = 32, n_nodes: int = 1000, pad: int = 2) -> Template:
    """Superellipsoid with semi-axes a, b, c and powers p1 (xy), p2 (z).

    The SDF is sampled as f/|grad f| (the approximation of
    LevelSetObjectGenerator.h:182-204) on an n_grid^3 grid whose spacing
    leaves `pad` layers outside the bounding box; values are stored float32.
    Nodes: Fibonacci directions projected radially onto the surface, ordered
    into compact 64-node chunks.
    """
    ee, en = 2.0 / p1, 2.0 / p2
    extent = 2.0 * max(a, b, c)
    h = extent / (n_grid - 1 - 2 * pad)
    o = -0.5 * h * (n_grid - 1)
    origin = np.array([o, o, o], dtype=np.float64)
    ax = o + h * np.arange(n_grid, dtype=np.float64)
    zz, yy, xx = np.meshgrid(ax, ax, ax, indexing="ij")  # x fastest after ravel
    pts = np.stack([xx, yy, zz], axis=-1).reshape(-1, 3)
    f, g = superellipsoid_value_grad(pts, a, b, c, ee, en)
    gn = np.linalg.norm(g, axis=1)
    phi = np.where(gn > 1e-12, f / np.where(gn > 1e-12, gn, 1.0), np.where(f <= 0.0, -0.5 * h, 0.5 * h))
    dirs = fibonacci_directions(n_nodes)
    nodes = _radial_surface_points(dirs, a, b, c, ee, en)
    order = chunk_order(nodes)
    return Template(origin=origin, spacing=float(h), size=(n_grid, n_grid, n_grid),
                    sdf=phi.astype(np.float32), nodes=nodes[order].astype(np.float32),
                    tris=_sphere_triangles(dirs, order))


def sphere_template(r: float, n_grid: int = 16, n_nodes: int = 400, pad: int = 2) -> Template:
    """Sphere of radius r with the exact SDF |p| - r (float32)."""
    h = 2.0 * r / (n_grid - 1 - 2 * pad)
    o = -0.5 * h * (n_grid - 1)
    ax = o + h * np.arange(n_grid, dtype=np.float64)
    zz, yy, xx = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([xx, yy, zz], axis=-1).reshape(-1, 3)
    phi = np.linalg.norm(pts, axis=1) - r
    dirs = fibonacci_directions(n_nodes)
    nodes = dirs * r
    order = chunk_order(nodes)
    return Template(origin=np.array([o, o, o]), spacing=float(h), size=(n_grid,) * 3,
                    sdf=phi.astype(np.float32), nodes=nodes[order].astype(np.float32),
                    tris=_sphere_triangles(dirs, order))


def plane_wall(xy_min, xy_max) -> Template:
    """Plane z = 0 (outward normal +z) as a 2x2x2 affine level set.

    Same construction as the reference's plane in
    apps/test/LSSphereBouncing.cpp:86-120: spacing = the larger in-plane
    extent, grid centred on the wall, phi = z, four corner nodes.
    """
    xy_min = np.asarray(xy_min, dtype=np.float64)
    xy_max = np.asarray(xy_max, dtype=np.float64)
    s = float(max(xy_max[0] - xy_min[0], xy_max[1] - xy_min[1]))
    centre = np.array([0.5 * (xy_min[0] + xy_max[0]), 0.5 * (xy_min[1] + xy_max[1]), 0.0])
    origin = centre - 0.5 * s
    sdf = np.zeros(8, dtype=np.float64)
    for k in range(2):
        for j in range(2):
            for i in range(2):
                sdf[i + 2 * (j + 2 * k)] = origin[2] + k * s
    nodes = np.array([[xy_min[0], xy_min[1], 0.0], [xy_max[0], xy_min[1], 0.0],
                      [xy_max[0], xy_max[1], 0.0], [xy_min[0], xy_max[1], 0.0]])
    return Template(origin=origin, spacing=s, size=(2, 2, 2), sdf=sdf.astype(np.float32),
                    nodes=nodes.astype(np.float32),
                    tris=np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32))


def shoemake_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (q0, q1, q2, q3), Shoemake's method
    (as apps/tutorial/LSParticleRolling.cpp:14-31)."""
    u = rng.random((n, 3))
    s1, s2 = np.sqrt(1.0 - u[:, 0]), np.sqrt(u[:, 0])
    a, b = 2.0 * math.pi * u[:, 1], 2.0 * math.pi * u[:, 2]
    return np.stack([s2 * np.cos(b), s1 * np.sin(a), s1 * np.cos(a), s2 * np.sin(b)], axis=1)


def jittered_lattice(counts, spacing: float, origin, jitter: float, rng: np.random.Generator):
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([i, j, k], axis=-1).reshape(-1, 3).astype(np.float64) * spacing
    base += np.asarray(origin, dtype=np.float64)
    return base + rng.uniform(-jitter, jitter, size=base.shape)


def template_radius(t: Template) -> float:
    return float(np.max(np.linalg.norm(t.nodes.astype(np.float64), axis=1)))


def template_box_reach(t: Template) -> float:
    """Largest distance from the template origin (0,0,0) to a corner of its SDF box."""
    lo = t.origin
    hi = t.origin + t.spacing * (np.asarray(t.size) - 1)
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    return float(np.max(np.linalg.norm(corners, axis=1)))


# ------------------------------------------------------------ meshes ----
def load_obj(path, verbose: bool = False):
    """Vertices (n, 3) float64 and triangles (m, 3) int32 of a Wavefront OBJ
    file, read by the native OBJLoader::loadOBJMesh (OBJLoader.h): polygons
    fan-split, 1-based and negative indices resolved. Raises OSError when the
    file cannot be read or is malformed."""
    import ctypes as C

    from .solver import solver_lib

    lib = solver_lib()
    h = lib.fdmx_obj_load(str(path).encode(), int(verbose))
    if not h:
        raise OSError(lib.fdmx_last_error().decode())
    try:
        v = np.zeros((lib.fdmx_obj_num_vertices(h), 3))
        f = np.zeros((lib.fdmx_obj_num_faces(h), 3), np.int32)
        lib.fdmx_obj_get(h, v.ctypes.data_as(C.POINTER(C.c_double)), f.ctypes.data_as(C.POINTER(C.c_int)))
    finally:
        lib.fdmx_obj_free(h)
    return v, f


class MeshParticle:
    """A closed triangle mesh as a level set (LevelSetObject::
    TriangleMeshParticle, LevelSetObjectGenerator.h:382-951): the grid of
    signed distances is sampled on the GPU."""

    def __init__(self, vertices, triangles, positive_inside: bool = False):
        import ctypes as C

        from .solver import solver_lib

        self._C = C
        self._lib = solver_lib()
        self.vertices = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1, 3))
        self.triangles = np.ascontiguousarray(np.asarray(triangles, np.int32).reshape(-1, 3))
        self.h = self._lib.fdmx_mesh_create(self._dp(self.vertices), len(self.vertices),
                                            self._ip(self.triangles), len(self.triangles),
                                            int(positive_inside))
        if not self.h:
            raise RuntimeError(self._lib.fdmx_last_error().decode())

    def _dp(self, a):
        return a.ctypes.data_as(self._C.POINTER(self._C.c_double))

    def _ip(self, a):
        return a.ctypes.data_as(self._C.POINTER(self._C.c_int))

    def _ok(self, r):
        if r != 0:
            raise RuntimeError(self._lib.fdmx_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self._lib.fdmx_mesh_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def valid(self) -> bool:
        return bool(self._lib.fdmx_mesh_is_valid(self.h))

    def build_grid(self, spacing: float, padding: int = 2):
        self._ok(self._lib.fdmx_mesh_build_grid(self.h, float(spacing), int(padding)))
        return self.grid()

    def build_grid_by_resolution(self, nodes_per_diameter: int = 50, padding: int = 2):
        self._ok(self._lib.fdmx_mesh_build_grid_by_resolution(self.h, int(nodes_per_diameter),
                                                                int(padding)))
        return self.grid()

    def grid(self):
        """(origin (3,), spacing, size (3,), values (nx*ny*nz,) x fastest), or
        None before a grid is built (or for an invalid mesh)."""
        o, sz, h = np.zeros(3), np.zeros(3, np.int32), np.zeros(1)
        self._lib.fdmx_mesh_grid_info(self.h, self._dp(o), self._ip(sz), self._dp(h))
        if sz.prod() == 0:
            return None
        vals = np.zeros(int(np.prod(sz.astype(np.int64))))
        self._lib.fdmx_mesh_grid_values(self.h, self._dp(vals))
        return o, float(h[0]), tuple(int(x) for x in sz), vals

    def evaluate(self, points):
        """Signed distance (negative inside) and gradient at points (n, 3)."""
        p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1, 3))
        val, grad = np.zeros(len(p)), np.zeros((len(p), 3))
        self._ok(self._lib.fdmx_mesh_evaluate(self.h, self._dp(p), len(p), self._dp(val),
                                              self._dp(grad)))
        return val, grad

    def output_grid_vtu(self, prefix: str):
        self._ok(self._lib.fdmx_mesh_output_grid_vtu(self.h, str(prefix).encode()))


def mesh_template(vertices, triangles, spacing: float = None, nodes_per_diameter: int = 50,
                  padding: int = 2) -> Template:
    """Template of a closed triangle mesh: its level-set grid (spacing, or the
    largest box extent over nodes_per_diameter) sampled on the GPU, and its
    vertices as surface nodes (ordered into compact 64-node chunks, the
    triangles renumbered to match), as the reference's tutorial feeds a
    TriangleMeshParticle to addParticle (apps/tutorial/LSParticleFalling.cpp:
    57-103)."""
    m = MeshParticle(vertices, triangles)
    try:
        g = m.build_grid(spacing, padding) if spacing else m.build_grid_by_resolution(
            nodes_per_diameter, padding)
    finally:
        m.close()
    if g is None:
        raise ValueError("invalid mesh: no level-set grid")
    origin, h, size, vals = g
    v = np.asarray(vertices, np.float64).reshape(-1, 3)
    order = chunk_order(v)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    tris = inv[np.asarray(triangles, np.int64).reshape(-1, 3)].astype(np.int32)
    return Template(origin=origin, spacing=h, size=size, sdf=vals.astype(np.float32),
                    nodes=v[order].astype(np.float32), tris=tris)


def icosphere(radius: float = 1.0, subdivisions: int = 2):
    """Vertices and outward triangles of a subdivided icosahedron on a sphere."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts) * radius, np.array(faces, np.int32)


def box_mesh(lx: float, ly: float, lz: float):
    """The 8 corners and 12 outward triangles of an axis-aligned box centred
    at the origin."""
    hx, hy, hz = 0.5 * lx, 0.5 * ly, 0.5 * lz
    v = np.array([[x, y, z] for z in (-hz, hz) for y in (-hy, hy) for x in (-hx, hx)])
    f = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4],
                  [2, 6, 3], [3, 6, 7], [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]], np.int32)
    return v, f


def torus_mesh(R: float = 1.0, r: float = 0.4, nu: int = 48, nv: int = 24):
    """Vertices and outward triangles of a torus about z (tube radius r,
    centre-line radius R) on an nu x nv parameter grid: a closed mesh that is
    not convex."""
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    uu, ww = np.meshgrid(u, w, indexing="ij")
    v = np.stack([(R + r * np.cos(ww)) * np.cos(uu), (R + r * np.cos(ww)) * np.sin(uu),
                  r * np.sin(ww)], -1).reshape(-1, 3)
    f = []
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            f += [(a, b, c), (a, c, d)]
    return v, np.array(f, np.int32)
