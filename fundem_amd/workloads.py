"""Seeded workloads: the configurations the solver is exercised and measured on.

Each builder returns a `Scenario` (inputs only). `apply(scenario, target)`
feeds it through the solver API method names, so the same scenario drives
the device solver and any object with the same methods.

Conventions recorded here:
  * "exponents p1, p2" are superellipsoid powers (ee = 2/p1, en = 2/p2);
  * damping ratio zeta maps to restitution e = exp(-pi zeta / sqrt(1 - zeta^2))
    (the inverse of contactKernel.cuh:50-51);
  * kt = 0.4 kn where a configuration does not give it;
  * periodic boxes put max.y equal to max.z, so the ghost-hash z test, which
    compares with max.y (mySpatialGrid.cu:236), coincides with the z bound.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import generators as gen


def restitution_from_damping(zeta: float) -> float:
    return math.exp(-math.pi * zeta / math.sqrt(1.0 - zeta * zeta))


@dataclass
class Scenario:
    name: str
    templates: list
    density: object  # one value, or one per template
    materials: list  # (kn, ks, mu, e)
    x: np.ndarray
    v: np.ndarray
    w: np.ndarray
    q: np.ndarray
    mat: np.ndarray
    geo: np.ndarray
    domain_min: np.ndarray
    domain_max: np.ndarray
    gravity: np.ndarray
    dt: float
    wall_templates: list = field(default_factory=list)
    wall_materials: list = field(default_factory=list)  # (mu, e)
    walls: list = field(default_factory=list)  # (x, v, w, q, mat, geo)
    periodic: tuple = (False, False)
    sector: bool = False  # the quarter-plane sector periodic boundary about min
    cell_size: float = 0.0
    bonds: list = field(default_factory=list)  # (i, j, point, normal, R, L, E, nu, st, c, mub)
    steps: int = 1
    # materials given as (kn, ks, mu, zeta) and wall materials as (mu, zeta)
    damping_ratio: bool = False
    # templates added with zero mass (fixed particles: invM = 0)
    fixed_templates: tuple = ()
    # explicit (mass, inertia6) per template instead of the density path
    # (which also shifts each template by its centroid)
    mass_properties: list = None

    @property
    def n(self) -> int:
        return len(self.x)


def apply(s: Scenario, t) -> None:
    """Feed a scenario into a solver-like object."""
    for m in s.materials:
        (t.add_material_damping if s.damping_ratio else t.add_material)(*m)
    dens = s.density if isinstance(s.density, (list, tuple)) else [s.density] * len(s.templates)
    for g, (d, tp) in enumerate(zip(dens, s.templates)):
        if g in s.fixed_templates:
            t.add_geometry(0.0, np.zeros(6), tp)
        elif s.mass_properties is not None:
            t.add_geometry(s.mass_properties[g][0], s.mass_properties[g][1], tp)
        else:
            t.add_geometry_density(d, tp)
    for m in s.wall_materials:
        (t.add_wall_material_damping if s.damping_ratio else t.add_wall_material)(*m)
    for tp in s.wall_templates:
        t.add_wall_geometry(tp)
    t.add_particles(s.x, s.v, s.w, s.q, s.mat, s.geo)
    for wl in s.walls:
        t.add_wall(*wl)
    for b in s.bonds:
        t.add_bond(*b)
    if any(s.periodic):
        t.set_periodic(*s.periodic)
    if s.sector:
        t.set_sector(True)
    if s.cell_size > 0:
        t.set_cell_size(s.cell_size)
    t.initialize(s.domain_min, s.domain_max, s.gravity, s.dt)


def apply_batch(scenarios, t) -> None:
    """Feed several scenarios into one solver as independent systems. They
    share the particle templates (added once, from the first scenario), the
    time step and gravity; materials, walls and wall templates are per system."""
    base = scenarios[0]
    dens = base.density if isinstance(base.density, (list, tuple)) else [base.density] * len(base.templates)
    for g, (d, tp) in enumerate(zip(dens, base.templates)):
        if g in base.fixed_templates:
            t.add_geometry(0.0, np.zeros(6), tp)
        elif base.mass_properties is not None:
            t.add_geometry(base.mass_properties[g][0], base.mass_properties[g][1], tp)
        else:
            t.add_geometry_density(d, tp)
    mat_off = wmat_off = wt_off = 0
    for s in scenarios:
        if s.templates is not base.templates and len(s.templates) != len(base.templates):
            raise ValueError("systems of one batch share their particle templates")
        if s.bonds:
            raise ValueError("bonded systems are not batched")
        if s.dt != base.dt or not np.array_equal(s.gravity, base.gravity):
            raise ValueError("systems of one batch share dt and gravity")
        if s.damping_ratio != base.damping_ratio:
            raise ValueError("systems of one batch give their materials the same way")
        t.add_system(s.domain_min, s.domain_max, s.periodic[0], s.periodic[1])
        for m in s.materials:
            (t.add_material_damping if s.damping_ratio else t.add_material)(*m)
        for m in s.wall_materials:
            (t.add_wall_material_damping if s.damping_ratio else t.add_wall_material)(*m)
        for tp in s.wall_templates:
            t.add_wall_geometry(tp)
        t.add_particles(s.x, s.v, s.w, s.q, np.asarray(s.mat, np.int32) + mat_off, s.geo)
        for pos, v, w, q, mat, geo in s.walls:
            t.add_wall(pos, v, w, q, mat + wmat_off, geo + wt_off)
        mat_off += len(s.materials)
        wmat_off += len(s.wall_materials)
        wt_off += len(s.wall_templates)
    if base.cell_size > 0:
        t.set_cell_size(base.cell_size)
    t.initialize(base.domain_min, base.domain_max, base.gravity, base.dt)


IDENT = np.array([1.0, 0.0, 0.0, 0.0])


def compressed_pour(scn: Scenario, make_solver, closing: float = 5e-3, max_steps: int = 20000,
                    check_every: int = 200):
    """The pour's lattice driven into a contact-dense state: every particle
    starts falling at a speed proportional to its height above the lowest,
    so the layers close in on one another within about `closing` seconds;
    once the particle-particle node contacts stop growing (checked every
    `check_every` steps) the state is returned at rest (v = w = 0), with the
    number of steps run and of particle node contacts. `make_solver()` gives
    an empty solver with the capacities the dense state needs."""
    import dataclasses
    v = scn.v.copy()
    v[:, 2] = -(scn.x[:, 2] - scn.x[:, 2].min()) / closing
    s = make_solver()
    apply(dataclasses.replace(scn, v=v), s)
    best, done, c = 0, 0, 0
    while done < max_steps:
        s.step(check_every)
        done += check_every
        c = int(s.contacts(0)["A"].size)
        if c >= 4 * scn.n or (best > scn.n and c < best):
            break
        best = max(best, c)
    if s.status() != 0:
        raise RuntimeError(f"capacity exceeded while compressing (status {s.status()})")
    st = s.state()
    s.close()
    dense = dataclasses.replace(scn, x=st["x"], q=st["q"], v=np.zeros_like(st["v"]),
                                w=np.zeros_like(st["w"]))
    return dense, done, c


def floor_wall(xy_min, xy_max, z0: float = 0.0):
    """A plane wall at height z0 covering the given xy range: the reference's
    plane construction (apps/test/LSSphereBouncing.cpp:86-127) built around
    the plane's own centre, and the wall placed at that centre. The wall's
    position must lie inside the domain, or its cell id is -1 and no node
    ever finds it."""
    c = np.array([0.5 * (xy_min[0] + xy_max[0]), 0.5 * (xy_min[1] + xy_max[1])])
    lo = np.asarray(xy_min, dtype=np.float64) - c
    hi = np.asarray(xy_max, dtype=np.float64) - c
    return gen.plane_wall(lo, hi), (np.array([c[0], c[1], z0]), np.zeros(3), np.zeros(3),
                                   IDENT, 0, 0)


def single_particle_on_wall(n_nodes: int = 2000, seed: int = 7, steps: int = 100000) -> Scenario:
    """One superellipsoid (semi-axes 10, 8, 6 mm, powers 4, 4, density 2500)
    rolling on a plane under gravity: 2000 nodes, 32^3 SDF, kn = 1e5 N/m,
    kt = 0.4 kn, damping ratio 0.3, mu = 0.3, v0 = 0.5 m/s along x, random
    orientation, dt = 1e-5 s."""
    rng = np.random.default_rng(seed)
    t = gen.superellipsoid_template(0.010, 0.008, 0.006, 4.0, 4.0, n_grid=32, n_nodes=n_nodes)
    q = gen.shoemake_quaternions(rng, 1)
    # start with the lowest node a hair above the floor
    R = _rotmat(q[0])
    lowest = float(np.min((t.nodes.astype(np.float64) @ R.T)[:, 2]))
    x = np.array([[0.0, 0.0, -lowest + 1e-5]])
    wt, wall = floor_wall([-0.5, -0.5], [0.5, 0.5])
    kn = 1e5
    return Scenario(
        name="single-particle-roll", templates=[t], density=2500.0,
        materials=[(kn, 0.4 * kn, 0.3, restitution_from_damping(0.3))],
        x=x, v=np.array([[0.5, 0.0, 0.0]]), w=np.zeros((1, 3)), q=q,
        mat=np.zeros(1, np.int32), geo=np.zeros(1, np.int32),
        domain_min=np.array([-0.5, -0.5, -0.5]), domain_max=np.array([0.5, 0.5, 0.5]),
        gravity=np.array([0.0, 0.0, -9.81]), dt=1e-5,
        wall_templates=[wt], wall_materials=[(0.3, restitution_from_damping(0.3))], walls=[wall],
        steps=steps)


def _rotmat(q):
    q0, q1, q2, q3 = q
    return np.array([
        [1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
        [2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1)],
        [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)],
    ])


def superellipsoid_family(count: int, seed: int, n_grid: int = 32, n_nodes: int = 1000,
                          axes=(0.8e-3, 1.2e-3), powers=(2.0, 8.0)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        a, b, c = rng.uniform(*axes, size=3)
        p1, p2 = rng.uniform(*powers, size=2)
        out.append(gen.superellipsoid_template(a, b, c, p1, p2, n_grid=n_grid, n_nodes=n_nodes))
    return out


def pour(counts=(100, 100, 100), spacing: float = None, n_templates: int = 16,
         n_nodes: int = 1000, n_grid: int = 32, seed: int = 2024, cell_size: float = 2.5e-3,
         kn: float = 1e4, mu: float = 0.5, zeta: float = 0.3, dt: float = 5e-7,
         steps: int = 10000, templates=None, clearance: float = 0.05,
         rest_on_floor: bool = True) -> Scenario:
    """Pour of superellipsoids (semi-axes U[0.8, 1.2] mm, powers U[2, 8],
    16 templates, 32^3 SDF, 1000 nodes) from a jittered lattice with random
    orientations, in a box periodic in x and y with a floor wall at z = 0,
    cell size 2.5 mm, kn = 1e4 N/m, ks = 0.4 kn, damping ratio 0.3, mu = 0.5,
    dt = 5e-7 s.

    The lattice spacing defaults to (1 + clearance) times the largest
    bounding diameter and the jitter to 1% of the spacing, so no two bounding
    spheres (hence no two level sets) overlap initially. With rest_on_floor
    every particle of the bottom layer is lowered until its lowest surface
    node touches the floor, so wall contacts are live from the first step;
    the layers above fall onto it."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = counts
    tmpl = templates if templates is not None else superellipsoid_family(
        n_templates, seed + 1, n_grid=n_grid, n_nodes=n_nodes)
    rmax = max(gen.template_radius(t) for t in tmpl)
    if spacing is None:
        spacing = 2.0 * rmax * (1.0 + clearance)
    jitter = 0.01 * spacing
    if spacing - 2.0 * np.sqrt(3.0) * jitter < 2.0 * rmax:
        raise ValueError("lattice spacing too small for non-overlapping particles")
    n = nx * ny * nz
    Lx, Ly = nx * spacing, ny * spacing
    H = nz * spacing + 2.0 * spacing
    H = max(H, Ly)
    z0 = rmax + jitter
    # y spans [H - Ly, H) so that max.y == max.z
    x = gen.jittered_lattice(counts, spacing, (0.5 * spacing, H - Ly + 0.5 * spacing, z0),
                             jitter, rng)
    dmin = np.array([0.0, H - Ly, 0.0])
    dmax = np.array([Lx, H, H])
    q = gen.shoemake_quaternions(rng, n)
    geo = rng.integers(0, len(tmpl), size=n).astype(np.int32)
    if rest_on_floor:
        bottom = np.arange(nx * ny)
        for i in bottom:
            R = _rotmat(q[i])
            low = float(np.min((tmpl[geo[i]].nodes.astype(np.float64) @ R.T)[:, 2]))
            x[i, 2] = -low
    wt, wall = floor_wall([0.0, H - Ly], [Lx, H])
    e = restitution_from_damping(zeta)
    return Scenario(
        name=f"pour-{n}", templates=tmpl, density=2500.0,
        materials=[(kn, 0.4 * kn, mu, e)],
        x=x, v=np.zeros((n, 3)), w=np.zeros((n, 3)), q=q,
        mat=np.zeros(n, np.int32), geo=geo,
        domain_min=dmin, domain_max=dmax, gravity=np.array([0.0, 0.0, -9.81]), dt=dt,
        wall_templates=[wt], wall_materials=[(mu, e)], walls=[wall],
        periodic=(True, True), cell_size=cell_size, steps=steps)


def packing(n: int, seed: int, templates, kn: float = 1e4, mu: float = 0.5, zeta: float = 0.3,
            dt: float = 5e-7, cell_size: float = 2.5e-3, clearance: float = 0.05) -> Scenario:
    """One packing of a batch: n particles of the shared templates on a
    jittered lattice (spacing 1.05 bounding diameters, no initial overlap)
    filling a box periodic in x and y, about twice as tall as wide, with a
    floor; the bottom layer rests on the floor."""
    rng = np.random.default_rng(seed)
    rmax = max(gen.template_radius(t) for t in templates)
    spacing = 2.0 * rmax * (1.0 + clearance)
    side = max(2, int(math.ceil((n / 2.0) ** (1.0 / 3.0))))
    layers = int(math.ceil(n / (side * side)))
    jitter = 0.01 * spacing
    L = side * spacing
    H = layers * spacing + 2.0 * spacing
    H = max(H, L)
    x = gen.jittered_lattice((side, side, layers), spacing,
                             (0.5 * spacing, H - L + 0.5 * spacing, rmax + jitter), jitter, rng)[:n]
    q = gen.shoemake_quaternions(rng, n)
    geo = rng.integers(0, len(templates), size=n).astype(np.int32)
    for i in range(min(n, side * side)):
        R = _rotmat(q[i])
        x[i, 2] = -float(np.min((templates[geo[i]].nodes.astype(np.float64) @ R.T)[:, 2]))
    dmin = np.array([0.0, H - L, 0.0])
    dmax = np.array([L, H, H])
    wt, wall = floor_wall([0.0, H - L], [L, H])
    e = restitution_from_damping(zeta)
    return Scenario(
        name=f"packing-{n}-{seed}", templates=templates, density=2500.0,
        materials=[(kn, 0.4 * kn, mu, e)],
        x=x, v=np.zeros((n, 3)), w=np.zeros((n, 3)), q=q,
        mat=np.zeros(n, np.int32), geo=geo,
        domain_min=dmin, domain_max=dmax, gravity=np.array([0.0, 0.0, -9.81]), dt=dt,
        wall_templates=[wt], wall_materials=[(mu, e)], walls=[wall],
        periodic=(True, True), cell_size=cell_size, steps=5000)


def batch_packings(count: int = 512, seed: int = 5, n_range=(500, 20000), templates=None,
                   kn_range=(5e3, 2e4), zeta_range=(0.1, 0.5), mu_range=(0.2, 0.8)):
    """The batched calibration sweep: `count` packings whose particle counts
    are log-uniform in n_range, sharing 16 superellipsoid templates (32^3 SDF,
    1000 nodes), each with its own kn, damping ratio and mu drawn uniformly
    from the given ranges."""
    rng = np.random.default_rng(seed)
    tmpl = templates if templates is not None else superellipsoid_family(16, seed + 1)
    ns = np.round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), size=count))).astype(int)
    out = []
    for k in range(count):
        out.append(packing(int(ns[k]), seed * 100003 + k, tmpl, kn=rng.uniform(*kn_range),
                           mu=rng.uniform(*mu_range), zeta=rng.uniform(*zeta_range)))
    return out


def column(counts=(8, 8, 64), seed: int = 4, n_templates: int = 8, n_nodes: int = 800,
           kn: float = 1e4, zeta: float = 0.3, mu: float = 0.5, dt: float = 1e-6,
           v_sd: float = 0.01, floor_side: int = 16, steps: int = 2000) -> Scenario:
    """The differentiated column collapse: counts[0] x counts[1] x counts[2]
    superellipsoids (8 templates with semi-axes U[0.8, 1.2] mm and powers
    U[2, 8], 32^3 SDF, 800 nodes) on a jittered lattice (spacing 1.05
    bounding diameters, no initial overlap) released above a rough floor: a
    floor wall carrying a floor_side^2 layer of fixed particles (the same
    shapes with zero mass, so invM = 0). Materials by damping ratio: kn,
    ks = 0.4 kn, mu, zeta for particles and the wall; initial velocities
    N(0, v_sd) per component; dt = 1e-6 s."""
    rng = np.random.default_rng(seed)
    tmpl = superellipsoid_family(n_templates, seed + 1, n_nodes=n_nodes)
    rmax = max(gen.template_radius(t) for t in tmpl)
    sp = 2.0 * rmax * 1.05
    nx, ny, nz = counts
    F = floor_side
    L = F * sp
    x0 = 0.5 * (L - nx * sp)
    y0 = 0.5 * (L - ny * sp)
    zf = rmax  # centres of the fixed layer
    cols = gen.jittered_lattice((nx, ny, nz), sp, (x0 + 0.5 * sp, y0 + 0.5 * sp, zf + sp * 1.05),
                                0.01 * sp, rng)
    n = len(cols)
    fx = np.array([[(i + 0.5) * sp, (j + 0.5) * sp, zf] for j in range(F) for i in range(F)])
    x = np.vstack([cols, fx])
    q = gen.shoemake_quaternions(rng, len(x))
    geo = np.concatenate([rng.integers(0, n_templates, size=n),
                          n_templates + rng.integers(0, n_templates, size=len(fx))]).astype(np.int32)
    for i in range(n, len(x)):  # fixed particles rest on the floor
        R = _rotmat(q[i])
        x[i, 2] = -float(np.min((tmpl[geo[i] - n_templates].nodes.astype(np.float64) @ R.T)[:, 2]))
    v = np.zeros((len(x), 3))
    v[:n] = rng.normal(0.0, v_sd, size=(n, 3))
    H = cols[:, 2].max() + 2 * sp
    dmin = np.array([0.0, 0.0, -sp])
    dmax = np.array([L, L, H])
    wt, wall = floor_wall([0.0, 0.0], [L, L])
    return Scenario(
        name=f"column-{n}", templates=tmpl + tmpl, density=2500.0,
        materials=[(kn, 0.4 * kn, mu, zeta)],
        x=x, v=v, w=np.zeros((len(x), 3)), q=q, mat=np.zeros(len(x), np.int32), geo=geo,
        domain_min=dmin, domain_max=dmax, gravity=np.array([0.0, 0.0, -9.81]), dt=dt,
        wall_templates=[wt], wall_materials=[(mu, zeta)], walls=[wall],
        damping_ratio=True, fixed_templates=tuple(range(n_templates, 2 * n_templates)),
        cell_size=0.0, steps=steps)


def cluster(n: int = 27, seed: int = 1, n_nodes: int = 256, n_grid: int = 16, n_templates: int = 3,
            periodic=(False, False), walls: bool = True, scale: float = 1e-3,
            spacing: float = 3.0, speed: float = 0.5, spin: float = 20.0, dt: float = 1e-6,
            kn: float = 1e4, mu: float = 0.4, zeta: float = 0.3, steps: int = 400,
            templates=None) -> Scenario:
    """A small cluster with random velocities that collides with itself, the
    floor and (when periodic) across the periodic faces within a few hundred
    steps. Lengths are in units of `scale` (semi-axes U[0.8, 1.2], lattice
    spacing `spacing`); at scale 1e-3 the inertia determinant falls below the
    reference's 1e-30 inverse threshold, so particles keep their spin, while
    at scale 1e-2 rotation is live. Used for step-by-step parity."""
    rng = np.random.default_rng(seed)
    tmpl = templates if templates is not None else superellipsoid_family(
        n_templates, seed + 100, n_grid=n_grid, n_nodes=n_nodes, axes=(0.8 * scale, 1.2 * scale))
    n_templates = len(tmpl)
    sp = spacing * scale
    side = int(math.ceil(n ** (1.0 / 3.0)))
    L = side * sp
    H = L + 2 * sp
    x = gen.jittered_lattice((side, side, side), sp, (0.5 * sp, H - L + 0.5 * sp, 0.75 * sp),
                             0.05 * sp, rng)[:n]
    # y spans [H - L, H) so that max.y == max.z
    dmin = np.array([0.0, H - L, 0.0])
    dmax = np.array([L, H, H])
    q = gen.shoemake_quaternions(rng, n)
    v = rng.normal(0.0, speed, size=(n, 3))
    w = rng.normal(0.0, spin, size=(n, 3))
    geo = rng.integers(0, n_templates, size=n).astype(np.int32)
    e = restitution_from_damping(zeta)
    wts, wls = [], []
    if walls:
        wt, wall = floor_wall([dmin[0], dmin[1]], [dmax[0], dmax[1]])
        wts, wls = [wt], [wall]
    return Scenario(
        name=f"cluster-{n}-{seed}", templates=tmpl, density=2500.0,
        materials=[(kn, 0.4 * kn, mu, e), (2 * kn, 0.8 * kn, 0.5 * mu, 0.9)],
        x=x, v=v, w=w, q=q, mat=(np.arange(n) % 2).astype(np.int32), geo=geo,
        domain_min=dmin, domain_max=dmax, gravity=np.array([0.0, 0.0, -9.81]), dt=dt,
        wall_templates=wts, wall_materials=[(mu, e)] if walls else [], walls=wls,
        periodic=tuple(periodic), cell_size=0.0, steps=steps)


def fcc_sites(cells, L: float, origin):
    """Sites of a face-centred cubic lattice of nearest-neighbour distance L
    (cube side L sqrt 2, four sites per cube), cells[0] x cells[1] x cells[2]
    cubes, ordered by cube (x fastest) and basis site."""
    a = L * math.sqrt(2.0)
    nx, ny, nz = cells
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    corner = np.stack([i, j, k], axis=-1).reshape(-1, 1, 3).astype(np.float64)
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    return ((corner + basis[None]) * a).reshape(-1, 3) + np.asarray(origin, np.float64)


def bonded_beam(counts=(20, 4, 4), radius: float = 0.5e-3, n_nodes: int = 400, n_grid: int = 16,
                seed: int = 3, strength_mean: float = 5e6, strength_sd: float = 0.5e6,
                tip_speed: float = 1.0, dt: float = 1e-7, steps: int = 50000,
                lattice: str = "cubic") -> Scenario:
    """Cantilever of bonded spheres (radius 0.5 mm, 16^3 SDF, 400 nodes) at
    nearest-neighbour distance L = 2r, on a simple cubic lattice (`counts`
    sites per axis, bonds to the 6 neighbours) or a face-centred cubic one
    (`counts` cubes per axis, four sites each, bonds to the 12 neighbours:
    about six bonds per site), with normal stiffness 1e9 Pa per area
    (E = 1e9 * L) and shear/normal 0.5 (bond radius L / sqrt(6)); tensile
    strength and cohesion ~ N(5, 0.5) MPa. The sites of the x = min plane are
    fixed (zero mass, zero velocity) and those of the x = max plane are
    driven at tip_speed along -z (zero mass)."""
    rng = np.random.default_rng(seed)
    L = 2.0 * radius
    sph = gen.sphere_template(radius, n_grid=n_grid, n_nodes=n_nodes)
    if lattice == "cubic":
        pts = gen.jittered_lattice(counts, L, (radius, radius, radius), 0.0, rng)
        nx, ny, nz = counts
        ii = np.arange(len(pts))
        ix, iy, iz = ii % nx, (ii // nx) % ny, ii // (nx * ny)
        pairs = []
        for dx, dy, dz in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            ok = (ix + dx < nx) & (iy + dy < ny) & (iz + dz < nz)
            i = ii[ok]
            pairs.append(np.stack([i, i + dx + nx * (dy + ny * dz)], axis=1))
        pairs = np.concatenate(pairs)
        pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    elif lattice == "fcc":
        from scipy.spatial import cKDTree
        pts = fcc_sites(counts, L, (radius, radius, radius))
        pairs = np.array(sorted(cKDTree(pts).query_pairs(1.001 * L)), dtype=np.int64)
    else:
        raise ValueError(f"unknown lattice {lattice!r}")
    n = len(pts)
    tol = 1e-3 * L
    geo = np.zeros(n, np.int32)
    geo[pts[:, 0] <= pts[:, 0].min() + tol] = 1  # clamp
    tip = pts[:, 0] >= pts[:, 0].max() - tol
    geo[tip] = 2                                  # driven tip
    v = np.zeros((n, 3))
    v[tip, 2] = -tip_speed
    bonds = []
    E = 1e9 * L
    Rb = L / math.sqrt(6.0)
    strengths = rng.normal(strength_mean, strength_sd, size=(len(pairs), 2))
    for (i, j), (st, coh) in zip(pairs, strengths):
        i, j = int(i), int(j)
        bonds.append((i, j, 0.5 * (pts[i] + pts[j]), pts[i] - pts[j], Rb, L, E, 0.3, float(st),
                      float(coh), 0.0))
    margin = 4 * L
    dmin = pts.min(axis=0) - margin
    dmax = pts.max(axis=0) + margin
    s = Scenario(
        name=f"beam-{n}", templates=[sph, sph, sph], density=[2500.0, 0.0, 0.0],
        materials=[(1e4, 4e3, 0.3, restitution_from_damping(0.3))],
        x=pts, v=v, w=np.zeros((n, 3)), q=np.tile(IDENT, (n, 1)),
        mat=np.zeros(n, np.int32), geo=geo, domain_min=dmin, domain_max=dmax,
        gravity=np.array([0.0, 0.0, -9.81]), dt=dt, bonds=bonds, steps=steps)
    return s
