"""Python binding of the C++ host API (LSSolver) through libfundem_solver.so.

`Solver` mirrors LSSolver's methods (kernel/LSSolver.h:14-467) with numpy
arrays for vectors. It runs only on the HIP path: constructing it raises if
the native libraries are missing or no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)

_LIBS = {}


class NativeLibraryMissing(RuntimeError):
    pass


def solver_lib(precision: str = "fp32"):
    """Load libfundem_solver.so and through it libfundem_amd.so (FP32 state
    storage), or their FP64-storage builds (precision "fp64"). Both builds
    export the same names: the FP64 pair is loaded with its own symbols
    bound ahead of any already loaded (RTLD_DEEPBIND), so one process can use
    both."""
    if precision in _LIBS:
        return _LIBS[precision]
    core, solver = _build.libs(precision)
    if not (os.path.exists(solver) and os.path.exists(core)):
        raise NativeLibraryMissing(
            f"native libraries not built ({solver}); run fundem_amd.build.build_all()")
    if precision == "fp32":
        C.CDLL(core, mode=C.RTLD_GLOBAL)
        lib = C.CDLL(solver)
    else:
        lib = C.CDLL(solver, mode=C.RTLD_LOCAL | os.RTLD_DEEPBIND)
    I, LL, D, P = C.c_int, C.c_longlong, C.c_double, C.c_void_p
    sig = {
        "fdmx_last_error": (C.c_char_p, []),
        "fdmx_create": (P, [I]),
        "fdmx_destroy": (None, [P]),
        "fdmx_add_material": (I, [P, D, D, D, D]),
        "fdmx_add_wall_material": (I, [P, D, D]),
        "fdmx_add_material_damping": (I, [P, D, D, D, D]),
        "fdmx_add_wall_material_damping": (I, [P, D, D]),
        "fdmx_num_materials": (I, [P, C.POINTER(I), C.POINTER(I)]),
        "fdmx_autotune": (I, [P, I, I, _dp, C.c_char_p, I]),
        "fdmx_contact_geometry": (LL, [P, I, _ip, _ip, _ip, _dp, _dp, _dp, LL]),
        "fdmx_create_bonds_from_contacts": (I, [P, D, D, D, D, D, D, D]),
        "fdmx_num_bonds": (LL, [P]),
        "fdmx_adjoint_create": (P, [P]),
        "fdmx_adjoint_destroy": (None, [P]),
        "fdmx_adjoint_set_contacts": (I, [P, I]),
        "fdmx_adjoint_forward": (I, [P, LL, LL]),
        "fdmx_adjoint_backward": (I, [P, I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp]),
        "fdmx_add_geometry": (I, [P, D, _dp, _dp, D, _ip, _dp, _dp, I, _ip, I]),
        "fdmx_add_geometry_density": (I, [P, D, _dp, D, _ip, _dp, _dp, I, _dp, _ip, I]),
        "fdmx_add_wall_geometry": (I, [P, _dp, D, _ip, _dp, _dp, I, _ip, I]),
        "fdmx_write_vtu": (I, [P, C.c_char_p, LL, LL, D]),
        "fdmx_obj_load": (P, [C.c_char_p, I]),
        "fdmx_obj_num_vertices": (LL, [P]),
        "fdmx_obj_num_faces": (LL, [P]),
        "fdmx_obj_get": (None, [P, _dp, _ip]),
        "fdmx_obj_free": (None, [P]),
        "fdmx_mesh_create": (P, [_dp, I, _ip, I, I]),
        "fdmx_mesh_destroy": (None, [P]),
        "fdmx_mesh_is_valid": (I, [P]),
        "fdmx_mesh_build_grid": (I, [P, D, I]),
        "fdmx_mesh_build_grid_by_resolution": (I, [P, I, I]),
        "fdmx_mesh_grid_info": (None, [P, _dp, _ip, _dp]),
        "fdmx_mesh_grid_values": (None, [P, _dp]),
        "fdmx_mesh_evaluate": (I, [P, _dp, I, _dp, _dp]),
        "fdmx_mesh_output_grid_vtu": (I, [P, C.c_char_p]),
        "fdmx_add_particles": (I, [P, I, _dp, _dp, _dp, _dp, _ip, _ip]),
        "fdmx_add_wall": (I, [P, _dp, _dp, _dp, _dp, I, I]),
        "fdmx_add_bond": (I, [P, I, I, _dp, _dp, D, D, D, D, D, D, D]),
        "fdmx_set_periodic": (I, [P, I, I]),
        "fdmx_set_sector": (I, [P, I]),
        "fdmx_set_global_damping": (I, [P, D]),
        "fdmx_set_constant_force": (I, [P, I, _dp, _dp]),
        "fdmx_set_cell_size": (I, [P, D]),
        "fdmx_set_capacity": (I, [P, I, I]),
        "fdmx_set_skin": (I, [P, D]),
        "fdmx_add_system": (I, [P, C.POINTER(D), C.POINTER(D), I, I]),
        "fdmx_pair_totals": (I, [P, C.POINTER(I)]),
        "fdmx_list_builds": (I, [P, C.POINTER(LL), C.POINTER(LL),
                                 C.POINTER(D)]),
        "fdmx_use_graph": (I, [P, I]),
        "fdmx_set_timing": (I, [P, I]),
        "fdmx_initialize": (I, [P, _dp, _dp, _dp, D]),
        "fdmx_advance": (I, [P, LL]),
        "fdmx_sync": (I, [P]),
        "fdmx_status": (I, [P, _up]),
        "fdmx_get_state": (I, [P, _dp, _dp, _dp, _dp, _dp, _dp]),
        "fdmx_num_particles": (LL, [P]),
        "fdmx_get_contacts": (LL, [P, I, _ip, _ip, _ip, _dp, LL]),
        "fdmx_get_bonds": (I, [P, _ip, _ip, _dp, _dp]),
        "fdmx_get_hash": (I, [P, I, _ip, _up, _ip, _ip, _ip]),
        "fdmx_get_grid": (I, [P, I, _dp, _dp, _dp, _ip]),
        "fdmx_get_geometry": (I, [P, I, _dp, _dp, _dp, _dp, _dp]),
        "fdmx_stage_times": (I, [P, C.c_char_p, I, _dp, C.POINTER(C.c_longlong), I]),
    }
    sig["fdmx_set_launch_tuning"] = (I, [P, _ip])
    sig["fdmx_get_launch_tuning"] = (I, [P, _ip])
    sig["fdmx_set_lookahead"] = (I, [P, I])
    sig["fdmx_state_bytes"] = (I, [])
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fdmx_state_bytes() != (8 if precision == "fp64" else 4):
        raise RuntimeError(f"{solver} does not store state in {precision}")
    _LIBS[precision] = lib
    return lib


def _d(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _dp_(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _ip_(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _tris(t):
    """(pointer, count) of a template's surface triangles, or (None, 0)."""
    tris = getattr(t, "tris", None)
    if tris is None or len(tris) == 0:
        return None, 0
    a = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 3))
    t._tris_keep = a  # kept alive for the call
    return _ip_(a), len(a)


class Solver:
    """LSSolver on the GPU (device `device`)."""

    def __init__(self, device: int = 0, precision: str = "fp32"):
        """precision: storage of the state a step carries over (velocities,
        spins, orientations, forces, torques, springs, bond constants):
        "fp32" (the production build) or "fp64" (the reference's all-FP64
        arithmetic)."""
        lib = solver_lib(precision)
        self.precision = precision
        self._lib = lib
        self.h = lib.fdmx_create(int(device))
        if not self.h:
            raise RuntimeError(lib.fdmx_last_error().decode())
        self._n = 0

    def close(self):
        if getattr(self, "h", None):
            self._lib.fdmx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ok(self, r):
        if r != 0:
            raise RuntimeError(self._lib.fdmx_last_error().decode())

    # ---------------------------------------------------------------- inputs
    def add_material(self, kn, ks, mu, e):
        self._ok(self._lib.fdmx_add_material(self.h, kn, ks, mu, e))

    def add_wall_material(self, mu, e):
        self._ok(self._lib.fdmx_add_wall_material(self.h, mu, e))

    def add_material_damping(self, kn, ks, mu, zeta):
        """A particle material given by its damping ratio instead of restitution."""
        self._ok(self._lib.fdmx_add_material_damping(self.h, kn, ks, mu, zeta))

    def add_wall_material_damping(self, mu, zeta):
        self._ok(self._lib.fdmx_add_wall_material_damping(self.h, mu, zeta))

    def contact_geometry(self, set_id):
        """Contact points, normals and overlaps of the latest step's contacts."""
        n = self._lib.fdmx_contact_geometry(self.h, set_id, None, None, None, None, None, None, 0)
        if n < 0:
            raise RuntimeError(self._lib.fdmx_last_error().decode())
        a, b, k = (np.zeros(n, np.int32) for _ in range(3))
        pt, nr, ov = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        self._lib.fdmx_contact_geometry(self.h, set_id, _ip_(a), _ip_(b), _ip_(k), _dp_(pt), _dp_(nr),
                                        _dp_(ov), n)
        return dict(A=a, B=b, k=k, point=pt, normal=nr, overlap=ov)

    def create_bonds_from_contacts(self, radius, length, youngs, poisson, tensile=0.0, cohesion=0.0,
                                   friction=0.0):
        self._ok(self._lib.fdmx_create_bonds_from_contacts(self.h, radius, length, youngs, poisson,
                                                           tensile, cohesion, friction))

    def write_vtu(self, directory, frame=0, step=0, time=0.0):
        """Write one output frame's .vtu files (LSSolver::output, LSSolver.h:717-745)
        under `directory`: LSParticle/LSObjectMesh_NNNN.vtu, LSParticle/Particle_NNNN.vtu,
        LSParticleInteraction/LSInteraction_NNNN.vtu (+ VBondedInteraction_NNNN.vtu),
        Wall/LSObjectMesh_NNNN.vtu, LSParticle-WallInteraction/LSInteraction_NNNN.vtu."""
        self._ok(self._lib.fdmx_write_vtu(self.h, str(directory).encode(), int(frame), int(step),
                                          float(time)))

    @property
    def num_bonds(self) -> int:
        return int(self._lib.fdmx_num_bonds(self.h))

    def autotune(self, trial_steps=20, tune_cell=True):
        """Pick skin, cell size and launch shapes for the fastest step on this
        system (results do not change); the solver is left where it was."""
        out = np.zeros(9)
        buf = C.create_string_buffer(16384)
        self._ok(self._lib.fdmx_autotune(self.h, int(trial_steps), int(bool(tune_cell)), _dp_(out), buf,
                                         len(buf)))
        return dict(skin=out[0], cell_size=out[1], sweep_waves=int(out[2]),
                    sweep_waves_per_cu=int(out[3]), law_block=int(out[4]), reduce_block=int(out[5]),
                    law_mode=int(out[7]), nodes_per_lane=int(out[8]), ms_per_step=out[6],
                    trials=buf.value.decode().splitlines())

    def set_launch_tuning(self, sweep_waves=0, sweep_waves_per_cu=0, law_block=0, reduce_block=0,
                          law_mode=0, nodes_per_lane=0, wall_groups=0):
        """This solver's launch shapes (results do not depend on them); zero
        keeps the process default of a field. wall_groups: 1 sweeps wall
        pairs four per wave whatever the system's size, 2 never (0: for wall
        sets sized for at least 65536 pairs)."""
        t = np.array([sweep_waves, sweep_waves_per_cu, law_block, reduce_block, law_mode,
                      nodes_per_lane, wall_groups], np.int32)
        self._ok(self._lib.fdmx_set_launch_tuning(self.h, _ip_(t)))

    def launch_tuning(self):
        t = np.zeros(7, np.int32)
        self._ok(self._lib.fdmx_get_launch_tuning(self.h, _ip_(t)))
        return dict(zip(("sweep_waves", "sweep_waves_per_cu", "law_block", "reduce_block",
                         "law_mode", "nodes_per_lane", "wall_groups"), (int(v) for v in t)))

    def set_lookahead(self, steps):
        """Steps queued ahead of the list decisions read back (0: one step
        at a time)."""
        self._ok(self._lib.fdmx_set_lookahead(self.h, int(steps)))

    def num_materials(self):
        a, b = C.c_int(0), C.c_int(0)
        self._ok(self._lib.fdmx_num_materials(self.h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def add_geometry(self, mass, inertia6, t):
        self._ok(self._lib.fdmx_add_geometry(
            self.h, mass, _dp_(_d(inertia6)), _dp_(_d(t.origin)), t.spacing,
            _ip_(np.asarray(t.size, np.int32)), _dp_(_d(t.sdf)), _dp_(_d(t.nodes)), len(t.nodes),
            *_tris(t)))

    def add_geometry_density(self, density, t):
        off = np.zeros(3)
        self._ok(self._lib.fdmx_add_geometry_density(
            self.h, density, _dp_(_d(t.origin)), t.spacing, _ip_(np.asarray(t.size, np.int32)),
            _dp_(_d(t.sdf)), _dp_(_d(t.nodes)), len(t.nodes), _dp_(off), *_tris(t)))
        return off

    def add_wall_geometry(self, t):
        self._ok(self._lib.fdmx_add_wall_geometry(
            self.h, _dp_(_d(t.origin)), t.spacing, _ip_(np.asarray(t.size, np.int32)),
            _dp_(_d(t.sdf)), _dp_(_d(t.nodes)), len(t.nodes), *_tris(t)))

    def add_particles(self, x, v, w, q, mat, geo):
        x, v, w, q = _d(x), _d(v), _d(w), _d(q)
        n = x.size // 3
        mat = np.ascontiguousarray(np.broadcast_to(np.asarray(mat, np.int32), (n,)))
        geo = np.ascontiguousarray(np.broadcast_to(np.asarray(geo, np.int32), (n,)))
        self._ok(self._lib.fdmx_add_particles(self.h, n, _dp_(x), _dp_(v), _dp_(w), _dp_(q),
                                              _ip_(mat), _ip_(geo)))
        self._n += n

    def add_wall(self, x, v, w, q, mat, geo):
        self._ok(self._lib.fdmx_add_wall(self.h, _dp_(_d(x)), _dp_(_d(v)), _dp_(_d(w)), _dp_(_d(q)),
                                         mat, geo))

    def add_bond(self, i, j, point, normal, radius, length, youngs, poisson, tensile=0.0,
                 cohesion=0.0, friction=0.0):
        self._ok(self._lib.fdmx_add_bond(self.h, i, j, _dp_(_d(point)), _dp_(_d(normal)), radius,
                                         length, youngs, poisson, tensile, cohesion, friction))

    def set_global_damping(self, c):
        """The tutorials' external-force hook (apps/tutorial/globalDamping.cu:4-16):
        F -= c |F| v/|v|, T -= c |T| w/|w| between the force reduction and the
        second half-kick; None turns it off."""
        self._ok(self._lib.fdmx_set_global_damping(self.h, -1.0 if c is None else float(c)))

    def set_constant_force(self, index, force, torque):
        """apps/tutorial/globalDamping.cu:38-49 on particle `index` (both vectors are
        added to its force, as that kernel does); index None turns it off."""
        self._ok(self._lib.fdmx_set_constant_force(self.h, -1 if index is None else int(index),
                                                   _dp_(_d(force)), _dp_(_d(torque))))

    def set_sector(self, on=True):
        """LSSolver::addPeriodicBoundarySector (LSSolver.h:410-413); off removes
        every periodic boundary."""
        self._ok(self._lib.fdmx_set_sector(self.h, int(on)))

    def set_periodic(self, x, y):
        self._ok(self._lib.fdmx_set_periodic(self.h, int(x), int(y)))

    def add_system(self, dmin, dmax, periodic_x=False, periodic_y=False):
        """Start an independent system with its own domain and periodic axes;
        bodies, walls and materials added afterwards belong to it."""
        mn = np.ascontiguousarray(dmin, dtype=np.float64)
        mx = np.ascontiguousarray(dmax, dtype=np.float64)
        k = self._lib.fdmx_add_system(self.h, _dp_(mn), _dp_(mx), int(bool(periodic_x)),
                                      int(bool(periodic_y)))
        if k < 0:
            self._ok(k)
        return k

    def set_cell_size(self, cs):
        self._ok(self._lib.fdmx_set_cell_size(self.h, float(cs)))

    def set_capacity(self, pairs_per_particle, records_per_particle):
        self._ok(self._lib.fdmx_set_capacity(self.h, int(pairs_per_particle), int(records_per_particle)))

    def set_skin(self, skin):
        """Extra reach of the candidate lists (negative: automatic); takes
        effect at initialize."""
        self._ok(self._lib.fdmx_set_skin(self.h, float(skin)))

    def list_builds(self):
        """(candidate-list builds, ghost-list builds, skin in use)."""
        b, gb, sk = C.c_longlong(), C.c_longlong(), C.c_double()
        self._ok(self._lib.fdmx_list_builds(self.h, C.byref(b), C.byref(gb), C.byref(sk)))
        return b.value, gb.value, sk.value

    def pair_totals(self):
        """Candidate pairs per set (particle, wall, ghost X, Y, XY) at the latest step."""
        out = (C.c_int * 5)()
        self._ok(self._lib.fdmx_pair_totals(self.h, out))
        return list(out)

    def use_graph(self, on=True):
        self._ok(self._lib.fdmx_use_graph(self.h, int(on)))

    def set_timing(self, on=True):
        self._ok(self._lib.fdmx_set_timing(self.h, int(on)))

    # ------------------------------------------------------------------ run
    def initialize(self, mn, mx, gravity, dt):
        self._ok(self._lib.fdmx_initialize(self.h, _dp_(_d(mn)), _dp_(_d(mx)), _dp_(_d(gravity)),
                                           float(dt)))

    def step(self, n=1):
        self._ok(self._lib.fdmx_advance(self.h, int(n)))

    def sync(self):
        self._ok(self._lib.fdmx_sync(self.h))

    def status(self) -> int:
        v = C.c_uint(0)
        self._ok(self._lib.fdmx_status(self.h, C.byref(v)))
        return int(v.value)

    # -------------------------------------------------------------- results
    @property
    def n(self) -> int:
        return int(self._lib.fdmx_num_particles(self.h))

    def state(self):
        n = self.n
        x, v, w, f, t = (np.zeros((n, 3)) for _ in range(5))
        q = np.zeros((n, 4))
        self._ok(self._lib.fdmx_get_state(self.h, _dp_(x), _dp_(v), _dp_(w), _dp_(q), _dp_(f), _dp_(t)))
        return dict(x=x, v=v, w=w, q=q, F=f, T=t)

    def contacts(self, set_id):
        n = self._lib.fdmx_get_contacts(self.h, set_id, None, None, None, None, 0)
        if n < 0:
            raise RuntimeError(self._lib.fdmx_last_error().decode())
        a, b, k = (np.zeros(n, np.int32) for _ in range(3))
        sp = np.zeros((n, 3))
        self._lib.fdmx_get_contacts(self.h, set_id, _ip_(a), _ip_(b), _ip_(k), _dp_(sp), n)
        return dict(A=a, B=b, k=k, spring=sp)

    def bonds(self, nbonds):
        act, brk = np.zeros(nbonds, np.int32), np.zeros(nbonds, np.int32)
        s, t = np.zeros(nbonds), np.zeros(nbonds)
        self._ok(self._lib.fdmx_get_bonds(self.h, _ip_(act), _ip_(brk), _dp_(s), _dp_(t)))
        return dict(active=act, break_step=brk, sigma=s, tau=t)

    def grid(self, walls=False):
        mn, mx, inv = np.zeros(3), np.zeros(3), np.zeros(3)
        size = np.zeros(3, np.int32)
        self._ok(self._lib.fdmx_get_grid(self.h, int(walls), _dp_(mn), _dp_(mx), _dp_(inv), _ip_(size)))
        return dict(min=mn, max=mx, inv=inv, size=size)

    def hash_table(self, n, walls=False):
        g = self.grid(walls)
        ng = int(np.prod(g["size"].astype(np.int64)))
        hv = np.zeros(n, np.int32)
        key = np.zeros(n, np.uint32)
        idx = np.zeros(n, np.int32)
        st, en = np.zeros(ng, np.int32), np.zeros(ng, np.int32)
        self._ok(self._lib.fdmx_get_hash(self.h, int(walls), _ip_(hv), key.ctypes.data_as(_up),
                                         _ip_(idx), _ip_(st), _ip_(en)))
        return dict(hash=hv, sorted_hash=key.astype(np.int64) - 1, index=idx, start=st, end=en)

    def geometry(self, g):
        r, im = np.zeros(1), np.zeros(1)
        I, Ii, o = np.zeros(6), np.zeros(6), np.zeros(3)
        self._ok(self._lib.fdmx_get_geometry(self.h, g, _dp_(r), _dp_(im), _dp_(I), _dp_(Ii), _dp_(o)))
        return dict(radius=r[0], inv_mass=im[0], inertia=I, inv_inertia=Ii, origin=o)

    def stage_times(self):
        """{stage: (total device ms, launches timed)} since set_timing(True)."""
        names = C.create_string_buffer(4096)
        ms = np.zeros(64)
        calls = np.zeros(64, np.int64)
        cnt = self._lib.fdmx_stage_times(self.h, names, 4096, _dp_(ms),
                                         calls.ctypes.data_as(C.POINTER(C.c_longlong)), 64)
        keys = names.value.decode().split("\n")[:cnt]
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(keys)}


class Adjoint:
    """Reverse-mode gradients of a scalar outcome of a solver's run
    (fundem::LSAdjoint): forward() advances the solver keeping checkpoints,
    backward() takes the adjoint of the final state (dJ/dx, dJ/dq, dJ/dv,
    dJ/dw per particle) and returns the gradients with respect to the initial
    state and the material parameters."""

    def __init__(self, solver: Solver, contacts_per_particle: int = 16):
        self._s = solver
        self._lib = solver._lib
        self.h = self._lib.fdmx_adjoint_create(solver.h)
        if not self.h:
            raise RuntimeError(self._lib.fdmx_last_error().decode())
        solver._ok(self._lib.fdmx_adjoint_set_contacts(self.h, int(contacts_per_particle)))

    def close(self):
        if getattr(self, "h", None):
            self._lib.fdmx_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def forward(self, steps: int, checkpoint_every: int):
        self._s._ok(self._lib.fdmx_adjoint_forward(self.h, int(steps), int(checkpoint_every)))

    def backward(self, x_bar=None, q_bar=None, v_bar=None, w_bar=None):
        n = self._s.n
        pm, wm = self._s.num_materials()

        def arr(a, k):
            return np.zeros(n * k) if a is None else _d(a)

        xb, qb, vb, wb = arr(x_bar, 3), arr(q_bar, 4), arr(v_bar, 3), arr(w_bar, 3)
        x0, v0, w0 = (np.zeros((n, 3)) for _ in range(3))
        q0 = np.zeros((n, 4))
        kn, ks, mu, zeta = (np.zeros(max(pm, 1)) for _ in range(4))
        wmu, wzeta = np.zeros(max(wm, 1)), np.zeros(max(wm, 1))
        self._s._ok(self._lib.fdmx_adjoint_backward(
            self.h, n, _dp_(xb), _dp_(qb), _dp_(vb), _dp_(wb), _dp_(x0), _dp_(q0), _dp_(v0),
            _dp_(w0), _dp_(kn), _dp_(ks), _dp_(mu), _dp_(zeta), _dp_(wmu), _dp_(wzeta)))
        return dict(x=x0, q=q0, v=v0, w=w0, kn=kn[:pm], ks=ks[:pm], mu=mu[:pm], zeta=zeta[:pm],
                    wall_mu=wmu[:wm], wall_zeta=wzeta[:wm])
