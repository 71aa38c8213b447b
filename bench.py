"""Benchmark: the level-set DEM pour of 1,000,000 superellipsoids on one GPU.

Workload (synthetic, seeded): 16 superellipsoid templates (semi-axes
U[0.8, 1.2] mm, powers U[2, 8], 32^3 FP32 SDF, 1000 surface nodes) on a
100^3 jittered lattice (spacing 1.05 bounding diameters, so no initial
overlaps; the bottom layer rests on the floor), random orientations, box
periodic in x and y, floor wall, cell size 2.5 mm, kn = 1e4 N/m,
ks = 0.4 kn, damping ratio 0.3, mu = 0.5, dt = 5e-7 s.

One step is one iteration of LSSolver's loop: kick, drift and cell ids,
the device's list-rebuild decision, list builds when needed (sort, broad
phase, ghost sets) or the kept lists, the narrow phase with contact forces
for the particle, wall and ghost sets, the reduction and the second kick.
The timed region covers exactly --steps steps with the device synchronised
on both sides and no instrumentation. A separate pass of --stage-steps
steps brackets every stage with HIP events on the solver's stream (the
stream its kernels run on); their average per launch gives the roofline of
the dominant kernel. The CPU baseline is the serial FP64 oracle
on a 512-particle sample of the same workload.

The same line carries, as extra keys (--extras 1, the default):
  dense                 the pour's lattice compressed until its layers touch,
                        restarted at rest and timed like the pour (the
                        particle-particle narrow phase under load);
  batch, batch_speedup  512 packings in one solver against the same
                        packings run one after another in plain solvers;
  adjoint, adjoint_over_forward
                        one adjoint step of the 4096-particle column collapse
                        (2000 steps, checkpoints every 100) against one
                        forward step.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--side", type=int, default=100, help="lattice side (side^3 particles)")
    p.add_argument("--graph", type=int, default=0,
                   help="replay captured graphs (no per-stage timing)")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--skin", type=float, default=-1.0,
                   help="candidate-list skin in m (negative: a tenth of the largest radius)")
    p.add_argument("--copy-probe", action="store_true",
                   help="first run a plain 1 GiB device copy (counter scaling under rocprofv3)")
    p.add_argument("--stage-steps", type=int, default=200,
                   help="pour: steps of the separate stage-timed pass (roofline)")
    p.add_argument("--autotune", type=int, default=1,
                   help="pick skin, cell size and launch shapes with the autotuner first (results "
                        "do not change)")
    p.add_argument("--settle", type=int, default=20000,
                   help="column: plain steps that pack the column before the timed runs")
    p.add_argument("--checkpoint", type=int, default=100, help="column: steps per checkpoint")
    p.add_argument("--graph-compare", type=int, default=1,
                   help="column: also time the forward step replayed from captured graphs")
    p.add_argument("--workload", choices=("pour", "dense", "batch", "column", "meshsdf"),
                   default="pour",
                   help="pour: the 1M-particle pour; batch: 512 packings in one solver against "
                        "plain runs of the same packings one after another; column: adjoint "
                        "against forward step cost; meshsdf: level-set grid of a closed mesh")
    p.add_argument("--batch-sample", type=int, default=32,
                   help="packings timed one by one for the sequential reference")
    p.add_argument("--extras", type=int, default=1,
                   help="pour: also measure the contact-dense state, the batch against "
                        "sequential plain runs and the adjoint against the forward step, as "
                        "extra keys of the same line")
    p.add_argument("--dense-closing", type=float, default=5e-3,
                   help="dense state: time in which the lattice's layers close in (s)")
    p.add_argument("--dense-max-steps", type=int, default=20000,
                   help="dense state: most compression steps")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per launch of the timed stages (FP64 positions, FP32
# v, w, q, F, T, 32-bit ids; template tables stay in cache):
#   step_begin   (kStepBegin alone; the decision kernel is timed as "decide")
#                per particle: read x 24, v 12, w 12, q 16, F 12, T 12,
#                template id 4, grid state at build 4, visibility at build 4
#                (= 100); write x 24, v 12, w 12, q 16, cell id 4 (= 68);
#                sort lists and visibility flags are written by list builds
#                only; a body whose template's inverse inertia is exactly
#                zero (its spin never changes) skips T read and w write
#                (168 - 24 = 144)
#   reduce_kick  per particle: pair counts read (particle set 8, wall 8,
#                three ghost sets 24, mirror counts 16), F T write 24,
#                template id 4, v w read and write 48, q read 16 (= 148;
#                148 - 40 = 108 for a body whose spin never changes: no w,
#                no q); per candidate pair with contacts, read from each
#                side: its record count 4 and sums 48 (bounded by 104 per
#                pair: every candidate pair counted)
#   narrow       per candidate pair: ids 8, both bodies' x q template id 88,
#                empty record range or the kept pair's record range and pair
#                sums 80, history ids 16 (= 192); per contact: velocities of
#                both bodies 48, history read 16 and write 16 (= 80)
def kernel_bytes(stage, n, pairs, contacts, fixed=0):
    """`fixed`: bodies whose template has an exactly zero inverse inertia."""
    if stage == "step_begin":
        return 168.0 * n - 24.0 * fixed
    if stage == "reduce_kick":
        return 148.0 * n - 40.0 * fixed + 104.0 * sum(pairs.values())
    if stage == "narrow":
        return 192.0 * sum(pairs.values()) + 80.0 * sum(contacts.values())
    return None


def measured_traffic(stage):
    """HBM bytes per launch of a stage from the committed counter summary
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, scaled
    by a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get("stages", {}).get(stage)
    return None if e is None else float(e["bytes_per_launch"])


def cpu_baseline(templates, budget_s):
    """Serial FP64 oracle (reference order, reference storage) on an 8^3
    sample of the same workload; particle-steps/s on one host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from fundem_amd import workloads as W
    scn = W.pour((8, 8, 8), templates=templates)
    o = po.OracleSim(False, False)
    W.apply(scn, o)
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter() - t0
    reps = int(max(2, min(400, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    o.step(reps)
    dt = time.perf_counter() - t0
    return {"value": scn.n * reps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"512-particle (8^3) slice of the pour, same templates, {reps} steps, "
                      f"serial FP64 oracle in the reference's order, {dt:.1f} s"}


def copy_probe():
    import torch
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copy probe: 3 copies of 1 GiB")


def autotune(s, a):
    """Run the autotuner from the solver's current state (it returns the
    solver to that state) and report its choice."""
    if not a.autotune:
        return None
    t0 = time.perf_counter()
    c = s.autotune(trial_steps=20)
    for line in c["trials"]:
        log("  tune " + line)
    log(f"autotuned in {time.perf_counter() - t0:.1f} s: skin {c['skin']:.3e} m, cell "
        f"{c['cell_size']:.3e} m (0: the cell rule), sweep {c['sweep_waves']} waves/workgroup, "
        f"{c['sweep_waves_per_cu']} waves/CU, {c['nodes_per_lane']} node(s) per lane, law mode "
        f"{c['law_mode']}, law block {c['law_block']}, reduce block "
        f"{c['reduce_block']} ({c['ms_per_step']:.3f} ms/step in the trials)")
    return {k: (float(v) if not isinstance(v, int) else v) for k, v in c.items() if k != "trials"}


def run_steps(s, steps, warmup, graph):
    s.use_graph(bool(graph))
    s.step(1)
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    return time.perf_counter() - t0


def batch_main(a):
    """512 packings (log-uniform 500..20000 particles, 16 shared templates,
    per-packing kn, damping ratio and mu) advanced by one solver, against the
    same packings run one after another in solvers of their own."""
    import numpy as np
    from fundem_amd import Solver
    from fundem_amd import workloads as W
    t0 = time.perf_counter()
    scns = W.batch_packings()
    n_total = sum(sc.n for sc in scns)
    log(f"batch of {len(scns)} packings, {n_total} particles, set-up {time.perf_counter() - t0:.1f} s")
    s = Solver(0)
    s.set_capacity(8, 24)
    s.set_skin(a.skin)
    t0 = time.perf_counter()
    W.apply_batch(scns, s)
    log(f"batch uploaded and initialised in {time.perf_counter() - t0:.1f} s")
    s.use_graph(bool(a.graph))
    s.step(1 + a.warmup)
    tuning = autotune(s, a)
    b0, _, _ = s.list_builds()
    elapsed = run_steps(s, a.steps, 0, a.graph)
    if s.status() != 0:
        raise RuntimeError(f"capacity exceeded in the batch (status {s.status()})")
    batched = n_total * a.steps / elapsed
    b, gb, skin = s.list_builds()
    b -= b0
    log(f"batched: {1e3 * elapsed / a.steps:.3f} ms/step, {batched:.3e} particle-steps/s; "
        f"candidate pairs {s.pair_totals()}, {b} list builds")
    # stage timing in a separate eager pass
    s.use_graph(False)
    s.set_timing(True)
    s.step(min(a.stage_steps, 50))
    s.sync()
    for k, (ms, calls) in sorted(s.stage_times().items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:18s} {ms / max(calls, 1) * 1e3:9.1f} us/launch  {calls} launches (stage-timed pass)")
    s.set_timing(False)
    del s
    rng = np.random.default_rng(17)
    pick = sorted(rng.choice(len(scns), size=min(a.batch_sample, len(scns)), replace=False))
    seq_time, seq_work = 0.0, 0
    for k in pick:
        sc = scns[k]
        p = Solver(0)
        p.set_capacity(8, 24)
        p.set_skin(a.skin)
        W.apply(sc, p)
        t = run_steps(p, a.steps, a.warmup, a.graph)
        if p.status() != 0:
            raise RuntimeError(f"capacity exceeded in packing {k}")
        seq_time += t
        seq_work += sc.n * a.steps
        del p
    sequential = seq_work / seq_time
    log(f"sequential plain runs ({len(pick)} packings): {sequential:.3e} particle-steps/s")
    out = {
        "metric": "batched vs sequential plain particle-steps/s (512 packings, one GPU)",
        "value": batched, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(scns)} periodic packings of 500..20000 superellipsoids "
                               f"(log-uniform, {n_total} in total), 16 shared templates (32^3 "
                               "SDF, 1000 nodes), per-packing kn in [5e3, 2e4], damping ratio "
                               "in [0.1, 0.5], mu in [0.2, 0.8], dt 5e-7 s",
                   "graph": bool(a.graph), "skin_m": skin, "list_builds": b, "tuning": tuning},
        "sequential": {"value": sequential, "unit": "particle-steps/s",
                       "sample": f"{len(pick)} of the {len(scns)} packings (seeded choice), each "
                                 f"in a solver of its own, {a.steps} steps each, one after another"},
        "speedup": batched / sequential,
    }
    return out


def column_main(a):
    """The differentiated column collapse: one adjoint step (the segment's
    recomputation, its contact tape and the reverse step) against one plain
    forward step, both eager, from a column packed by a settling run."""
    import dataclasses
    import numpy as np
    from fundem_amd import Adjoint, Solver
    from fundem_amd import workloads as W
    sc = W.column()
    s0 = Solver(0)
    W.apply(sc, s0)
    t0 = time.perf_counter()
    s0.step(a.settle)
    st = s0.state()
    log(f"column of {sc.n} bodies settled for {a.settle} steps in {time.perf_counter() - t0:.1f} s; "
        f"contacts {[int(s0.contacts(k)['A'].size) for k in (0, 1)]}")
    del s0
    sc = dataclasses.replace(sc, x=st["x"], q=st["q"], v=st["v"], w=st["w"])
    steps = a.steps
    f = Solver(0)
    W.apply(sc, f)
    fwd = run_steps(f, steps, a.warmup, False) / steps
    contacts = [int(f.contacts(k)["A"].size) for k in (0, 1)]
    del f
    fwd_graph = None
    if a.graph_compare:
        f = Solver(0)
        W.apply(sc, f)
        fwd_graph = run_steps(f, steps, a.warmup, True) / steps
        del f
    g = Solver(0)
    W.apply(sc, g)
    adj = Adjoint(g, contacts_per_particle=16)
    g.sync()
    t0 = time.perf_counter()
    adj.forward(steps, a.checkpoint)
    g.sync()
    t1 = time.perf_counter()
    fin = g.state()
    masses = []
    for i in range(sc.n):
        gi = int(sc.geo[i])
        masses.append(0.0 if gi in sc.fixed_templates else 1.0)
    vbar = np.asarray(masses)[:, None] * fin["v"]
    t2 = time.perf_counter()
    grad = adj.backward(None, None, vbar, None)
    t3 = time.perf_counter()
    fwd_ckpt = (t1 - t0) / steps
    back = (t3 - t2) / steps
    graph_note = "" if fwd_graph is None else f" (captured graphs {1e3 * fwd_graph:.3f})"
    log(f"forward {1e3 * fwd:.3f} ms/step{graph_note}, forward with "
        f"checkpoints {1e3 * fwd_ckpt:.3f} ms/step, "
        f"adjoint {1e3 * back:.3f} ms/step (ratio {back / fwd:.2f}); contacts at the last step "
        f"{contacts}; dJ/dkn {grad['kn'][0]:.4e} dJ/dzeta {grad['zeta'][0]:.4e} dJ/dmu {grad['mu'][0]:.4e}")
    out = {
        "metric": "adjoint step cost / forward step cost (column collapse, one GPU)",
        "value": back / fwd, "unit": "x", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * back, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"column of {sc.n - len([1 for m in masses if m == 0.0])} "
                               "superellipsoids (8 templates, 32^3 SDF, 800 nodes) on a rough floor "
                               f"({len([1 for m in masses if m == 0.0])} fixed bodies), packed by "
                               f"{a.settle} settling steps, dt 1e-6 s, kn 1e4, zeta 0.3, mu 0.5; "
                               f"checkpoint every {a.checkpoint} steps; J = sum of |v|^2 / 2 over the moving bodies",
                   "forward_ms_per_step": 1e3 * fwd,
                   "forward_graph_ms_per_step": None if fwd_graph is None else 1e3 * fwd_graph,
                   "forward_with_checkpoints_ms_per_step":
                   1e3 * fwd_ckpt, "contacts_last_step": contacts},
    }
    return out


def meshsdf_main(a):
    """Level-set grid of a closed triangle mesh (TriangleMeshParticle::
    buildGridByResolution at its default 50 nodes per diameter): a torus of
    5120 triangles (about the reference tutorial's bunny, 4968), timed as whole
    grid builds (triangle upload, kernel, values back). The oracle's serial
    build of a 20-node-per-diameter grid of the same mesh is the CPU baseline."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from fundem_amd import generators as G
    v, f = G.torus_mesh(1.0, 0.4, 80, 32)
    m = G.MeshParticle(v, f)
    for _ in range(max(1, a.warmup)):
        g = m.build_grid_by_resolution(50, 2)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g = m.build_grid_by_resolution(50, 2)
    dt = (time.perf_counter() - t0) / a.steps
    nodes = g[3].size
    pairs = nodes * len(f)
    ext = (v.max(axis=0) - v.min(axis=0)).max()
    c0 = time.perf_counter()
    ref = po.mesh_grid(v, f, ext / 20, 2)
    cdt = time.perf_counter() - c0
    cpairs = ref[2].size * len(f)
    sample = m.build_grid(ext / 20, 2)
    assert np.array_equal(sample[3], ref[2]), "device grid differs from the oracle's"
    log(f"mesh grid {g[2]} ({nodes} nodes) x {len(f)} triangles: {1e3 * dt:.2f} ms per grid; "
        f"oracle {ref[1]} grid in {cdt:.2f} s")
    out = {
        "metric": "mesh level-set grid: node-triangle evaluations per second (one GPU)",
        "value": pairs / dt, "unit": "node-triangle pairs/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"closed torus mesh R 1, r 0.4, {len(v)} vertices, {len(f)} triangles; "
                               f"grid of 50 nodes per diameter, 2 padding layers: {g[2]} = {nodes} "
                               "nodes; one step = one whole grid build"},
        "cpu_baseline": {"value": cpairs / cdt, "unit": "node-triangle pairs/s", "cores": 1,
                         "kind": "port", "sample": f"oracle grid of 20 nodes per diameter, "
                                                   f"{ref[1]} = {ref[2].size} nodes, {cdt:.1f} s"},
    }
    print(json.dumps(out), flush=True)


def pour_main(a):
    """The 1M-particle pour: the headline line (value, roofline of the
    dominant stage, per-stage rooflines, whole-step HBM fraction)."""
    from fundem_amd import Solver
    from fundem_amd import workloads as W

    t0 = time.perf_counter()
    scn = W.pour((a.side, a.side, a.side))
    log(f"workload {scn.name}: {scn.n} particles, setup {time.perf_counter() - t0:.1f} s")
    s = Solver(0)
    s.set_capacity(8, 24)
    s.set_skin(a.skin)
    t0 = time.perf_counter()
    W.apply(scn, s)
    log(f"uploaded and initialised in {time.perf_counter() - t0:.1f} s")
    s.use_graph(bool(a.graph))
    s.step(1)
    tuning = autotune(s, a)
    # warm-up after tuning: the tuner leaves the solver at its saved point
    # with a list build requested, which the warm-up consumes
    s.step(max(a.warmup, 1))
    s.sync()
    if s.status() != 0:
        raise RuntimeError(f"capacity exceeded during warm-up (status {s.status()})")
    b0, gb0, skin = s.list_builds()
    t0 = time.perf_counter()
    s.step(a.steps)
    s.sync()
    elapsed = time.perf_counter() - t0
    b1, gb1, _ = s.list_builds()
    status = s.status()
    if status != 0:
        raise RuntimeError(f"capacity exceeded during the timed steps (status {status})")
    stages, sb0, sb1 = stage_pass(s, a.stage_steps)
    contacts = {k: int(s.contacts(k)["A"].size) for k in range(5)}
    pairs = dict(enumerate(s.pair_totals()))
    n = scn.n
    ms_step = 1e3 * elapsed / a.steps
    value = n * a.steps / elapsed
    log(f"{ms_step:.3f} ms/step, {value:.3e} particle-steps/s; contacts at the last step "
        f"{contacts}, candidate pairs {pairs}; skin {skin:.3e} m, {b1 - b0} list builds and "
        f"{gb1 - gb0} ghost-list builds in {a.steps} steps")
    log_stages(stages, a.stage_steps)
    roof, stage_roof = rooflines(stages, n, pairs, contacts, spin_fixed(s, scn.geo))
    # algorithmic bytes of a whole step: state 224 B per particle and 20 B per
    # contact record read and written, plus the template tables
    c_total = sum(contacts.values())
    step_bytes = 224.0 * n + 2 * 20.0 * c_total + 16 * (32 ** 3 * 4 + 1000 * 12)
    rebuild = rebuild_cost(stages, a.stage_steps, sb1 - sb0)
    out = {
        "metric": "particle-steps/s (1M level-set particle pour, one GPU)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"pour of {n} superellipsoids (16 templates, 32^3 FP32 SDF, 1000 "
                               "nodes), periodic x/y + floor, cell 2.5 mm, dt 5e-7 s; FP64 "
                               "positions, FP32 v/w/q/F/T, FP64 arithmetic",
                   "particles": n, "contacts_last_step": contacts, "candidate_pairs": pairs,
                   "graph": bool(a.graph), "skin_m": skin, "list_builds": b1 - b0,
                   "ghost_list_builds": gb1 - gb0, "tuning": tuning},
        "step_hbm_fraction": step_bytes * a.steps / elapsed / (HBM_PEAK_GBS * 1e9),
        "roofline": roof,
        "stage_roofline": stage_roof,
        "list_rebuild": rebuild,
    }
    s.close()
    return out, scn.templates


def stage_pass(s, steps):
    """Stage timing (HIP events around every stage on the solver's stream)
    in a separate eager pass, so that the timed steps carry no
    instrumentation. Returns the stage times and the list-build counters
    around the pass."""
    s.use_graph(False)
    s.set_timing(True)
    b0, _, _ = s.list_builds()
    s.step(steps)
    s.sync()
    b1, _, _ = s.list_builds()
    stages = s.stage_times()
    s.set_timing(False)
    return stages, b0, b1


def log_stages(stages, steps):
    for k, (ms, calls) in sorted(stages.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:18s} {ms / max(calls, 1) * 1e3:9.1f} us/launch  {calls} launches  "
            f"{ms / steps * 1e3:8.1f} us/step (stage-timed pass)")


REBUILD_STAGES = ("sort_cells", "broad_pp", "broad_pw", "broad_ghost", "periodic_build")


def rebuild_cost(stages, steps, builds):
    """Time of the list-build stages in the stage-timed pass: per build and
    amortised per step."""
    ms = sum(stages[k][0] for k in REBUILD_STAGES if k in stages)
    return {"builds": builds, "steps": steps, "us_per_build": 1e3 * ms / builds if builds else None,
            "us_per_step_amortised": 1e3 * ms / steps}


def spin_fixed(s, geo):
    """Bodies whose template's inverse inertia is exactly zero (the solver's
    own mass properties: a millimetre template's inertia determinant falls
    below the inversion threshold)."""
    import numpy as np
    geo = np.asarray(geo)
    fixed = [g for g in np.unique(geo) if not np.any(s.geometry(int(g))["inv_inertia"])]
    return int(np.isin(geo, fixed).sum())


def rooflines(stages, n, pairs, contacts, fixed=0, counters=True):
    """Roofline of the dominant modelled stage (by total time) and of every
    modelled stage: algorithmic bytes per launch over the average launch.
    `counters`: attach the pour's measured traffic (profiles/traffic.json);
    other workloads get null."""
    kb = lambda k: kernel_bytes(k, n, pairs, contacts, fixed)
    traffic = measured_traffic if counters else (lambda k: None)
    modeled = {k: v for k, v in stages.items() if kb(k) is not None}
    stage_roof = []
    for k, (ms, calls) in sorted(modeled.items(), key=lambda kv: -kv[1][0]):
        nb = kb(k)
        avg_s = ms / max(calls, 1) * 1e-3
        ach = nb / avg_s / 1e9 if avg_s > 0 else 0.0
        stage_roof.append({"stage": k, "avg_launch_us": avg_s * 1e6, "bytes_per_launch": nb,
                           "achieved": ach, "frac": ach / HBM_PEAK_GBS,
                           "traffic": traffic(k)})
    roof = None
    if modeled:
        dom, (ms, calls) = max(modeled.items(), key=lambda kv: kv[1][0])
        nb = kb(dom)
        avg_s = ms / max(calls, 1) * 1e-3
        achieved = nb / avg_s / 1e9 if avg_s > 0 else 0.0
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic(dom), "kernel": dom,
                "avg_launch_us": avg_s * 1e6, "bytes_per_launch": nb}
    return roof, stage_roof


def dense_main(a, templates):
    """The pour's lattice driven into a contact-dense state: every particle
    starts falling at a speed proportional to its height above the lowest,
    so the layers close in on one another within about --dense-closing
    seconds and collide at about (lattice spacing) / closing; once the
    particle-particle node contacts stop growing (checked every 200 steps),
    that state is restarted at rest (v = w = 0) and timed like the pour."""
    from fundem_amd import Solver
    from fundem_amd import workloads as W
    scn = W.pour((a.side, a.side, a.side), templates=templates)

    def make_solver():
        s = Solver(0)
        s.set_capacity(12, 64)
        return s

    t0 = time.perf_counter()
    dense, done, c = W.compressed_pour(scn, make_solver, a.dense_closing, a.dense_max_steps)
    log(f"dense: compressed for {done} steps in {time.perf_counter() - t0:.1f} s, {c} "
        "particle-particle node contacts")
    s = Solver(0)
    s.set_capacity(12, 64)
    s.set_skin(a.skin)
    W.apply(dense, s)
    s.step(1)
    tuning = autotune(s, a)
    s.step(max(a.warmup, 1))
    s.sync()
    steps = max(a.steps, 50)
    b0, gb0, skin = s.list_builds()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    elapsed = time.perf_counter() - t0
    b1, gb1, _ = s.list_builds()
    if s.status() != 0:
        raise RuntimeError(f"capacity exceeded in the dense state (status {s.status()})")
    stages, sb0, sb1 = stage_pass(s, a.stage_steps)
    contacts = {k: int(s.contacts(k)["A"].size) for k in range(5)}
    pairs = dict(enumerate(s.pair_totals()))
    n = dense.n
    log(f"dense: {1e3 * elapsed / steps:.3f} ms/step, {n * steps / elapsed:.3e} particle-steps/s; "
        f"contacts {contacts}, candidate pairs {pairs}, {b1 - b0} list builds in {steps} steps")
    log_stages(stages, a.stage_steps)
    roof, stage_roof = rooflines(stages, n, pairs, contacts, spin_fixed(s, dense.geo), counters=False)
    s.close()
    return {"metric": "particle-steps/s (1M pour lattice compressed until its layers touch)",
            "value": n * steps / elapsed, "unit": "particle-steps/s", "steps": steps,
            "ms_per_step": 1e3 * elapsed / steps,
            "state": f"layers closed in at (spacing)/{a.dense_closing} s for {done} steps, then "
                     "restarted at rest",
            "contacts_last_step": contacts, "candidate_pairs": pairs, "skin_m": skin,
            "list_builds": b1 - b0, "ghost_list_builds": gb1 - gb0, "tuning": tuning,
            "list_rebuild": rebuild_cost(stages, a.stage_steps, sb1 - sb0),
            "roofline": roof, "stage_roofline": stage_roof}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the one-GPU build "
              f"does not shard)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "meshsdf":
        meshsdf_main(a)
        return
    if a.workload == "column":
        print(json.dumps(column_main(a)), flush=True)
        return
    if a.workload == "batch":
        print(json.dumps(batch_main(a)), flush=True)
        return
    if a.workload == "dense":
        print(json.dumps(dense_main(a, None)), flush=True)
        return
    if a.copy_probe:
        copy_probe()
    out, templates = pour_main(a)
    if a.extras:
        # the batch speedup, the adjoint's cost and the contact-dense state,
        # as extra keys of the same line
        out["dense"] = dense_main(a, templates)
        ab = argparse.Namespace(**vars(a))
        ab.steps = max(a.steps, 100)
        batch = batch_main(ab)
        out["batch"] = batch
        out["batch_speedup"] = batch["speedup"]
        ac = argparse.Namespace(**vars(a))
        ac.steps, ac.checkpoint, ac.warmup = 2000, 100, max(a.warmup, 1)
        col = column_main(ac)
        out["adjoint"] = col
        out["adjoint_over_forward"] = col["value"]
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(templates, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
