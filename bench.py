"""Benchmark: the level-set DEM pour of 1,000,000 superellipsoids on one GPU.

Workload (synthetic, seeded): 16 superellipsoid templates (semi-axes
U[0.8, 1.2] mm, powers U[2, 8], 32^3 FP32 SDF, 1000 surface nodes) on a
100^3 jittered lattice, random orientations, box periodic in x and y, floor
wall, cell size 2.5 mm, kn = 1e4 N/m, mu = 0.5, dt = 5e-7 s.

One step is one iteration of LSSolver's loop (kick, drift, hash, sort,
broad/narrow phase and contact forces for particle, wall and ghost sets,
reduction, kick). The timed region covers exactly --steps steps with the
device synchronised on both sides; every stage is bracketed by HIP events on
the solver's stream, which give the per-stage device times used for the
roofline of the dominant stage. The CPU baseline is the serial FP64 oracle
on a 512-particle sample of the same workload.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--side", type=int, default=100, help="lattice side (side^3 particles)")
    p.add_argument("--graph", type=int, default=0, help="replay captured two-step graphs")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget")
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic bytes per launch of each stage (FP32 state, 32-bit ids):
#   integrate_hash  per particle: read x v w q F T (76) + template id (4),
#                   write x v w q (52) + hash, key, index (12)          = 144
#   sort_cells      per particle: 3 radix passes x (key read for the
#                   histogram 4 + key/index read 8 + write 8) + cell marks
#                   (sorted key read 4, start/end write 8, previous key
#                   write 4, previous-cell reset read 4 + write 8)       =  88
#   narrow_pp/pw/ghost  per contact: history read 16 + write 16; per pair:
#                   both poses (2 x 28) + pair sums (72) + pair record 20
#   reduce_kick     per particle: F T write (24) + v w read/write (48) +
#                   q read (16) + template id (4) + pair-sum reads
#                   (72 per owned or mirrored pair)
def stage_bytes(stage, n, pairs, contacts):
    if stage == "integrate_hash":
        return 144.0 * n
    if stage == "sort_cells":
        return 88.0 * n
    if stage.startswith("narrow"):
        return 32.0 * contacts + 148.0 * pairs
    if stage == "reduce_kick":
        return 92.0 * n + 144.0 * pairs
    if stage.startswith("broad"):
        return 20.0 * n + 12.0 * pairs
    return 0.0


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


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path shards by "
              f"slab, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from fundem_amd import Solver
    from fundem_amd import workloads as W

    t0 = time.perf_counter()
    scn = W.pour((a.side, a.side, a.side))
    log(f"workload {scn.name}: {scn.n} particles, setup {time.perf_counter() - t0:.1f} s")
    s = Solver(0)
    s.set_capacity(8, 24)
    t0 = time.perf_counter()
    W.apply(scn, s)
    log(f"uploaded and initialised in {time.perf_counter() - t0:.1f} s")
    s.use_graph(bool(a.graph))
    s.step(1)
    s.step(a.warmup)
    s.sync()
    if s.status() != 0:
        raise RuntimeError(f"capacity exceeded during warm-up (status {s.status()})")
    s.set_timing(not a.graph)
    s.sync()
    t0 = time.perf_counter()
    s.step(a.steps)
    s.sync()
    elapsed = time.perf_counter() - t0
    status = s.status()
    if status != 0:
        raise RuntimeError(f"capacity exceeded during the timed steps (status {status})")
    stages = s.stage_times() if not a.graph else {}
    cnt = {k: s.contacts(k)["A"].size for k in range(5)}
    contacts = sum(cnt.values())
    n = scn.n
    ms_step = 1e3 * elapsed / a.steps
    value = n * a.steps / elapsed
    log(f"{ms_step:.3f} ms/step, {value:.3e} particle-steps/s, contacts at last step {cnt}")
    for k, (ms, calls) in sorted(stages.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:16s} {ms / max(calls, 1) * 1e3:9.1f} us/launch  {calls} launches")

    roof = None
    if stages:
        dom, (ms, calls) = max(stages.items(), key=lambda kv: kv[1][0])
        pairs_est = contacts  # pairs with contacts are at most the contact count
        bytes_launch = stage_bytes(dom, n, pairs_est, contacts)
        avg_s = ms / max(calls, 1) * 1e-3
        achieved = bytes_launch / avg_s / 1e9 if avg_s > 0 else 0.0
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "avg_launch_us": avg_s * 1e6, "bytes_per_launch": bytes_launch}
    step_bytes = 224.0 * n + 2 * 20.0 * contacts + 16 * (32 ** 3 * 4 + 1000 * 12)
    out = {
        "metric": "particle-steps/s (1M level-set particle pour, one GPU)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"pour {n} superellipsoids (16 templates, 32^3 FP32 SDF, 1000 nodes), "
                               "periodic x/y + floor, cell 2.5 mm, dt 5e-7 s; FP32 state, FP64 arithmetic",
                   "particles": n, "contacts_last_step": cnt, "graph": bool(a.graph)},
        "step_hbm_fraction": step_bytes * a.steps / elapsed / (HBM_PEAK_GBS * 1e9),
        "roofline": roof,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scn.templates, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
