"""A compiled application of the C++ host API, written as the reference's
tutorials are (tests/apps/tutorial_app.cpp: an LSSolver subclass whose
addExternalForceTorque launches the tutorials' global damping on
getLSParticle().force()/torque()/velocity()/angularVelocity() with its
num_device(), gridDim() and blockDim(); particles added by
addParticleCluster; solve() with output frames), run on the GPU: its final
state equals the oracle's bit for bit after the same steps (the pre-loop build
and the numStep + 1 iterations of LSSolver::solve, LSSolver.h:440-467), its
launch shapes follow LSParticle.h:379-381, and the contact copies of
SolidInteraction's HostRef getters equal the oracle's contacts."""
import os
import struct
import subprocess

import numpy as np
import pytest

import pyoracle as po
from fundem_amd import build
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

DAMPING = 0.2
MAX_GPU_THREAD = 256  # LSSolver's default maxGPUThread


def scene():
    """Centimetre superellipsoids (rotation live) of three templates thrown
    at a floor; one material; particles ordered by template, one cluster each."""
    scn = W.cluster(n=27, speed=1.0, scale=1e-2, dt=1e-5, kn=1e5, spin=5.0)
    order = np.argsort(scn.geo, kind="stable")
    return scn, order


def write_grid(f, t):
    f.write(struct.pack("<3dd3i", *np.asarray(t.origin, np.float64), float(t.spacing),
                        *[int(s) for s in t.size]))
    f.write(np.asarray(t.sdf, np.float64).astype("<f8").tobytes())
    f.write(struct.pack("<i", len(t.nodes)))
    f.write(np.asarray(t.nodes, np.float64).astype("<f8").tobytes())
    tris = np.asarray(t.tris, np.int32).reshape(-1, 3)
    f.write(struct.pack("<i", len(tris)))
    f.write(tris.astype("<i4").tobytes())


def write_input(path, scn, order, max_time, num_frame):
    kn, ks, mu, e = scn.materials[0]
    wmu, we = scn.wall_materials[0]
    with open(path, "wb") as f:
        f.write(struct.pack("<dd", DAMPING, 2500.0))
        f.write(struct.pack("<4d", kn, ks, mu, e))
        f.write(struct.pack("<2d", wmu, we))
        f.write(struct.pack("<i", len(scn.templates)))
        for g, t in enumerate(scn.templates):
            write_grid(f, t)
            idx = order[scn.geo[order] == g]
            f.write(struct.pack("<i", len(idx)))
            for i in idx:
                f.write(np.concatenate([scn.x[i], scn.v[i], scn.w[i], scn.q[i]]).astype("<f8").tobytes())
        write_grid(f, scn.wall_templates[0])
        f.write(np.asarray(scn.walls[0][0], "<f8").tobytes())
        f.write(np.concatenate([scn.domain_min, scn.domain_max, scn.gravity]).astype("<f8").tobytes())
        f.write(struct.pack("<ddi", scn.dt, max_time, num_frame))


def test_tutorial_app_solve_equals_oracle(tmp_path):
    assert os.path.exists(build.APP_BIN), "test application not built"
    scn, order = scene()
    n = scn.n
    max_time, num_frame = 2e-3, 4
    inp, out, frames = tmp_path / "scene.bin", tmp_path / "state.bin", tmp_path / "frames"
    write_input(inp, scn, order, max_time, num_frame)
    print(f"tutorial app: {n} particles, maxTime {max_time}, dt {scn.dt}", flush=True)
    r = subprocess.run([build.APP_BIN, str(inp), str(out), str(frames)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Computation Completed" in r.stdout
    data = np.frombuffer(out.read_bytes(), "<f8")
    x0 = data[:3 * n].reshape(n, 3)
    fin = data[3 * n:3 * n + 19 * n].reshape(n, 19)
    rest = data[22 * n:]
    # the oracle, fed the same clusters (geometry per template, density path)
    # at the positions addParticleCluster gave them
    o = po.OracleSim(True, True)
    kn, ks, mu, e = scn.materials[0]
    o.add_material(kn, ks, mu, e)
    for t in scn.templates:
        o.add_geometry_density(2500.0, t)
    o.add_wall_material(*scn.wall_materials[0])
    o.add_wall_geometry(scn.wall_templates[0])
    o.add_particles(x0, scn.v[order], scn.w[order], scn.q[order], np.zeros(n, np.int32),
                    scn.geo[order])
    wx, wv, ww, wq, wm, wg = scn.walls[0]
    o.add_wall(wx, wv, ww, wq, wm, wg)
    o.set_global_damping(DAMPING)
    o.initialize(scn.domain_min, scn.domain_max, scn.gravity, scn.dt)
    num_step = int(max_time / scn.dt) + 1
    o.step(num_step + 1)
    os_ = o.state()
    got = dict(x=fin[:, 0:3], v=fin[:, 3:6], w=fin[:, 6:9], q=fin[:, 9:13], F=fin[:, 13:16],
               T=fin[:, 16:19])
    for key in ("x", "v", "w", "q", "F", "T"):
        assert np.array_equal(got[key], os_[key]), f"{key} differs from the oracle"
    # launch shapes (LSParticle.h:379-381)
    block = min(MAX_GPU_THREAD, n)
    assert rest[0] == (n + block - 1) // block and rest[1] == block
    # SolidInteraction's host copies after the last frame
    pos = 2
    node_base = np.concatenate([[0], np.cumsum([len(scn.templates[g].nodes) for g in scn.geo[order]])])
    seen = 0
    for s in (po.SET_PP, po.SET_PW):
        m = int(rest[pos])
        gd, bd = rest[pos + 1], rest[pos + 2]
        rec = rest[pos + 3:pos + 3 + 8 * m].reshape(m, 8)
        pos += 3 + 8 * m
        oc = o.contacts(s)
        assert m == oc["A"].size, s
        assert np.array_equal(rec[:, 0].astype(np.int64), node_base[oc["A"]] + oc["k"]), s
        assert np.array_equal(rec[:, 1].astype(np.int64), oc["B"]), s
        np.testing.assert_allclose(rec[:, 2:5], oc["point"], rtol=0, atol=1e-15)
        np.testing.assert_allclose(rec[:, 5:8], oc["normal"], rtol=0, atol=1e-12)
        pb = min(MAX_GPU_THREAD, m) if m else 1
        assert bd == pb and gd == ((m + pb - 1) // pb if m else 1), s
        seen += m
    assert pos == rest.size
    assert seen > 0
    # the frames solve() wrote
    vtus = list((tmp_path / "frames_phase1").rglob("*.vtu"))
    assert len(vtus) >= 5 * (num_frame + 1), len(vtus)
