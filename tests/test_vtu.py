"""The .vtu writer (fundem_amd/include/fundem/VTUWriter.h), the on-disk
format of the reference's output frames (LSParticle.h:426-756,
SolidInteraction.h:153-316, VBondedInteraction.h:262-448): a host program
writes known arrays, which are decoded here with Python's base64 and XML
modules (tests/vtu_read.py) and compared value for value; every binary block
must equal base64([UInt64 byte count][bytes]) as Python encodes it. Also the
surface triangulations the generators give templates for the mesh frames:
closed (Euler characteristic 2, every edge shared by two triangles) and
wound outward."""
import base64
import os
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import vtu_read
from fundem_amd import generators as G

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("vtu")
    exe = d / "vtu_writer_check"
    subprocess.run([cxx, "-O1", "-std=c++17", os.path.join(HERE, "vtu_writer_check.cpp"), "-o",
                    str(exe)], check=True)
    out = d / "frames" / "sub"
    r = subprocess.run([str(exe), str(out), "x"], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return out


def _known():
    N = 4
    i3 = np.arange(3 * N)
    pts = (np.float32(0.25) * i3.astype(np.float32) - np.float32(1.0)).astype(np.float32)
    vel = (np.float32(1e-3) * (i3 * i3).astype(np.float32)).astype(np.float32)
    q = (np.arange(4 * N).astype(np.float32) / np.float32(7.0)).astype(np.float32)
    e = (np.arange(N).astype(np.float32) * np.float32(1.5e-9)).astype(np.float32)
    pid = (100 - 3 * np.arange(N)).astype(np.int32)
    return pts.reshape(-1, 3), vel.reshape(-1, 3), q.reshape(-1, 4), e, pid


def test_mesh_piece_round_trip(written):
    f = vtu_read.read(written / "Mesh_0003.vtu")
    pts, vel, q, e, pid = _known()
    assert f["time"] == 1.25 and f["step"] == 7
    assert f["n_points"] == 4 and f["n_cells"] == 2 and f["vectors"] == "velocity"
    assert np.array_equal(f["points"], pts)
    assert np.array_equal(f["cells"]["connectivity"], [0, 1, 2, 1, 2, 3])
    assert np.array_equal(f["cells"]["offsets"], [3, 6])
    assert np.array_equal(f["cells"]["types"], [5, 5])
    pd = f["point_data"]
    assert list(pd) == ["particleID", "velocity", "orientation", "energy"]
    assert np.array_equal(pd["particleID"], pid)
    assert np.array_equal(pd["velocity"], vel)
    assert np.array_equal(pd["orientation"], q)
    assert np.array_equal(pd["energy"], e)


def test_binary_blocks_are_standard_base64(written):
    pts, vel, q, e, pid = _known()
    expect = {
        None: pts, "particleID": pid, "velocity": vel, "orientation": q, "energy": e,
        "connectivity": np.array([0, 1, 2, 1, 2, 3], np.int32),
        "offsets": np.array([3, 6], np.int32), "types": np.array([5, 5], np.uint8),
    }
    root = ET.parse(written / "Mesh_0003.vtu").getroot()
    pads = set()
    seen = 0
    for el in root.iter("DataArray"):
        if el.get("format") != "binary":
            continue
        raw = np.ascontiguousarray(expect[el.get("Name")]).tobytes()
        ref = base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode()
        assert el.text.strip() == ref, el.get("Name")
        pads.add(len(ref) - len(ref.rstrip("=")))
        seen += 1
    assert seen == 8 and pads == {0, 1, 2}, pads


def test_points_and_empty_pieces(written):
    f = vtu_read.read(written / "Points_0012.vtu")
    pts, vel, _, _, _ = _known()
    assert f["time"] == 0.5 and f["step"] == 11 and f["n_cells"] == 0
    assert all(a.size == 0 for a in f["cells"].values())
    assert f["vectors"] is None
    assert np.array_equal(f["points"], pts) and np.array_equal(f["point_data"]["velocity"], vel)
    z = vtu_read.read(written / "Empty_0000.vtu")
    assert z["n_points"] == 0 and z["points"].shape == (0, 3)
    assert z["point_data"]["contactNormal"].size == 0


@pytest.mark.parametrize("make", [
    lambda: G.sphere_template(1e-3, n_nodes=400),
    lambda: G.superellipsoid_template(1.0, 0.8, 0.6, 4.0, 3.0, n_grid=24, n_nodes=500),
    lambda: G.plane_wall([-1, -1], [2, 1]),
])
def test_template_triangulations_closed_and_outward(make):
    t = make()
    tri = t.tris
    assert tri is not None and tri.dtype == np.int32 and tri.min() >= 0
    assert tri.max() < len(t.nodes)
    nodes = t.nodes.astype(np.float64)
    a, b, c = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
    nrm = np.cross(b - a, c - a)
    if len(tri) == 2:  # the plane wall: two triangles facing +z, covering its rectangle
        assert np.all(nrm[:, 2] > 0)
        area = 0.5 * np.linalg.norm(nrm, axis=1).sum()
        assert np.isclose(area, 3.0 * 2.0)
        return
    # closed surface of genus 0: V - E + F = 2, each edge in exactly two triangles
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert len(nodes) - len(counts) + len(tri) == 2
    # outward: the signed volume (divergence theorem) equals the enclosed volume > 0
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    assert vol > 0
    assert np.all(np.einsum("ij,ij->i", nrm, a + b + c) > 0)
