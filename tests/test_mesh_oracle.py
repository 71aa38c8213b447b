"""The mesh level-set oracle (oracle/meshsdf.cpp, restating
LevelSetObjectGenerator.h:122-224,454-914) pinned by what the geometry
fixes, and the native OBJ loader (OBJLoader.h) on files written here:
  * an axis-aligned box mesh gives the box's exact signed distance and the
    outward normal of the nearest face as gradient;
  * an icosphere lies between its inscribed and circumscribed spheres, so its
    distance is bounded by theirs and its sign is fixed outside that shell;
  * the closest point of a single triangle is the projection onto a convex
    set: inside the triangle, with (p - c).(x - c) <= 0 for every vertex x;
  * the solid-angle inside test on a torus (not convex, genus 1) agrees with
    the torus's own membership away from its surface;
  * grid placement: origin on a multiple of the spacing, at least `padding`
    (at least one) spacings outside the mesh box on every side, less than
    one more; sign flip with positive_inside; no grid for an invalid mesh."""
import math

import numpy as np
import pytest

import pyoracle as po
from fundem_amd import generators as G

torus_mesh = G.torus_mesh


def box_sdf(p, half):
    q = np.abs(p) - half
    return np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)


def test_box_mesh_is_the_exact_box_distance():
    half = np.array([0.5, 1.0, 1.5])
    v, f = G.box_mesh(*(2 * half))
    rng = np.random.default_rng(1)
    p = rng.uniform(-2.5, 2.5, size=(4000, 3))
    val, grad = po.mesh_points(v, f, p)
    np.testing.assert_allclose(val, box_sdf(p, half), rtol=0, atol=1e-13)
    # gradient: outside, (p - c)/|p - c|; inside, the nearest face's normal
    q = np.abs(p) - half
    out = q.max(axis=1) > 0
    c = np.clip(p, -half, half)
    gout = (p - c)[out] / np.linalg.norm((p - c)[out], axis=1)[:, None]
    np.testing.assert_allclose(grad[out], gout, rtol=0, atol=1e-12)
    srt = np.sort(q, axis=1)
    clear = ~out & (srt[:, 2] - srt[:, 1] > 1e-6)  # one nearest face
    ax = np.argmax(q, axis=1)
    gin = np.zeros_like(p)
    gin[np.arange(len(p)), ax] = np.sign(p[np.arange(len(p)), ax])
    np.testing.assert_allclose(grad[clear], gin[clear], rtol=0, atol=1e-12)


def test_icosphere_between_inscribed_and_circumscribed_spheres():
    v, f = G.icosphere(1.0, 2)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(b - a, c - a)
    r_in = np.min(np.abs(np.einsum("ij,ij->i", n / np.linalg.norm(n, axis=1)[:, None], a)))
    rng = np.random.default_rng(2)
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rad = rng.uniform(0.0, 2.0, size=3000)
    p = d * rad[:, None]
    val, _ = po.mesh_points(v, f, p)
    outside, inside = rad > 1.0 + 1e-9, rad < r_in - 1e-9
    assert np.all(val[outside] > 0) and np.all(val[inside] < 0)
    assert np.all(val[outside] >= rad[outside] - 1.0 - 1e-12)
    assert np.all(val[outside] <= rad[outside] - r_in + 1e-12)
    assert np.all(val[inside] >= -(1.0 - rad[inside]) - 1e-12)
    assert np.all(val[inside] <= -(r_in - rad[inside]) + 1e-12)


def test_single_triangle_closest_point_is_the_projection():
    rng = np.random.default_rng(3)
    for _ in range(40):
        tri = rng.normal(size=(3, 3))
        p = rng.normal(scale=2.0, size=(200, 3))
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n /= np.linalg.norm(n)
        p = p[np.abs((p - tri[0]) @ n) > 1e-3]  # off the plane: solid angle below 2 pi
        val, grad = po.mesh_points(tri, np.array([[0, 1, 2]]), p)
        assert np.all(val > 0)
        cp = p - val[:, None] * grad
        # inside the triangle (barycentric coordinates >= 0)
        T = np.stack([tri[1] - tri[0], tri[2] - tri[0]], 1)
        bc, *_ = np.linalg.lstsq(T, (cp - tri[0]).T, rcond=None)
        assert np.all(bc > -1e-9) and np.all(bc.sum(axis=0) < 1 + 1e-9)
        # variational inequality of the projection onto a convex set
        for x in tri:
            assert np.all(np.einsum("ij,ij->i", p - cp, x - cp) <= 1e-9)


def test_torus_inside_test_matches_membership():
    R, r = 1.0, 0.4
    v, f = torus_mesh(R, r)
    rng = np.random.default_rng(4)
    p = rng.uniform([-1.6, -1.6, -0.6], [1.6, 1.6, 0.6], size=(3000, 3))
    exact = np.hypot(np.hypot(p[:, 0], p[:, 1]) - R, p[:, 2]) - r
    far = np.abs(exact) > 0.02  # beyond the facets' deviation from the torus
    val, _ = po.mesh_points(v, f, p[far])
    assert np.array_equal(val < 0, exact[far] < 0)
    assert np.sum(exact[far] < 0) > 200  # both sides are exercised


@pytest.mark.parametrize("h,pad", [(0.1, 2), (0.137, 3), (0.25, 0)])
def test_grid_placement_and_values(h, pad):
    v, f = G.icosphere(0.8, 1)
    v = v + np.array([0.31, -0.2, 0.05])
    origin, size, vals = po.mesh_grid(v, f, h, pad)
    layers = max(1, pad)
    assert np.allclose(origin / h, np.round(origin / h), atol=1e-9)
    last = origin + (np.array(size) - 1) * h
    assert np.all(origin <= v.min(axis=0) - layers * h + 1e-12)
    assert np.all(origin > v.min(axis=0) - (layers + 1) * h - 1e-12)
    assert np.all(last >= v.max(axis=0) + layers * h - 1e-12)
    assert np.all(last < v.max(axis=0) + (layers + 1) * h + 1e-12)
    nx, ny, nz = size
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = origin + h * np.stack([i, j, k], -1).reshape(-1, 3)
    val, grad = po.mesh_points(v, f, pts)
    np.testing.assert_allclose(vals, val / np.linalg.norm(grad, axis=1), rtol=1e-15)
    flipped = po.mesh_grid(v, f, h, pad, positive_inside=True)[2]
    np.testing.assert_array_equal(flipped, -vals)


def test_invalid_meshes_have_no_grid():
    v, f = G.icosphere(1.0, 0)
    assert po.mesh_grid(v, f, 0.0) is None
    bad = f.copy()
    bad[3, 1] = bad[3, 0]
    assert po.mesh_grid(v, bad, 0.2) is None
    out = f.copy()
    out[0, 0] = len(v)
    assert po.mesh_grid(v, out, 0.2) is None
    val, grad = po.mesh_points(v, bad, np.zeros((2, 3)))
    assert np.all(val == 1.0) and np.all(grad == 0.0)


OBJ = """# a cube with quads, normals and texture records
o cube
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
vn 0 0 1
vt 0.5 0.5
v -1 -1 1
v 1 -1 1
   v 1 1 1   
v -1 1 1
f 1 4 3 2
f 5/1/1 6/1/1 7/1/1 8/1/1
g side
f 1//1 2//1 6//1
f -7 -2 -3
usemtl none
f 2 3 7 6
f 3/1 4/1 8/1 7/1
f 4 1 5 8
"""


def test_obj_loader_records_indices_and_fans(tmp_path):
    path = tmp_path / "cube.obj"
    path.write_text(OBJ)
    v, f = G.load_obj(path)
    assert v.shape == (8, 3) and f.dtype == np.int32
    assert np.array_equal(v[6], [1, 1, 1])
    expect = [(0, 3, 2), (0, 2, 1), (4, 5, 6), (4, 6, 7), (0, 1, 5), (1, 6, 5),
              (1, 2, 6), (1, 6, 5), (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]
    assert [tuple(t) for t in f] == expect


@pytest.mark.parametrize("body,what", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n", "zero index"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "out of range"),
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", "two vertices"),
    ("v 0 0\nf 1 2 3\n", "short vertex"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -4\n", "negative out of range"),
])
def test_obj_loader_rejects_malformed_files(tmp_path, body, what):
    path = tmp_path / "bad.obj"
    path.write_text(body)
    with pytest.raises(OSError):
        G.load_obj(path)


def test_obj_loader_missing_file(tmp_path):
    with pytest.raises(OSError):
        G.load_obj(tmp_path / "none.obj")
