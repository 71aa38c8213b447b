"""Mesh level sets on the GPU (fdm_launchMeshSignedDistanceGrid / Points
behind LevelSetObject::TriangleMeshParticle) against the oracle
(oracle/meshsdf.cpp): the same grid placement and the same values bit for
bit, for a box, an icosphere and a torus (not convex), with and without
positive_inside, plus point queries; the grid file of outputGridVTU; and a
mesh template (vertices as surface nodes) in a full simulation, bit-exact
against the oracle's step."""
import numpy as np
import pytest

import pyoracle as po
import vtu_read
from fundem_amd import generators as G
from fundem_amd import workloads as W

pytestmark = pytest.mark.gpu

MESHES = {
    "box": lambda: G.box_mesh(1.0, 2.0, 3.0),
    "icosphere": lambda: G.icosphere(1.3, 3),
    "torus": lambda: G.torus_mesh(),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("positive_inside", [False, True])
def test_grid_equals_oracle(name, positive_inside):
    v, f = MESHES[name]()
    v = v + np.array([0.013, -0.021, 0.007])
    m = G.MeshParticle(v, f, positive_inside=positive_inside)
    assert m.valid
    for res in (20, 37):
        origin, h, size, vals = m.build_grid_by_resolution(res, 2)
        ext = (v.max(axis=0) - v.min(axis=0)).max()
        o2, s2, v2 = po.mesh_grid(v, f, ext / res, 2, positive_inside)
        assert np.array_equal(origin, o2) and size == s2 and h == ext / res
        bad = np.flatnonzero(vals != v2)
        assert bad.size == 0, (name, res, bad[:10], vals[bad[:10]], v2[bad[:10]])
    rng = np.random.default_rng(5)
    p = rng.uniform(v.min(axis=0) - 0.5, v.max(axis=0) + 0.5, size=(3000, 3))
    val, grad = m.evaluate(p)
    val2, grad2 = po.mesh_points(v, f, p, positive_inside)
    assert np.array_equal(val, val2) and np.array_equal(grad, grad2)


def test_invalid_mesh_and_grid_file(tmp_path):
    v, f = G.icosphere(1.0, 1)
    bad = f.copy()
    bad[0, 2] = bad[0, 1]
    m = G.MeshParticle(v, bad)
    assert not m.valid and m.build_grid(0.2) is None
    m = G.MeshParticle(v, f)
    origin, h, size, vals = m.build_grid(0.25)
    m.output_grid_vtu(str(tmp_path / "sphere"))
    g = vtu_read.read(tmp_path / "sphereLSGrid.vtu")
    assert g["n_points"] == vals.size
    assert np.array_equal(g["point_data"]["signedDistance"], vals.astype(np.float32))
    ijk = g["point_data"]["gridIJK"]
    np.testing.assert_allclose(g["points"], (origin + h * ijk).astype(np.float32), rtol=1e-6)
    assert np.array_equal(ijk[1], [1, 0, 0]) and np.array_equal(ijk[size[0]], [0, 1, 0])


def test_mesh_template_particles_bitexact():
    """Tori built from their mesh (SDF on the GPU, vertices as nodes) dropped
    as a small cluster onto a floor: device and oracle steps agree bit for bit."""
    from fundem_amd import Solver
    v, f = G.torus_mesh(1.0e-3, 0.4e-3, 32, 16)
    t = G.mesh_template(v, f, nodes_per_diameter=24)
    assert t.tris is not None and len(t.nodes) == len(v)
    scn = W.cluster(n=8, speed=1.0, templates=[t], spacing=2.6, steps=300)
    scn.x[:, 2] -= scn.x[:, 2].min() - 1.0e-3  # the lower layer reaches the floor
    g = Solver()
    W.apply(scn, g)
    o = po.OracleSim(True, True)  # FP32 state storage and canonical fold, as the device
    W.apply(scn, o)
    contacts = 0
    for _ in range(30):
        g.step(10)
        o.step(10)
        contacts += g.contacts(po.SET_PP)["A"].size + g.contacts(po.SET_PW)["A"].size
        a, b = g.state(), o.state()
        for k in ("x", "v", "w", "q"):
            assert np.array_equal(a[k], b[k]), k
    assert contacts > 0
