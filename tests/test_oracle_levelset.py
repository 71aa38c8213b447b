"""Pins of the oracle's level-set sampling (the trilinear value and the Q1
gradient of LSParticleContactDetectionKernel.cuh:4-77, sampled as
LSParticleContactDetectionKernel.cu:72-119 does) on curved fields, where an
affine field cannot tell a wrong gradient apart.

The value is checked against scipy's trilinear interpolant on the same nodes.
Inside a cell that interpolant is linear in each coordinate, so a central
difference gives its partial derivatives exactly up to rounding. The Q1
gradient is stated through it: d/dx at the point, but d/dy and d/dz on the
cell's lower x face (the reference truncates the fractional x to 0)."""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import pyoracle as O


def field(size, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = size
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    curved = 0.3 * (x - 2.1) ** 2 - 0.2 * (y - 3.3) * (z - 1.7) + 0.05 * x * y * z
    return (curved + rng.normal(0.0, 0.5, size=(nz, ny, nx))).reshape(-1)  # x fastest


def cell_points(size, origin, h, n, seed, margin=0.05):
    rng = np.random.default_rng(seed)
    cells = np.stack([rng.integers(0, s - 1, n) for s in size], axis=1)
    frac = rng.uniform(margin, 1.0 - margin, (n, 3))
    return cells, np.asarray(origin) + (cells + frac) * h


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_value_and_q1_gradient_on_a_curved_field(seed):
    size, origin, h = (6, 7, 8), np.array([-0.31, 0.12, -0.05]), 0.037
    sdf = field(size, seed)
    interp = RegularGridInterpolator(
        [origin[d] + h * np.arange(size[d]) for d in range(3)],
        sdf.reshape(size[2], size[1], size[0]).transpose(2, 1, 0), method="linear")
    cells, pts = cell_points(size, origin, h, 400, seed + 10)
    val, grad, inside = O.sample_level_set(sdf, size, origin, h, pts)
    assert inside.all()
    scale = np.abs(sdf).max()
    np.testing.assert_allclose(val, interp(pts), rtol=0, atol=1e-12 * scale)

    eps = 1e-4 * h
    def d(at, axis):
        e = np.zeros(3)
        e[axis] = eps
        return (interp(at + e) - interp(at - e)) / (2 * eps) * h  # per cell

    face = pts.copy()
    face[:, 0] = origin[0] + cells[:, 0] * h  # the cell's lower x face
    np.testing.assert_allclose(grad[:, 0], d(pts, 0), rtol=0, atol=1e-7 * scale)
    np.testing.assert_allclose(grad[:, 1], d(face, 1), rtol=0, atol=1e-7 * scale)
    np.testing.assert_allclose(grad[:, 2], d(face, 2), rtol=0, atol=1e-7 * scale)
    # the field is curved enough for Q1 to matter: the true y and z slopes at
    # the point differ from the face's
    assert np.abs(grad[:, 1] - d(pts, 1)).max() > 1e-2 * scale
    assert np.abs(grad[:, 2] - d(pts, 2)).max() > 1e-2 * scale


def test_points_outside_the_grid_cells_are_not_sampled():
    size, origin, h = (4, 4, 4), np.zeros(3), 0.5
    sdf = field(size, 3)
    top = origin + h * (np.array(size) - 1)
    pts = np.array([[-1e-9, 0.7, 0.7], [0.7, 0.7, top[2]], [0.7, top[1] + 0.1, 0.7],
                    [0.7, 0.7, 0.7], [top[0] - 1e-9, 0.2, 0.2]])
    val, grad, inside = O.sample_level_set(sdf, size, origin, h, pts)
    assert inside.tolist() == [False, False, False, True, True]
    assert (val[~inside] == 0).all() and (grad[~inside] == 0).all()
