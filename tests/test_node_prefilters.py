"""The narrow phase's plane-wall rejection before the node transform
(contact.hip, sweepPair), checked in FP64 with the device's rotation formula
(v' = v + q0 t + u x t, t = 2 u x v) and orientations stored in FP32 as the
solver stores them: the plane's value along the node coordinates,
c0 + u.s with g = R_T a, u = R_A^T g, c0 = b + g.(x_A - x_T), equals the
plane at the node's local point R_T^T (R_A s + x_A - x_T) to within a 1e-12
fraction of the values involved, far below the 1e-9 slack the test keeps;
the derivation uses that the rotation formula's transpose is the
conjugate's for any quaternion, unit or not."""
import numpy as np


def qrot(q, v):
    u = q[..., 1:]
    t = 2.0 * np.cross(u, v)
    return v + q[..., :1] * t + np.cross(u, t)


def qrotT(v, q):
    c = q.copy()
    c[..., 1:] *= -1.0
    return qrot(c, v)


def stored_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32).astype(np.float64)


def test_transpose_of_rotation_formula_is_the_conjugate():
    rng = np.random.default_rng(1)
    q = rng.normal(size=(1000, 4)) * rng.uniform(0.5, 2.0, size=(1000, 1))  # not unit
    a, b = rng.normal(size=(1000, 3)), rng.normal(size=(1000, 3))
    lhs = np.einsum("ij,ij->i", qrot(q, a), b)
    rhs = np.einsum("ij,ij->i", a, qrotT(b, q))
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_plane_prefilter_matches_the_plane_at_the_local_point():
    rng = np.random.default_rng(3)
    n = 200000
    qA = stored_quaternions(rng, n)
    qT = stored_quaternions(rng, n)
    qT[: n // 2] = [1.0, 0.0, 0.0, 0.0]  # floors are not turned
    h = 1e-3
    a = rng.normal(size=(n, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.uniform(-0.5, 0.5, size=n)
    xA = rng.uniform(-0.4, 0.4, size=(n, 3))
    xT = rng.uniform(-0.4, 0.4, size=(n, 3))
    s = rng.normal(size=(n, 3)) * 2e-3
    local = qrotT(qrot(qA, s) + xA - xT, qT)
    exact = b + np.einsum("ij,ij->i", a, local)
    g = qrot(qT, a)
    u = qrotT(g, qA)
    c0 = b + np.einsum("ij,ij->i", g, xA - xT)
    fast = c0 + np.einsum("ij,ij->i", u, s)
    scale = np.abs(b) + np.linalg.norm(local, axis=1) + h
    assert np.all(np.abs(fast - exact) <= 1e-12 * scale)
