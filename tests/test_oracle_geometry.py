"""Pins of the oracle's envelope arithmetic against closed forms.

Octahedron::set_coordinate_transformation_matrix, ::cell_is_captured and
::capture_cell (Octahedron.cpp:98-310) have no reference tests; these checks use
what the geometry fixes, chosen so that a transposed rotation, a sign or
octant-table error, or a dropped term fails at least one of them.
"""
import ctypes as C

import numpy as np
import pytest

F = C.POINTER(C.c_double)
TRUNC = float(np.sqrt(np.float32(3.0)))
SQRT2F = float(np.sqrt(np.float32(2.0)))


def _p(a):
    return a.ctypes.data_as(F)


def rot(oracle, e):
    e = np.asarray(e, np.float64)
    r = np.empty(9)
    oracle.lib().oracle_rotation(_p(e), _p(r))
    return r


def child(oracle, d, L, h=1.0, trunc=TRUNC, R=None):
    d = np.ascontiguousarray(d, np.float64)
    out = np.empty(4)
    oracle.lib().oracle_child_envelope(_p(d), float(L), float(h), float(trunc),
                                       _p(np.ascontiguousarray(R, np.float64)), _p(out))
    return out


def test_rotation_is_proper_orthonormal(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        e = rng.uniform([0, 0, 0], [2 * np.pi, np.pi, np.pi / 2])
        R = rot(oracle, e).reshape(3, 3)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1.0) < 1e-13


def test_rotation_zxz_closed_form(oracle):
    """Z-X-Z Euler rotation: R = Rz(alpha) Rx(beta) Rz(gamma)."""
    def rz(t):
        return np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])

    def rx(t):
        return np.array([[1, 0, 0], [0, np.cos(t), -np.sin(t)], [0, np.sin(t), np.cos(t)]])

    rng = np.random.default_rng(1)
    for _ in range(100):
        a, b, g = rng.uniform([0, 0, 0], [2 * np.pi, np.pi, np.pi / 2])
        np.testing.assert_allclose(rot(oracle, [a, b, g]).reshape(3, 3), rz(a) @ rx(b) @ rz(g),
                                   atol=1e-14)


def test_capture_test_is_l1_ball_in_grain_frame(oracle):
    """Inside iff ||R^T (x - c)||_1 <= L, evaluated independently with numpy."""
    rng = np.random.default_rng(2)
    L = oracle.lib()
    n_in = 0
    for _ in range(4000):
        R = rot(oracle, rng.uniform([0, 0, 0], [2 * np.pi, np.pi, np.pi / 2]))
        c = rng.uniform(-1, 1, 3)
        x = rng.uniform(-3, 3, 3)
        Lg = rng.uniform(0.1, 4.0)
        norm = np.abs(R.reshape(3, 3).T @ (x - c)).sum()
        if abs(norm - Lg) < 1e-9:
            continue
        env = np.array([c[0], c[1], c[2], Lg])
        got = L.oracle_capture_test(_p(env), _p(R), _p(np.ascontiguousarray(x)))
        assert got == int(norm <= Lg)
        n_in += got
    assert 200 < n_in < 3800
    # zero length never captures, even at the centre
    env = np.array([0.0, 0.0, 0.0, 0.0])
    assert L.oracle_capture_test(_p(env), _p(rot(oracle, [0, 0, 0])), _p(np.zeros(3))) == 0


def test_child_at_parent_centre(oracle):
    """d = 0: A is the face centre, equidistant from the face's vertices; both
    projections equal L/sqrt2.  Below the truncation both are replaced by half
    the edge, so Lnew = sqrt2f*L/sqrt2 ~ L and the shift ~ 0; above it Lnew is
    sqrt2f times the truncation length."""
    R = rot(oracle, [0.3, 1.1, 0.7])
    for L in (0.2, 0.5, 1.0, 2.0):
        out = child(oracle, [0.0, 0.0, 0.0], L, R=R)
        assert abs(out[3] - L) < 1e-6 * L
        assert np.linalg.norm(out[:3]) < 1e-6 * L
    for L in (3.0, 5.0):
        out = child(oracle, [0.0, 0.0, 0.0], L, R=R)
        assert out[3] == SQRT2F * TRUNC
        assert abs(np.linalg.norm(out[:3]) - (L - out[3]) / np.sqrt(3)) < 1e-12 * L


@pytest.mark.parametrize("octant", range(8))
def test_child_shift_along_nearest_vertex(oracle, octant):
    """For a cell near a vertex direction of its octant, S1 is that vertex: the
    child centre moves along that axis of the grain frame (a column of R), by
    (L - Lnew)/sqrt3 with the vertex's sign."""
    sx = -1 if octant & 1 else 1
    sy = -1 if octant & 2 else 1
    sz = -1 if octant & 4 else 1
    rng = np.random.default_rng(10 + octant)
    for axis in range(3):
        R = rot(oracle, rng.uniform([0, 0, 0], [2 * np.pi, np.pi, np.pi / 2])).reshape(3, 3)
        L = 2.5
        loc = np.array([0.1 * sx, 0.12 * sy, 0.08 * sz])
        loc[axis] = 1.6 * (sx, sy, sz)[axis]
        d = R @ loc
        out = child(oracle, d, L, R=R.reshape(-1))
        shift_local = R.T @ out[:3]
        expect = np.zeros(3)
        expect[axis] = (sx, sy, sz)[axis] * (L - out[3]) / np.sqrt(3)
        np.testing.assert_allclose(shift_local, expect, atol=1e-12)
        assert 0 < out[3] <= SQRT2F * TRUNC


def test_child_length_bounds_and_truncation(oracle):
    """Lnew <= sqrt2f * trunc * h; a smaller truncation coefficient (the
    envelope.max_half_diagonal_length option) lowers the bound."""
    rng = np.random.default_rng(3)
    for trunc in (TRUNC, SQRT2F * 1.0 * 0.5):
        for _ in range(2000):
            R = rot(oracle, rng.uniform([0, 0, 0], [2 * np.pi, np.pi, np.pi / 2]))
            L = rng.uniform(0.2, 6.0)
            loc = rng.uniform(-1, 1, 3)
            loc *= rng.uniform(0, L) / np.abs(loc).sum()
            d = R.reshape(3, 3) @ loc
            out = child(oracle, d, L, trunc=trunc, R=R)
            assert 0 <= out[3] <= SQRT2F * trunc
