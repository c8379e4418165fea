"""Geometry of the laser antenna's particles (load_laser_species,
tiling.F90:1806-1893), checked against closed forms on the CPU."""
import math

import numpy as np

from picsar_amd.laser import antenna_plane, emax_from_a0


def test_axis_aligned_antenna():
    """Normal along z through spot: one particle per (x, y) cell centre at
    z = spot_z, antenna-frame coordinates x - spot_x and y - spot_y, weight
    eps0 * Emax / 0.01 * dx * dy."""
    xmin, d, n = (-2.0, 0.5, 1.0), (0.5, 0.25, 1.5), (8, 6, 10)
    spot = (0.3, 1.1, 4.0)
    pos, w, xx, yy = antenna_plane(xmin, d, n, spot, (0, 0, 1), (1, 0, 0), (0, 1, 0), 0.2)
    assert len(w) == 8 * 6
    xs = xmin[0] + (np.arange(8) + 0.5) * d[0]
    ys = xmin[1] + (np.arange(6) + 0.5) * d[1]
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    np.testing.assert_allclose(pos[:, 0], X.ravel(), rtol=0, atol=1e-15)
    np.testing.assert_allclose(pos[:, 1], Y.ravel(), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(pos[:, 2], 4.0)
    np.testing.assert_allclose(xx, X.ravel() - spot[0], atol=1e-15)
    np.testing.assert_allclose(yy, Y.ravel() - spot[1], atol=1e-15)
    np.testing.assert_allclose(w, 0.2 / 0.01 * 0.5 * 0.25, rtol=1e-15)


def test_tilted_antenna():
    """Normal (0, 0.6, 0.8): particles lie on the plane, inside the box along
    z, and (xx, yy) are their coordinates on the orthonormal pair
    (pv1, pv2 = vec x pv1)."""
    xmin, d, n = (0.0, 0.0, 0.0), (0.5, 0.5, 0.5), (12, 20, 24)
    vec = np.array([0.0, 0.6, 0.8])
    pv1 = np.array([1.0, 0.0, 0.0])
    pv2 = np.cross(vec, pv1)
    spot = np.array([3.0, 5.0, 6.0])
    pos, w, xx, yy = antenna_plane(xmin, d, n, spot, vec, pv1, pv2, 1.0)
    assert 0 < len(w) <= 12 * 20
    np.testing.assert_allclose((pos - spot) @ vec, 0.0, atol=1e-13)
    assert (pos[:, 2] >= 0.0).all() and (pos[:, 2] < 12.0).all()
    np.testing.assert_allclose(xx, (pos - spot) @ pv1, atol=1e-13)
    np.testing.assert_allclose(yy, (pos - spot) @ pv2, atol=1e-13)
    # sx(k) = |d_k / pv1_k| for pv1_k != 0, else 1 (tiling.F90:1843-1848): here
    # sx = (0.5, 1, 1), sy = (1, 0.5/0.8, 0.5/0.6)
    np.testing.assert_allclose(w, 1.0 / 0.01 * 0.5 * (0.5 / 0.8), rtol=1e-15)


def test_emax_from_a0():
    assert math.isclose(emax_from_a0(2.0, 32.0), 2.0 * 2 * math.pi / 32.0, rel_tol=1e-15)
