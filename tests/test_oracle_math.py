"""Pins of the oracle that do not come from the reference's golden vectors.

These fix what the mathematics fixes: partition of unity and linear
reproduction of the B-spline shapes, the discrete continuity equation of the
Esirkepov current, the moment of the direct current, the Boris rotation, the
Vay force-free drift, the stable sort against numpy's stable argsort, and the
periodic guard copy and fold against an independent modular-index
formulation.  They cover the two guard routines (which need MPI in the
reference and cannot be built from its sources here) and the charge
deposition the continuity test uses.
"""
import numpy as np
import pytest

NX, NY, NZ, NG = 9, 8, 7, 3
SHAPE = (NX + 1 + 2 * NG, NY + 1 + 2 * NG, NZ + 1 + 2 * NG)


def zeros():
    return np.zeros(SHAPE, order="F")


def node_coords(d=1.0):
    i = np.arange(-NG, NX + NG + 1)
    j = np.arange(-NG, NY + NG + 1)
    k = np.arange(-NG, NZ + NG + 1)
    return np.meshgrid(i, j, k, indexing="ij")


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("lo", [True, False])
def test_gather_constant_and_linear(orc, order, lo):
    rng = np.random.default_rng(order)
    n = 300
    xp, yp, zp = (rng.uniform(0, m, n) for m in (NX, NY, NZ))
    I, J, K = node_coords()
    # constant fields come back exactly up to rounding of the weights
    F = [np.asfortranarray(np.full(SHAPE, v)) for v in (1.0, -2.0, 3.0, 0.5, 4.0, -1.5)]
    e = [np.zeros(n) for _ in range(6)]
    assert orc.geteb3d_energy_conserving(xp, yp, zp, *e, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG,
                                         NG, order, order, order, *F, lo) == 0
    for v, out in zip((1.0, -2.0, 3.0, 0.5, 4.0, -1.5), e):
        np.testing.assert_allclose(out, v, rtol=0, atol=4e-15 * abs(v) * 8)
    # Ex linear in its two nodal directions is reproduced exactly (first
    # moment of the B-spline); Ey is staggered in y and nodal in x, z.
    ex = np.asfortranarray(0.3 + 0.7 * J + -0.2 * K + 0.0 * I)
    ey = np.asfortranarray(-0.1 + 0.4 * I + 0.9 * K)
    e = [np.zeros(n) for _ in range(6)]
    G = [ex, ey] + [zeros() for _ in range(4)]
    orc.geteb3d_energy_conserving(xp, yp, zp, *e, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG,
                                  order, order, order, *G, lo)
    np.testing.assert_allclose(e[0], 0.3 + 0.7 * yp - 0.2 * zp, atol=1e-12)
    np.testing.assert_allclose(e[1], -0.1 + 0.4 * xp + 0.9 * zp, atol=1e-12)


def test_boris_pure_b_rotation(orc):
    rng = np.random.default_rng(3)
    n = 200
    u = [rng.standard_normal(n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    bz = rng.uniform(0.5, 2.0, n)
    z = np.zeros(n)
    q, m, dt = -1.0, 1.0, 0.3
    uu = [a.copy() for a in u]
    orc.boris_push_u_3d(*uu, g.copy(), z, z, z, z, z, bz, q, m, dt)
    # |u| conserved; rotation about z by 2*atan(q*B*dt/(2*m*gamma))
    np.testing.assert_allclose(uu[0] ** 2 + uu[1] ** 2 + uu[2] ** 2,
                               u[0] ** 2 + u[1] ** 2 + u[2] ** 2, rtol=1e-14)
    np.testing.assert_array_equal(uu[2], u[2])
    ang = np.arctan2(uu[1], uu[0]) - np.arctan2(u[1], u[0])
    ang = (ang + np.pi) % (2 * np.pi) - np.pi
    expect = -2 * np.arctan(q * bz * dt * g / (2 * m))
    np.testing.assert_allclose(ang, expect, atol=1e-12)


def test_vay_force_free(orc):
    # E = -v x B: a particle moving with the E x B drift is not accelerated
    rng = np.random.default_rng(4)
    n = 100
    vx = rng.uniform(-0.5, 0.5, n)
    gam = 1 / np.sqrt(1 - vx ** 2)
    ux, uy, uz = vx * gam, np.zeros(n), np.zeros(n)
    bz = rng.uniform(0.5, 3, n)
    ey = vx * bz  # E + v x B = 0 -> Ey = vx*Bz
    z = np.zeros(n)
    u = [ux.copy(), uy.copy(), uz.copy()]
    gi = 1 / gam
    orc.ebcancelpush3d(*u, gi, z, ey, z, z, z, bz, -1.0, 1.0, 0.7, 0)
    np.testing.assert_allclose(u[0], ux, rtol=1e-13)
    np.testing.assert_allclose(u[1], 0, atol=1e-13)
    np.testing.assert_allclose(gi, 1 / gam, rtol=1e-13)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_esirkepov_continuity(orc, order):
    """(rho_new - rho_old)/dt + div J = 0 at round-off, rho of the same order."""
    rng = np.random.default_rng(10 + order)
    n = 400
    dt = 0.5
    x, y, z = (rng.uniform(2.0, m - 2.0, n) for m in (NX, NY, NZ))
    u = [rng.uniform(-1.5, 1.5, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 1.5, n)
    q = -0.8
    J = [zeros() for _ in range(3)]
    assert orc.depose_esirkepov(*J, x, y, z, *u, g, w, q, 0, 0, 0, dt, 1, 1, 1, NX, NY, NZ, NG,
                                NG, NG, order) == 0
    rho1, rho0 = zeros(), zeros()
    xo, yo, zo = x - dt * u[0] * g, y - dt * u[1] * g, z - dt * u[2] * g
    assert orc.depose_rho(rho1, x, y, z, w, q, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, order) == 0
    assert orc.depose_rho(rho0, xo, yo, zo, w, q, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, order) == 0
    # Jx(i) lives at i+1/2: div J(i) = Jx(i) - Jx(i-1) + ...
    div = np.zeros(SHAPE)
    div[1:, :, :] += J[0][1:, :, :] - J[0][:-1, :, :]
    div[:, 1:, :] += J[1][:, 1:, :] - J[1][:, :-1, :]
    div[:, :, 1:] += J[2][:, :, 1:] - J[2][:, :, :-1]
    res = (rho1 - rho0) / dt + div
    scale = np.abs(rho1).max() / dt
    assert np.abs(res[1:, 1:, 1:]).max() < 1e-13 * scale
    # first moment: sum Jx = sum q w vx  (unit cells)
    np.testing.assert_allclose(J[0].sum(), (q * w * u[0] * g).sum(), rtol=1e-12)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_direct_moment(orc, order):
    rng = np.random.default_rng(20 + order)
    n = 300
    x, y, z = (rng.uniform(1.0, m - 1.0, n) for m in (NX, NY, NZ))
    u = [rng.uniform(-1, 1, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 1.5, n)
    J = [zeros() for _ in range(3)]
    dx, dy, dz = 0.5, 0.75, 1.5
    assert orc.depose_direct(*J, x * dx, y * dy, z * dz, *u, g, w, 1.3, 0, 0, 0, 0.4, dx, dy,
                             dz, NX, NY, NZ, NG, NG, NG, order, False) == 0
    vol = dx * dy * dz
    for c in range(3):
        np.testing.assert_allclose(J[c].sum() * vol, (1.3 * w * u[c] * g).sum(), rtol=1e-12)


def test_fd_weights_closed_form(orc):
    np.testing.assert_array_equal(orc.fd_weights(2), [1.0])
    np.testing.assert_allclose(orc.fd_weights(4), [9 / 8, -1 / 24], rtol=1e-15)


def test_sort_matches_stable_argsort(orc):
    rng = np.random.default_rng(5)
    n = 2000
    P = [rng.uniform(-0.5, 8.4, n), rng.uniform(-0.5, 8.4, n), rng.uniform(-0.5, 8.4, n)]
    P[0][:100] = 3.0  # ties
    P += [rng.standard_normal(n) for _ in range(4)] + [np.arange(n, dtype=np.float64)]
    box = (-0.5, -0.5, -0.5, 9.5, 9.5, 9.5, 1.0, 1.0, 1.0)
    key = (np.floor(P[2] + 0.5) * 100 + np.floor(P[1] + 0.5) * 10 + np.floor(P[0] + 0.5))
    perm = np.argsort(key, kind="stable")
    expect = [p[perm] for p in P]
    bad, keys = orc.particle_bin_sorting(*P, *box, return_keys=True)
    assert bad == 0
    np.testing.assert_array_equal(keys, key.astype(np.int64) + 1)
    for a, b in zip(P, expect):
        np.testing.assert_array_equal(a, b)


def test_sort_out_of_range_is_reported(orc):
    # the reference checks only the key range [1, nx3*ny3*nz3] (sorting.F90:305)
    P = [np.array([0.2, 50.0, 0.2]), np.array([0.2, 0.2, 0.2]), np.array([0.2, 0.2, -0.5])] + \
        [np.zeros(3) for _ in range(5)]
    bad = orc.particle_bin_sorting(*P, 0.0, 0.0, 0.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0)
    assert bad == 2


def _wrap_all(a):
    """Periodic image of the interior [0, n) of every axis, by modular index."""
    I = (np.arange(-NG, NX + NG + 1) % NX) + NG
    J = (np.arange(-NG, NY + NG + 1) % NY) + NG
    K = (np.arange(-NG, NZ + NG + 1) % NZ) + NG
    return a[np.ix_(I, J, K)]


def test_field_bc_periodic(orc):
    rng = np.random.default_rng(6)
    f = np.asfortranarray(rng.standard_normal(SHAPE))
    g = f.copy(order="F")
    orc.field_bc_periodic(g, NX, NY, NZ, NG, NG, NG)
    np.testing.assert_array_equal(g, _wrap_all(f))
    h = g.copy(order="F")
    orc.field_bc_periodic(h, NX, NY, NZ, NG, NG, NG)
    np.testing.assert_array_equal(h, g)  # idempotent


def test_current_bc_periodic(orc):
    rng = np.random.default_rng(7)
    a = np.asfortranarray(rng.standard_normal(SHAPE))
    b = a.copy(order="F")
    orc.current_bc_periodic(b, NX, NY, NZ, NG, NG, NG)
    # every interior node (0..n-1 on each axis) holds the sum over its images
    acc = np.zeros((NX, NY, NZ))
    I = np.arange(-NG, NX + NG + 1)
    J = np.arange(-NG, NY + NG + 1)
    K = np.arange(-NG, NZ + NG + 1)
    np.add.at(acc, np.ix_(I % NX, J % NY, K % NZ), a)
    np.testing.assert_allclose(b[NG:NG + NX, NG:NG + NY, NG:NG + NZ], acc, rtol=1e-13, atol=1e-13)
    # node n on a 1D line receives a(n) + a(0) + a(n) (summation_bcs order)
    c = np.zeros(SHAPE, order="F")
    c[NG + NX, NG + 2, NG + 2] = 1.0
    c[NG + 0, NG + 2, NG + 2] = 10.0
    orc.current_bc_periodic(c, NX, NY, NZ, NG, NG, NG)
    assert c[NG + 0, NG + 2, NG + 2] == 11.0
    assert c[NG + NX, NG + 2, NG + 2] == 12.0
