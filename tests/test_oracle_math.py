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


@pytest.mark.parametrize("orders", [(0, 1, 2), (3, 2, 1), (1, 3, 0), (2, 2, 3), (0, 0, 0)])
def test_generic_esirkepov_continuity(orc, orders):
    """pxr_depose_jxjyjz_esirkepov_n with one order per axis satisfies the
    continuity equation with pxr_depose_rho_n of the same orders, and its
    first moment is q w v along every axis of order 1 or more; the generic gather with those orders reproduces
    constant fields to the rounding of its single-precision cubic constants."""
    rng = np.random.default_rng(sum(o * 4 ** a for a, o in enumerate(orders)))
    n = 400
    dt = 0.5
    x, y, z = (rng.uniform(2.0, m - 2.0, n) for m in (NX, NY, NZ))
    u = [rng.uniform(-1.5, 1.5, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 1.5, n)
    q = -0.8
    J = [zeros() for _ in range(3)]
    assert orc.depose_esirkepov_n(*J, x, y, z, *u, g, w, q, 0, 0, 0, dt, 1, 1, 1, NX, NY, NZ,
                                  NG, NG, NG, *orders) == 0
    rho1, rho0 = zeros(), zeros()
    v = [a * g for a in u]
    xo, yo, zo = x - dt * v[0], y - dt * v[1], z - dt * v[2]
    for r, p in ((rho1, (x, y, z)), (rho0, (xo, yo, zo))):
        assert orc.depose_rho_n(r, *p, w, q, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG,
                                *orders) == 0
    div = np.zeros(SHAPE)
    div[1:, :, :] += J[0][1:, :, :] - J[0][:-1, :, :]
    div[:, 1:, :] += J[1][:, 1:, :] - J[1][:, :-1, :]
    div[:, :, 1:] += J[2][:, :, 1:] - J[2][:, :, :-1]
    res = (rho1 - rho0) / dt + div
    scale = np.abs(rho1).max() / dt
    assert np.abs(res[1:, 1:, 1:]).max() < 1e-13 * scale
    for a in range(3):
        # an order-0 axis moves charge only when the nearest node changes
        if orders[a] >= 1:
            np.testing.assert_allclose(J[a].sum(), (q * w * v[a]).sum(), rtol=1e-12)
    if min(orders) >= 1:
        F = [np.asfortranarray(np.full(SHAPE, c)) for c in (1.0, -2.0, 3.0, 0.5, 4.0, -1.5)]
        for lo in (True, False):
            e = [np.zeros(n) for _ in range(6)]
            assert orc.geteb3d_n(x, y, z, *e, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, *orders,
                                 *F, l_lower_order_in_v=lo) == 0
            for c, out in zip((1.0, -2.0, 3.0, 0.5, 4.0, -1.5), e):
                np.testing.assert_allclose(out, c, rtol=1e-6 if 3 in orders else 1e-14)


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


LASER = dict(emax1=2.0, emax2=0.7, k0=2 * np.pi / 32.0, tau=25.0, t_peak=40.0)


def _laser(orc, win, xx, yy, t, polangle=0.0, tord=2, w0=64.0, pv1=(1.0, 0, 0),
           pv2=(0, 1.0, 0)):
    qz = 1j * w0 ** 2 * np.pi / (2 * np.pi / LASER["k0"])
    return orc.laser_profile(win, np.atleast_1d(np.asarray(xx, float)),
                             np.atleast_1d(np.asarray(yy, float)), LASER["emax1"],
                             LASER["emax2"], np.array(pv1), np.array(pv2), LASER["k0"], qz,
                             LASER["tau"], t, LASER["t_peak"], tord, polangle)


def test_laser_gaussian_closed_forms(orc):
    """Gaussian antenna (laser_pusher_manager_3d.F90:368-415), q_z = i pi w0^2
    / lambda: amp = Im[e^{i k0 (t - t_peak) - r^2/w0^2 - ((t - t_peak)/tau)^n}
    (E1 p1 + E2 e^{2 pi i polangle} p2)]."""
    k0, tau, tp = LASER["k0"], LASER["tau"], LASER["t_peak"]
    # on axis at the peak: the phase is 0, only the polangle-shifted term is left
    a = _laser(orc, 0, 0.0, 0.0, tp, polangle=0.25)[0]
    np.testing.assert_allclose(a, [0.0, LASER["emax2"], 0.0], atol=1e-15)
    # a quarter period later, off axis, fourth-order time envelope
    dtq = np.pi / (2 * k0)
    r2 = 10.0 ** 2 + 7.0 ** 2
    a = _laser(orc, 0, 10.0, 7.0, tp + dtq, tord=4)[0]
    env = np.exp(-r2 / 64.0 ** 2 - (dtq / tau) ** 4)
    np.testing.assert_allclose(a, [LASER["emax1"] * env, LASER["emax2"] * env, 0.0],
                               rtol=1e-14, atol=1e-15)
    # temporal_order 0: no time envelope
    a = _laser(orc, 0, 0.0, 0.0, tp + dtq + 500.0 * 2 * np.pi / k0, tord=0)[0]
    np.testing.assert_allclose(a, [LASER["emax1"], LASER["emax2"], 0.0], rtol=1e-12)


def test_laser_hanning_closed_forms(orc):
    """Hanning antenna (laser_pusher_manager_3d.F90:426-478): amp =
    Re[e^{i k0 (t - t_peak) - r^2/w0^2} (E1 p1 + E2 e^{2 pi i polangle} p2)]
    (1 - cos(pi t / t_peak)) / 2 for t <= 2 t_peak, 0 after."""
    tp = LASER["t_peak"]
    a = _laser(orc, 1, 0.0, 0.0, tp, polangle=0.125)[0]
    np.testing.assert_allclose(a, [LASER["emax1"], LASER["emax2"] * np.cos(np.pi / 4), 0.0],
                               rtol=1e-15)
    t = 10.0
    r2 = 3.0 ** 2 + 5.0 ** 2
    a = _laser(orc, 1, 3.0, -5.0, t)[0]
    ph = LASER["k0"] * (t - tp)
    win = 0.5 - 0.5 * np.cos(np.pi * t / tp)
    np.testing.assert_allclose(a, np.array([LASER["emax1"], LASER["emax2"], 0.0]) *
                               np.exp(-r2 / 64.0 ** 2) * np.cos(ph) * win, rtol=1e-13,
                               atol=1e-15)
    assert not _laser(orc, 1, 0.0, 0.0, 2 * tp + 1e-9).any()


def test_laser_pusher_moves_by_the_profile(orc):
    """laserp_pusher_* (laser_pusher_manager_3d.F90:195-356): u = charge *
    amp * 0.01 c / emax from the profile at pid columns 2 and 3, gaminv = 1,
    x += dt u."""
    rng = np.random.default_rng(11)
    n = 40
    pid = np.asfortranarray(np.stack([np.full(n, 3.0), rng.standard_normal(n) * 30,
                                      rng.standard_normal(n) * 30], axis=1))
    x0 = [rng.random(n) * 10 for _ in range(3)]
    pv1, pv2 = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    qz = 1j * 64.0 ** 2 * np.pi / 32.0
    for win in (0, 1):
        x = [a.copy() for a in x0]
        u = [np.zeros(n) for _ in range(3)]
        g = np.zeros(n)
        emax = 2.5
        orc.laserp_pusher(win, pid, *x, *u, g, 0.5, emax, LASER["emax1"], LASER["emax2"], pv1,
                          pv2, LASER["k0"], qz, -1.0, LASER["tau"], 33.0, LASER["t_peak"], 2,
                          0.0)
        amp = orc.laser_profile(win, pid[:, 1].copy(), pid[:, 2].copy(), LASER["emax1"],
                                LASER["emax2"], pv1, pv2, LASER["k0"], qz, LASER["tau"], 33.0,
                                LASER["t_peak"], 2, 0.0)
        for d in range(3):
            np.testing.assert_array_equal(u[d], -1.0 * amp[:, d] * (0.01 / emax))
            np.testing.assert_array_equal(x[d], x0[d] + 0.5 * u[d])
        np.testing.assert_array_equal(g, 1.0)


def _interior(a, lo, hi):
    """nodes lo..hi[d] of each axis d of a grid array"""
    return a[NG + lo:NG + hi[0] + 1, NG + lo:NG + hi[1] + 1, NG + lo:NG + hi[2] + 1]


def test_divergence_of_a_yee_curl_vanishes(orc):
    """div E by backward differences (calc_field_div) is unchanged by the Yee
    E push without current, and div B by forward differences
    (calc_field_divB) by the B push: the discrete divergence of the discrete
    curl is zero.  A difference taken the other way round breaks both."""
    rng = np.random.default_rng(5)
    d = (0.9, 1.1, 1.3)
    dt = 0.4
    F = {k: np.asfortranarray(rng.standard_normal(SHAPE)) for k in
         ("ex", "ey", "ez", "bx", "by", "bz")}
    J = [zeros() for _ in range(3)]
    lo, hi = (0, 0, 0), (NX, NY, NZ)
    before, after = zeros(), zeros()
    orc.calc_field_div(before, F["ex"], F["ey"], F["ez"], NX, NY, NZ, NG, NG, NG, *d)
    orc.push_em3d_evec(lo, hi, *(F[k] for k in ("ex", "ey", "ez", "bx", "by", "bz")), *J, dt,
                       dt / d[0], dt / d[1], dt / d[2], NX, NY, NZ, NG, NG, NG)
    orc.calc_field_div(after, F["ex"], F["ey"], F["ez"], NX, NY, NZ, NG, NG, NG, *d)
    # E moved at the nodes 0..n: div E at 1..n sees only updated values
    diff = _interior(after - before, 1, (NX, NY, NZ))
    assert np.abs(diff).max() < 1e-13 * np.abs(_interior(before, 1, (NX, NY, NZ))).max()
    assert np.abs(_interior(after - before, 0, (NX, NY, NZ))).max() > 1e-3  # guards not moved
    orc.calc_field_divb(before, F["bx"], F["by"], F["bz"], NX, NY, NZ, NG, NG, NG, *d)
    orc.push_em3d_bvec(lo, hi, *(F[k] for k in ("ex", "ey", "ez", "bx", "by", "bz")),
                       dt / d[0], dt / d[1], dt / d[2], NX, NY, NZ, NG, NG, NG)
    orc.calc_field_divb(after, F["bx"], F["by"], F["bz"], NX, NY, NZ, NG, NG, NG, *d)
    m = (NX - 1, NY - 1, NZ - 1)  # div B at n reads the guard node n+1
    diff = _interior(after - before, 1, m)
    assert np.abs(diff).max() < 1e-13 * np.abs(_interior(before, 1, m)).max()


def test_divergence_of_linear_fields(orc):
    """Fields linear in the node index have divergence a/dx + b/dy + c/dz at
    every node of the loops; the nodes outside keep their value."""
    d = (0.9, 1.1, 1.3)
    i, j, k = node_coords()
    a, b, c = 0.7, -1.9, 2.3
    fx, fy, fz = (np.asfortranarray(v * 1.0) for v in (a * i, b * j, c * k))
    want = a / d[0] + b / d[1] + c / d[2]
    for fn, lo in ((orc.calc_field_div, 0), (orc.calc_field_divb, 1)):
        out = np.asfortranarray(np.full(SHAPE, 3.5))
        fn(out, fx, fy, fz, NX, NY, NZ, NG, NG, NG, *d)
        inner = _interior(out, lo, (NX, NY, NZ))
        np.testing.assert_allclose(inner, want, rtol=1e-13)
        assert (out == 3.5).sum() == out.size - inner.size


def test_divergence_norms(orc):
    """get_norm_divErho and get_loc_norm_2 (diags.F90:1222-1345) against
    numpy sums over the nodes 1..n of each axis."""
    rng = np.random.default_rng(8)
    dv, rho = (np.asfortranarray(rng.standard_normal(SHAPE)) for _ in range(2))
    s = (slice(NG + 1, NG + NX + 1), slice(NG + 1, NG + NY + 1), slice(NG + 1, NG + NZ + 1))
    want = np.sqrt(((dv[s] - rho[s]) ** 2).sum())
    assert orc.norm_diverho(dv, rho, NX, NY, NZ, NG, NG, NG) == pytest.approx(want, rel=1e-14)
    assert orc.loc_norm_2(dv, NX, NY, NZ, NG, NG, NG) == pytest.approx((dv[s] ** 2).sum(),
                                                                       rel=1e-14)
