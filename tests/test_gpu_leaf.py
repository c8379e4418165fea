"""GPU parity of the leaf kernels against the oracle and the golden vectors.

Every call goes through the C ABI (libpxrhip.so via picsar_amd.picsarpy).
Bars: pushers, position push, FDTD stencils, guard routines and the sort are
bit-exact; gathers and deposits (different summation order, FP64 atomics)
agree to 1e-12 relative in norm; cell indices are exact (a wrong index moves
a whole shape weight and fails the 1e-12 bar by many orders of magnitude).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, NZ, NG = 20, 18, 14, 3
SHAPE = (NX + 1 + 2 * NG, NY + 1 + 2 * NG, NZ + 1 + 2 * NG)
REL = 1e-12


@pytest.fixture(scope="module")
def dev():
    from picsar_amd import _lib
    _lib.load()
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return _lib


@pytest.fixture(scope="module")
def pxr(dev):
    from picsar_amd.picsarpy import picsar
    picsar.l_nodalgrid = False
    return picsar


def T(dev, a):
    from picsar_amd.devmem import DeviceArray
    return DeviceArray.from_numpy(np.asarray(a, dtype=np.float64))


def TG(dev, a):
    from picsar_amd.devmem import DeviceGrid
    g = DeviceGrid(a.shape[0] - 1, a.shape[1] - 1, a.shape[2] - 1, 0)
    g.upload(a)
    return g


def H(t):
    return t.numpy()


def HG(t):
    return t.numpy()


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(
        np.linalg.norm(np.asarray(b).ravel()), 1e-300)


def particles(rng, n, lo=0.0, margin=0.0):
    x = rng.uniform(lo + margin, NX - margin, n)
    y = rng.uniform(lo + margin, NY - margin, n)
    z = rng.uniform(lo + margin, NZ - margin, n)
    k = n // 10
    x[:k] = np.floor(x[:k])  # on nodes
    y[k:2 * k] = np.floor(y[k:2 * k]) + 0.5
    return x, y, z


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("lo", [True, False])
@pytest.mark.parametrize("nodal", [False, True])
def test_gather(dev, pxr, orc, order, lo, nodal):
    rng = np.random.default_rng(100 + order)
    n = 30000
    dx, dy, dz = 0.6, 0.8, 1.1
    xmin, ymin, zmin = -2.0, 1.0, 0.5
    x, y, z = particles(rng, n)
    xp, yp, zp = xmin + x * dx, ymin + y * dy, zmin + z * dz
    F = [np.asfortranarray(rng.standard_normal(SHAPE)) for _ in range(6)]
    e_ref = [np.full(n, 0.25) for _ in range(6)]  # accumulate into non-zero
    assert orc.geteb3d_energy_conserving(xp, yp, zp, *e_ref, xmin, ymin, zmin, dx, dy, dz, NX,
                                         NY, NZ, NG, NG, NG, order, order, order, *F, lo,
                                         nodal) == 0
    e = [T(dev, np.full(n, 0.25)) for _ in range(6)]
    pxr.l_nodalgrid = nodal
    try:
        pxr.geteb3d_energy_conserving(n, T(dev, xp), T(dev, yp), T(dev, zp), *e, xmin, ymin, zmin,
                                      dx, dy, dz, NX, NY, NZ, NG, NG, NG, order, order, order,
                                      *[TG(dev, f) for f in F], False, lo, 8, 0)
    finally:
        pxr.l_nodalgrid = False
    for c in range(6):
        assert rel(H(e[c]), e_ref[c]) < REL, (c, rel(H(e[c]), e_ref[c]))


def test_gather_golden(dev, pxr, golden, ggrid):
    d = golden("gather")
    g = ggrid
    n = len(d["xp"])
    F = [TG(dev, d[k]) for k in ("exg", "eyg", "ezg", "bxg", "byg", "bzg")]
    for order in (1, 2, 3):
        for lo in (1, 0):
            e = [T(dev, np.zeros(n)) for _ in range(6)]
            pxr.geteb3d_energy_conserving(n, T(dev, d["xp"]), T(dev, d["yp"]), T(dev, d["zp"]),
                                          *e, g["xmin"], g["ymin"], g["zmin"], g["dx"], g["dy"],
                                          g["dz"], g["nx"], g["ny"], g["nz"], g["ng"], g["ng"],
                                          g["ng"], order, order, order, *F, False, bool(lo), 8, 0)
            out = np.stack([H(t) for t in e])
            assert rel(out, d[f"o{order}_lo{lo}_a0"]) < REL
            assert rel(out, d[f"o{order}_lo{lo}_a1"]) < REL


def test_gather_numpy_arrays_roundtrip(dev, pxr, golden, ggrid):
    """numpy inputs are staged on the device and outputs written back."""
    d = golden("gather")
    g = ggrid
    n = len(d["xp"])
    e = [np.zeros(n) for _ in range(6)]
    F = [np.asfortranarray(d[k]) for k in ("exg", "eyg", "ezg", "bxg", "byg", "bzg")]
    pxr.geteb3d_energy_conserving(n, d["xp"], d["yp"], d["zp"], *e, g["xmin"], g["ymin"],
                                  g["zmin"], g["dx"], g["dy"], g["dz"], g["nx"], g["ny"], g["nz"],
                                  g["ng"], g["ng"], g["ng"], 3, 3, 3, *F, False, True, 8, 0)
    assert rel(np.stack(e), d["o3_lo1_a1"]) < REL


def test_pushers_golden_bitexact(dev, pxr, golden):
    d = golden("push")
    E = [T(dev, d[k]) for k in ("ex", "ey", "ez", "bx", "by", "bz")]
    q, m, dt = float(d["q"]), float(d["m"]), float(d["dt"])
    n = len(d["ux"])
    u = [T(dev, d[k]) for k in ("ux", "uy", "uz")]
    g = T(dev, d["gaminv"])
    pxr.pxr_boris_push_u_3d(n, *u, g, *E, q, m, dt)
    np.testing.assert_array_equal(np.stack([H(t) for t in u + [g]]), d["boris"])
    for which in (0, 1, 2):
        u = [T(dev, d[k]) for k in ("ux", "uy", "uz")]
        g = T(dev, d["gaminv"])
        pxr.pxr_ebcancelpush3d(n, *u, g, *E, q, m, dt, which)
        np.testing.assert_array_equal(np.stack([H(t) for t in u + [g]]), d[f"vay{which}"])
    g = T(dev, np.zeros(n))
    pxr.pxr_set_gamma(n, T(dev, d["ux"]), T(dev, d["uy"]), T(dev, d["uz"]), g)
    np.testing.assert_array_equal(H(g), d["set_gamma"])
    x = [T(dev, d[k]) for k in ("x", "y", "z")]
    pxr.pxr_pushxyz(n, *x, T(dev, d["ux"]), T(dev, d["uy"]), T(dev, d["uz"]), T(dev, d["gaminv"]),
                    dt)
    np.testing.assert_array_equal(np.stack([H(t) for t in x]), d["pushxyz"])


@pytest.mark.parametrize("case", range(7))
def test_laser_pusher(dev, pxr, orc, golden, case):
    """laserp_pusher_gaussian / laserp_pusher_hanning on the GPU: the momenta
    are charge * amp * 0.01 / emax with amp from gaussian_profile /
    hanning_profile (golden vectors of the reference's own routines) to a few
    ulps, gaminv = 1 and x += dt u, against the oracle's pusher."""
    d = golden("laser")
    win, tord, pol, t, qr, qi = d[f"case{case}"]
    xx, yy = d["xx"], d["yy"]
    n = len(xx)
    pid = np.asfortranarray(np.stack([np.full(n, 2.0), xx, yy], axis=1))
    rng = np.random.default_rng(case)
    x0 = [rng.random(n) * 8.0 for _ in range(3)]
    emax, charge, dt = 2.4, -1.0, 0.55
    args = (emax, float(d["emax1"]), float(d["emax2"]), d["pv1"], d["pv2"], float(d["k0"]),
            complex(qr, qi), charge)
    X = [T(dev, a) for a in x0]
    U = [T(dev, np.zeros(n)) for _ in range(3)]
    G = T(dev, np.zeros(n))
    P = T(dev, pid.ravel(order="F"))
    if win == 0:
        pxr.laserp_pusher_gaussian(n, 3, P, *X, *U, G, dt, 8, *args, float(d["tau"]), float(t),
                                   float(d["t_peak"]), int(tord), float(pol))
    else:
        pxr.laserp_pusher_hanning(n, 3, P, *X, *U, G, dt, 8, *args, float(t),
                                  float(d["t_peak"]), int(tord), float(pol))
    x = [a.copy() for a in x0]
    u = [np.zeros(n) for _ in range(3)]
    g = np.zeros(n)
    orc.laserp_pusher(int(win), pid, *x, *u, g, dt, *args, float(d["tau"]), float(t),
                      float(d["t_peak"]), int(tord), float(pol))
    amp = d[f"amp{case}"]
    scale = max(np.abs(amp).max() * 0.01 / emax, 1e-300)
    for k in range(3):
        assert np.abs(H(U[k]) - charge * amp[:, k] * (0.01 / emax)).max() <= 16e-16 * scale
        assert np.abs(H(U[k]) - u[k]).max() <= 16e-16 * scale
        assert np.abs(H(X[k]) - x[k]).max() <= 1e-15 * 8.0
    np.testing.assert_array_equal(H(G), 1.0)


def test_pushers_large_bitexact(dev, pxr, orc):
    rng = np.random.default_rng(7)
    n = 200000
    u0 = [rng.standard_normal(n) * s for s in (0.5, 0.5, 5.0)]
    E = [rng.standard_normal(n) for _ in range(6)]
    g0 = 1 / np.sqrt(1 + u0[0] ** 2 + u0[1] ** 2 + u0[2] ** 2)
    uu = [a.copy() for a in u0]
    gg = g0.copy()
    orc.ebcancelpush3d(*uu, gg, *E, 1.0, 1836.0, 0.5, 0)
    u = [T(dev, a) for a in u0]
    g = T(dev, g0)
    pxr.pxr_ebcancelpush3d(n, *u, g, *[T(dev, a) for a in E], 1.0, 1836.0, 0.5, 0)
    for a, b in zip([H(t) for t in u + [g]], uu + [gg]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("algo", [0, 3])
def test_deposit(dev, pxr, orc, order, algo):
    rng = np.random.default_rng(200 + order)
    n = 40000
    dt, dx, dy, dz = 0.45, 0.9, 1.0, 1.2
    xmin, ymin, zmin = 0.3, -1.0, 2.0
    x, y, z = particles(rng, n, margin=1.0)
    xp, yp, zp = xmin + x * dx, ymin + y * dy, zmin + z * dz
    u = [rng.uniform(-1.5, 1.5, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 2.0, n)
    q = -1.1
    J0 = [np.asfortranarray(rng.standard_normal(SHAPE) * 0.1) for _ in range(3)]
    Jr = [j.copy(order="F") for j in J0]
    args = (xp, yp, zp, *u, g, w, q, xmin, ymin, zmin, dt, dx, dy, dz, NX, NY, NZ, NG, NG, NG,
            order)
    bad = orc.depose_esirkepov(*Jr, *args) if algo == 0 else orc.depose_direct(*Jr, *args, False)
    assert bad == 0
    J = [TG(dev, j) for j in J0]
    pxr.depose_jxjyjz(*J, n, *[T(dev, a) for a in (xp, yp, zp, *u, g, w)], q, xmin, ymin, zmin,
                      dt, dx, dy, dz, NX, NY, NZ, NG, NG, NG, order, order, order, algo)
    for c in range(3):
        got = HG(J[c])
        assert rel(got - J0[c], Jr[c] - J0[c]) < REL, (c, rel(got - J0[c], Jr[c] - J0[c]))


def test_deposit_golden(dev, pxr, golden, ggrid):
    d = golden("deposit")
    g = ggrid
    n = len(d["xp"])
    for order in (1, 2, 3):
        for algo in (0, 3):
            ref = d[f"o{order}_a{algo}"]
            J = [TG(dev, np.zeros(ref.shape[1:], order="F")) for _ in range(3)]
            pxr.depose_jxjyjz(*J, n, *[T(dev, d[k]) for k in ("xp", "yp", "zp", "ux", "uy", "uz",
                                                            "gaminv", "w")],
                              float(d["q"]), g["xmin"], g["ymin"], g["zmin"], float(d["dt"]),
                              g["dx"], g["dy"], g["dz"], g["nx"], g["ny"], g["nz"], g["ng"],
                              g["ng"], g["ng"], order, order, order, algo)
            got = np.stack([HG(t) for t in J])
            assert rel(got, ref) < REL, (order, algo, rel(got, ref))


@pytest.mark.parametrize("order", [1, 2, 3])
def test_rho(dev, pxr, orc, golden, ggrid, order):
    """pxr_depose_rho_n against the oracle (random particles, accumulating
    into a non-zero rho) and the reference's golden vectors."""
    rng = np.random.default_rng(300 + order)
    n = 40000
    dx, dy, dz = 0.9, 1.0, 1.2
    xmin, ymin, zmin = 0.3, -1.0, 2.0
    x, y, z = particles(rng, n, margin=1.0)
    xp, yp, zp = xmin + x * dx, ymin + y * dy, zmin + z * dz
    w = rng.uniform(0.5, 2.0, n)
    r0 = np.asfortranarray(rng.standard_normal(SHAPE) * 0.1)
    rr = r0.copy(order="F")
    assert orc.depose_rho(rr, xp, yp, zp, w, -1.1, xmin, ymin, zmin, dx, dy, dz, NX, NY, NZ, NG,
                          NG, NG, order) == 0
    R = TG(dev, r0)
    pxr.pxr_depose_rho_n(R, n, *[T(dev, a) for a in (xp, yp, zp, w)], -1.1, xmin, ymin, zmin,
                         dx, dy, dz, NX, NY, NZ, NG, NG, NG, order, order, order)
    assert rel(HG(R) - r0, rr - r0) < REL
    d, g = golden("rho"), ggrid
    ref = d[f"o{order}_n"]
    R = TG(dev, np.zeros(ref.shape, order="F"))
    pxr.pxr_depose_rho_n(R, len(d["xp"]), *[T(dev, d[k]) for k in ("xp", "yp", "zp", "w")],
                         float(d["q"]), g["xmin"], g["ymin"], g["zmin"], g["dx"], g["dy"],
                         g["dz"], g["nx"], g["ny"], g["nz"], g["ng"], g["ng"], g["ng"], order,
                         order, order)
    assert rel(HG(R), ref) < REL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_esirkepov_continuity_on_gpu(dev, pxr, order):
    """Discrete continuity of the GPU Esirkepov current: with rho of the same
    order deposited on the GPU at the old and new positions,
    (rho_new - rho_old)/dt + div J vanishes to round-off at every node.
    Particles move up to ~0.8 cell, so they cross cell faces."""
    rng = np.random.default_rng(400 + order)
    n = 50000
    dt, dx, dy, dz = 0.5, 1.0, 1.0, 1.0
    x, y, z = particles(rng, n, margin=2.0)
    u = [rng.uniform(-1.5, 1.5, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 1.5, n)
    q = -0.8
    J = [TG(dev, np.zeros(SHAPE, order="F")) for _ in range(3)]
    pxr.depose_jxjyjz(*J, n, *[T(dev, a) for a in (x, y, z, *u, g, w)], q, 0.0, 0.0, 0.0, dt, dx,
                      dy, dz, NX, NY, NZ, NG, NG, NG, order, order, order, 0)
    xo, yo, zo = x - dt * (u[0] * g), y - dt * (u[1] * g), z - dt * (u[2] * g)
    rho = []
    for (a, b, c) in ((x, y, z), (xo, yo, zo)):
        R = TG(dev, np.zeros(SHAPE, order="F"))
        pxr.pxr_depose_rho_n(R, n, *[T(dev, v) for v in (a, b, c, w)], q, 0.0, 0.0, 0.0, dx, dy,
                             dz, NX, NY, NZ, NG, NG, NG, order, order, order)
        rho.append(HG(R))
    Jx, Jy, Jz = (HG(j) for j in J)
    div = np.zeros(SHAPE)
    div[1:, :, :] += Jx[1:, :, :] - Jx[:-1, :, :]
    div[:, 1:, :] += Jy[:, 1:, :] - Jy[:, :-1, :]
    div[:, :, 1:] += Jz[:, :, 1:] - Jz[:, :, :-1]
    res = (rho[0] - rho[1]) / dt + div
    scale = np.abs(rho[0]).max() / dt
    assert np.abs(res[1:, 1:, 1:]).max() < 1e-12 * scale, np.abs(res[1:, 1:, 1:]).max() / scale


def test_fdtd_golden_bitexact(dev, pxr, golden, ggrid):
    d = golden("fdtd")
    g = ggrid
    nx, ny, nz, ng = g["nx"], g["ny"], g["nz"], g["ng"]
    lo, hi = (-1, -1, -1), (nx + 1, ny + 1, nz + 1)
    alo, ahi = (-ng, -ng, -ng), (nx + ng, ny + ng, nz + ng)
    mudt, sx, sy, sz = 0.45, 0.45 / g["dx"], 0.45 / g["dy"], 0.45 / g["dz"]
    F = [TG(dev, d[f"in{i}"]) for i in range(9)]
    pxr.pxrpush_em3d_evec(lo, hi, lo, hi, lo, hi, *sum(([f, alo, ahi] for f in F), []), mudt, sx,
                          sy, sz)
    np.testing.assert_array_equal(np.stack([HG(f) for f in F[:3]]), d["yee_e"])
    F = [TG(dev, d[f"in{i}"]) for i in range(9)]
    pxr.pxrpush_em3d_bvec(lo, hi, lo, hi, lo, hi, *sum(([f, alo, ahi] for f in F[:6]), []),
                          0.5 * sx, 0.5 * sy, 0.5 * sz)
    np.testing.assert_array_equal(np.stack([HG(f) for f in F[3:6]]), d["yee_b"])
    for no in (2, 4, 6):
        c = d[f"fdw{no}"]
        F = [TG(dev, d[f"in{i}"]) for i in range(9)]
        pxr.pxrpush_em3d_evec_norder(*F, mudt, sx * c, sy * c, sz * c, nx, ny, nz, no, no, no, ng,
                                     ng, ng, 1, 1, 1)
        np.testing.assert_array_equal(np.stack([HG(f) for f in F[:3]]), d[f"no{no}_e"])
        F = [TG(dev, d[f"in{i}"]) for i in range(9)]
        pxr.pxrpush_em3d_bvec_norder(*F[:6], 0.5 * sx * c, 0.5 * sy * c, 0.5 * sz * c, nx, ny, nz,
                                     no, no, no, ng, ng, ng, 1, 1, 1)
        np.testing.assert_array_equal(np.stack([HG(f) for f in F[3:6]]), d[f"no{no}_b"])


@pytest.mark.parametrize("nodal", [False, True])
@pytest.mark.parametrize("orders,ng,nxs", [((4, 4, 4), 3, 1), ((2, 2, 2), 2, 0),
                                           ((6, 6, 6), 4, 2), ((8, 8, 8), 4, 1),
                                           ((4, 2, 6), 3, 1)])
def test_fdtd_norder_bitexact(dev, pxr, orc, nodal, orders, ng, nxs):
    """pxrpush_em3d_{e,b}vec_norder against the oracle, bit for bit, on a grid
    deeper than one z chunk of the register-blocked kernel (full stencils up
    to order 8 on both staggerings) and with unequal orders per axis (the
    per-node kernel)."""
    from picsar_amd.sim import _fd_weights
    rng = np.random.default_rng(21 + sum(orders))
    nx, ny, nz = 37, 9, 140
    shp = (nx + 1 + 2 * ng, ny + 1 + 2 * ng, nz + 1 + 2 * ng)
    A = [np.asfortranarray(rng.standard_normal(shp)) for _ in range(9)]
    d = (0.9, 1.1, 1.3)
    dt = 0.3
    coef = [dt / d[a] * _fd_weights(orders[a]) for a in range(3)]
    want = [x.copy(order="F") for x in A]
    orc.push_em3d_evec_norder(*want, dt, *coef, nx, ny, nz, *orders, ng, ng, ng, nxs, nxs, nxs,
                              l_nodal=nodal)
    got = [x.copy(order="F") for x in A]
    pxr.pxrpush_em3d_evec_norder(*got, dt, *coef, nx, ny, nz, *orders, ng, ng, ng, nxs, nxs, nxs,
                                 nodal)
    for a, b in zip(got[:3], want[:3]):
        np.testing.assert_array_equal(a, b)
    want = [x.copy(order="F") for x in A]
    orc.push_em3d_bvec_norder(*want[:6], *coef, nx, ny, nz, *orders, ng, ng, ng, nxs, nxs, nxs,
                              l_nodal=nodal)
    got = [x.copy(order="F") for x in A]
    pxr.pxrpush_em3d_bvec_norder(*got[:6], *coef, nx, ny, nz, *orders, ng, ng, ng, nxs, nxs, nxs,
                                 nodal)
    for a, b in zip(got[3:6], want[3:6]):
        np.testing.assert_array_equal(a, b)


def test_guards_bitexact(dev, pxr, orc):
    rng = np.random.default_rng(9)
    a = np.asfortranarray(rng.standard_normal(SHAPE))
    b = a.copy(order="F")
    orc.field_bc_periodic(b, NX, NY, NZ, NG, NG, NG)
    t = TG(dev, a)
    pxr.field_bc(t, NG, NG, NG, NX, NY, NZ)
    np.testing.assert_array_equal(HG(t), b)
    b = a.copy(order="F")
    orc.current_bc_periodic(b, NX, NY, NZ, NG, NG, NG)
    t = TG(dev, a)
    pxr.summation_bcs(t, NG, NG, NG, NX, NY, NZ)
    np.testing.assert_array_equal(HG(t), b)


def test_field_divergence_bitexact(dev, pxr, orc, golden):
    """calc_field_div / calc_field_divB: bit for bit against the golden
    vectors of the reference's routines and against the oracle; the nodes
    outside the loops keep their value."""
    d = golden("div")
    nx, ny, nz, g = (int(d[k]) for k in ("nx", "ny", "nz", "ng"))
    h = tuple(float(d[k]) for k in ("dx", "dy", "dz"))
    F = [TG(dev, np.asfortranarray(d[k])) for k in ("fx", "fy", "fz")]
    for name, fn in (("dive", pxr.calc_field_div), ("divb", pxr.calc_field_divB)):
        out = TG(dev, np.asfortranarray(np.full(d["fx"].shape, 7.0)))
        fn(out, *F, nx, ny, nz, g, g, g, *h)
        np.testing.assert_array_equal(HG(out), d[name])
    rng = np.random.default_rng(12)
    A = [np.asfortranarray(rng.standard_normal(SHAPE)) for _ in range(3)]
    h = (0.7, 1.2, 0.95)
    for ofn, gfn in ((orc.calc_field_div, pxr.calc_field_div),
                     (orc.calc_field_divb, pxr.calc_field_divB)):
        want = np.asfortranarray(np.zeros(SHAPE))
        ofn(want, *A, NX, NY, NZ, NG, NG, NG, *h)
        got = np.asfortranarray(np.zeros(SHAPE))
        gfn(got, *A, NX, NY, NZ, NG, NG, NG, *h)  # numpy arrays: staged and written back
        np.testing.assert_array_equal(got, want)


def test_divergence_norms(dev, pxr, orc):
    """get_norm_divErho and get_loc_norm_2 against the oracle's sums (the
    GPU sums in another order)."""
    rng = np.random.default_rng(13)
    dv, rho = (np.asfortranarray(rng.standard_normal(SHAPE)) for _ in range(2))
    D, R = TG(dev, dv), TG(dev, rho)
    got = pxr.get_norm_divErho(D, R, NX, NY, NZ, NG, NG, NG)
    assert got == pytest.approx(orc.norm_diverho(dv, rho, NX, NY, NZ, NG, NG, NG), rel=1e-13)
    got = pxr.get_loc_norm_2(D, NX, NY, NZ, NG, NG, NG)
    assert got == pytest.approx(orc.loc_norm_2(dv, NX, NY, NZ, NG, NG, NG), rel=1e-13)


def test_divergence_argument_errors(dev, pxr):
    from picsar_amd import PxrHipError
    from picsar_amd._lib import ERR_ARG
    a = [TG(dev, np.zeros((6, 6, 6), order="F")) for _ in range(4)]
    with pytest.raises(PxrHipError) as ei:  # no guard: the differences would leave the array
        pxr.calc_field_div(*a, 5, 5, 5, 0, 0, 0, 1.0, 1.0, 1.0)
    assert ei.value.code == ERR_ARG
    with pytest.raises(PxrHipError) as ei:
        pxr.calc_field_divB(*a, 5, 5, 5, 0, 0, 0, 1.0, 1.0, 1.0)
    assert ei.value.code == ERR_ARG


@pytest.mark.parametrize("case", ["", "2"])
def test_sort_golden_bitexact(dev, pxr, golden, case):
    d = golden("sort")
    P = [T(dev, d[f"in{i}"]) for i in range(8)]
    n = len(d["in0"])
    pxr.pxr_particle_bin_sorting(n, *P, 1, *d["box" + case])
    np.testing.assert_array_equal(np.stack([H(t) for t in P]), d["out" + case])


def test_sort_large_bitexact(dev, pxr, orc):
    rng = np.random.default_rng(11)
    n = 150000
    # a 16^3 tile with half-cell membership margins, many ties per cell
    P = [rng.uniform(-0.5, 16.5, n) for _ in range(3)] + [rng.standard_normal(n) for _ in range(4)]
    P.append(np.arange(n, dtype=np.float64))
    box = (-0.5, -0.5, -0.5, 17.5, 17.5, 17.5, 1.0, 1.0, 1.0)
    Q = [p.copy() for p in P]
    assert orc.particle_bin_sorting(*Q, *box) == 0
    D = [T(dev, p) for p in P]
    pxr.pxr_particle_bin_sorting(n, *D, 1, *box)
    for a, b in zip([H(t) for t in D], Q):
        np.testing.assert_array_equal(a, b)


def test_errors_are_reported(dev, pxr):
    from picsar_amd import PxrHipError
    from picsar_amd._lib import ERR_ARG, ERR_BOUNDS
    n = 3
    xp = T(dev, np.array([1.0, 2.0, 500.0]))  # last one far outside the grid
    z = [T(dev, np.array([1.0, 2.0, 3.0])) for _ in range(2)]
    e = [T(dev, np.zeros(n)) for _ in range(6)]
    F = [TG(dev, np.ones(SHAPE, order="F")) for _ in range(6)]
    with pytest.raises(PxrHipError) as ei:
        pxr.geteb3d_energy_conserving(n, xp, *z, *e, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, 1,
                                      1, 1, *F, False, True, 8, 0)
    assert ei.value.code == ERR_BOUNDS
    # the two in-range particles were still gathered
    np.testing.assert_allclose(H(e[0])[:2], 1.0, rtol=1e-15)
    assert H(e[0])[2] == 0.0
    with pytest.raises(PxrHipError) as ei:
        pxr.geteb3d_energy_conserving(n, xp, *z, *e, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, 4,
                                      4, 4, *F, False, True, 8, 0)
    assert ei.value.code == ERR_ARG  # the reference's shapes stop at order 3
    # sort: a key out of range moves nothing
    P = [T(dev, np.array([0.5, 0.5, 9.0])) for _ in range(3)] + \
        [T(dev, np.array([1.0, 2.0, 3.0])) for _ in range(5)]
    with pytest.raises(PxrHipError) as ei:
        pxr.pxr_particle_bin_sorting(3, *P, 1, 0, 0, 0, 2, 2, 2, 1, 1, 1)
    assert ei.value.code == ERR_BOUNDS
    np.testing.assert_array_equal(H(P[7]), [1.0, 2.0, 3.0])
    # np = 0 is a no-op
    pxr.pxr_pushxyz(0, *[T(dev, np.zeros(1)) for _ in range(7)], 0.1)


def test_torch_tensors_interoperate():
    """torch ROCm tensors are accepted in place; the call synchronises the
    device on both sides (torch runs its own HIP runtime).  The check runs
    in a fresh child process that starts torch's runtime before the
    library's (tests/torch_interop_child.py): in this process the library's
    runtime came first and torch would not see the GPU."""
    import os
    import subprocess
    import sys
    pytest.importorskip("torch")
    child = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_interop_child.py")
    r = subprocess.run([sys.executable, child], capture_output=True, text=True, timeout=240)
    out = r.stdout.strip().splitlines()
    assert r.returncode == 0 and out and out[-1].startswith("OK"), (r.returncode, r.stdout,
                                                                     r.stderr[-2000:])
    assert float(out[-1].split()[1]) < REL


def test_convert_f32_f64(dev):
    """pxrhip_convert_f32_f64 widens float32 values exactly."""
    from picsar_amd.devmem import DeviceArray
    rng = np.random.default_rng(31)
    h = (rng.standard_normal(10007) * 1e3).astype(np.float32)
    src = DeviceArray.from_numpy(h)
    dst = DeviceArray(len(h))
    dev.call("pxrhip_convert_f32_f64", src.ptr, dst.ptr, len(h))
    np.testing.assert_array_equal(dst.numpy(), h.astype(np.float64))
    dev.call("pxrhip_convert_f32_f64", src.ptr, dst.ptr, 0)  # n = 0: nothing to do
