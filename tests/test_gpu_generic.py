"""GPU parity of the generic-order leaf routines: one shape order per axis.

The reference's dispatchers call pxr_gete3d_n/pxr_getb3d_n_energy_conserving
(field_gathering_on_3d.F90:697, 1042), pxr_depose_jxjyjz_esirkepov_n
(esirkepov_3d.F90:3480) and pxr_depose_rho_n (charge_deposition_3d.F90:2081)
for unequal orders, order 0 (deposition), field_gathe_algo 2 and
current_depo_algo 1 (field_gathering_manager_3d.F90:397-700,
current_deposition_manager_3d.F90:106-208).  Every call goes through the C
ABI.  Bars: the gather is bit-exact against the golden vectors and the
oracle; the deposits hold every particle's contributions bit for bit and
differ only in the order of the FP64 atomic sums, so they agree to 1e-13
relative in norm.
"""
import numpy as np
import pytest

from tests.test_gpu_leaf import NG, NX, NY, NZ, SHAPE, H, HG, T, TG, particles, rel

pytestmark = pytest.mark.gpu

DEP_REL = 1e-13


@pytest.fixture(scope="module")
def dev():
    from picsar_amd import _lib
    _lib.load()
    _lib.call("pxrhip_init", 0)
    return _lib


@pytest.fixture(scope="module")
def pxr(dev):
    from picsar_amd.picsarpy import picsar
    picsar.l_nodalgrid = False
    return picsar


def test_generic_gather_golden_bitexact(dev, pxr, golden, ggrid):
    d = golden("generic")
    g = ggrid
    n = len(d["g_xp"])
    F = [TG(dev, d[k]) for k in ("exg", "eyg", "ezg", "bxg", "byg", "bzg")]
    P = [T(dev, d[k]) for k in ("g_xp", "g_yp", "g_zp")]
    for c, (nox, noy, noz, lo, nod, l4) in enumerate(d["gather_cases"]):
        pxr.l_nodalgrid = bool(nod)
        try:
            e = [T(dev, np.zeros(n)) for _ in range(6)]
            pxr.geteb3d_energy_conserving(n, *P, *e, g["xmin"], g["ymin"], g["zmin"], g["dx"],
                                          g["dy"], g["dz"], g["nx"], g["ny"], g["nz"], g["ng"],
                                          g["ng"], g["ng"], int(nox), int(noy), int(noz), *F,
                                          bool(l4), bool(lo), 8, 2)
        finally:
            pxr.l_nodalgrid = False
        np.testing.assert_array_equal(np.stack([H(t) for t in e]), d[f"gather{c}"],
                                      err_msg=f"case {c}")


@pytest.mark.parametrize("orders", [(1, 2, 3), (3, 1, 2), (2, 3, 1), (3, 3, 1), (2, 2, 2)])
@pytest.mark.parametrize("lo,nodal,l4", [(True, False, False), (False, False, False),
                                         (True, True, False), (True, False, True)])
def test_generic_gather_against_oracle(dev, pxr, orc, orders, lo, nodal, l4):
    """40,000 particles, accumulating into non-zero outputs; with l4symtry the
    particles span both signs of x and y.  Equal orders go through
    field_gathe_algo 2, unequal ones through the default algorithm."""
    rng = np.random.default_rng(sum(o * 4 ** a for a, o in enumerate(orders)))
    n = 40000
    dx, dy, dz = 0.9, 1.0, 1.2
    x, y, z = particles(rng, n, margin=1.0)
    xmin, ymin, zmin = 0.3, -1.0, 2.0
    if l4:  # both signs of x and y, which the routine folds onto x, y >= 0
        x = rng.uniform(-(NX - 1.0), NX - 1.0, n)
        y = rng.uniform(-(NY - 1.0), NY - 1.0, n)
        xmin, ymin = 0.0, 0.0
    xp, yp, zp = xmin + x * dx, ymin + y * dy, zmin + z * dz
    F = [np.asfortranarray(rng.standard_normal(SHAPE)) for _ in range(6)]
    e0 = [rng.standard_normal(n) for _ in range(6)]
    er = [a.copy() for a in e0]
    assert orc.geteb3d_n(xp, yp, zp, *er, xmin, ymin, zmin, dx, dy, dz, NX, NY, NZ, NG, NG, NG,
                         *orders, *F, l4symtry=l4, l_lower_order_in_v=lo, l_nodal=nodal) == 0
    algo = 2 if len(set(orders)) == 1 else 0
    e = [T(dev, a) for a in e0]
    pxr.l_nodalgrid = nodal
    try:
        pxr.geteb3d_energy_conserving(n, T(dev, xp), T(dev, yp), T(dev, zp), *e, xmin, ymin,
                                      zmin, dx, dy, dz, NX, NY, NZ, NG, NG, NG, *orders,
                                      *[TG(dev, f) for f in F], l4, lo, 8, algo)
    finally:
        pxr.l_nodalgrid = False
    np.testing.assert_array_equal(np.stack([H(t) for t in e]), np.stack(er))


def test_generic_deposit_golden(dev, pxr, golden, ggrid):
    """current_depo_algo 1 runs pxr_depose_jxjyjz_esirkepov_n for every
    order, the equal ones included."""
    d = golden("generic")
    g = ggrid
    n = len(d["d_xp"])
    P = [T(dev, d[k]) for k in ("d_xp", "d_yp", "d_zp", "ux", "uy", "uz", "gaminv", "w")]
    for c, o in enumerate(d["esirk_cases"]):
        ref = d[f"esirk{c}"]
        J = [TG(dev, np.zeros(ref.shape[1:], order="F")) for _ in range(3)]
        pxr.depose_jxjyjz(*J, n, *P, float(d["q"]), g["xmin"], g["ymin"], g["zmin"],
                          float(d["dt"]), g["dx"], g["dy"], g["dz"], g["nx"], g["ny"], g["nz"],
                          g["ng"], g["ng"], g["ng"], *map(int, o), 1)
        got = np.stack([HG(t) for t in J])
        assert rel(got, ref) < DEP_REL, (tuple(o), rel(got, ref))


@pytest.mark.parametrize("orders,algo", [((0, 1, 2), 0), ((1, 2, 3), 0), ((3, 2, 1), 2),
                                         ((3, 3, 0), 3), ((0, 0, 0), 0), ((3, 3, 3), 1),
                                         ((2, 2, 2), 1)])
def test_generic_deposit_against_oracle(dev, pxr, orc, orders, algo):
    rng = np.random.default_rng(400 + sum(o * 4 ** a for a, o in enumerate(orders)))
    n = 40000
    dt, dx, dy, dz = 0.45, 0.9, 1.0, 1.2
    xmin, ymin, zmin = 0.3, -1.0, 2.0
    x, y, z = particles(rng, n, margin=1.0)
    xp, yp, zp = xmin + x * dx, ymin + y * dy, zmin + z * dz
    u = [rng.uniform(-1.5, 1.5, n) for _ in range(3)]
    gi = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 2.0, n)
    q = -1.1
    J0 = [np.asfortranarray(rng.standard_normal(SHAPE) * 0.1) for _ in range(3)]
    Jr = [j.copy(order="F") for j in J0]
    assert orc.depose_esirkepov_n(*Jr, xp, yp, zp, *u, gi, w, q, xmin, ymin, zmin, dt, dx, dy,
                                  dz, NX, NY, NZ, NG, NG, NG, *orders) == 0
    J = [TG(dev, j) for j in J0]
    pxr.depose_jxjyjz(*J, n, *[T(dev, a) for a in (xp, yp, zp, *u, gi, w)], q, xmin, ymin, zmin,
                      dt, dx, dy, dz, NX, NY, NZ, NG, NG, NG, *orders, algo)
    for c in range(3):
        got = HG(J[c])
        assert rel(got - J0[c], Jr[c] - J0[c]) < DEP_REL, (c, rel(got - J0[c], Jr[c] - J0[c]))


def test_generic_rho_golden(dev, pxr, golden, ggrid):
    d = golden("generic")
    g = ggrid
    n = len(d["r_xp"])
    P = [T(dev, d[k]) for k in ("r_xp", "r_yp", "r_zp")]
    for c, (nox, noy, noz, lpw, l4) in enumerate(d["rho_cases"]):
        ref = d[f"rho{c}"]
        R = TG(dev, np.zeros(ref.shape, order="F"))
        w = d["w"] if lpw else d["w"][:1]
        pxr.pxr_depose_rho_n(R, n, *P, T(dev, w), 0.7, g["xmin"], g["ymin"], g["zmin"], g["dx"],
                             g["dy"], g["dz"], g["nx"], g["ny"], g["nz"], g["ng"], g["ng"],
                             g["ng"], int(nox), int(noy), int(noz), bool(lpw), bool(l4))
        assert rel(HG(R), ref) < DEP_REL, (c, rel(HG(R), ref))


@pytest.mark.parametrize("orders", [(0, 1, 2), (3, 2, 1), (0, 0, 0), (1, 3, 3)])
def test_generic_rho_against_oracle(dev, pxr, orc, orders):
    rng = np.random.default_rng(500 + sum(o * 4 ** a for a, o in enumerate(orders)))
    n = 40000
    dx, dy, dz = 0.9, 1.0, 1.2
    xmin, ymin, zmin = 0.3, -1.0, 2.0
    x, y, z = particles(rng, n, margin=1.0)
    xp, yp, zp = xmin + x * dx, ymin + y * dy, zmin + z * dz
    w = rng.uniform(0.5, 2.0, n)
    r0 = np.asfortranarray(rng.standard_normal(SHAPE) * 0.1)
    rr = r0.copy(order="F")
    assert orc.depose_rho_n(rr, xp, yp, zp, w, -1.1, xmin, ymin, zmin, dx, dy, dz, NX, NY, NZ,
                            NG, NG, NG, *orders) == 0
    R = TG(dev, r0)
    pxr.pxr_depose_rho_n(R, n, *[T(dev, a) for a in (xp, yp, zp, w)], -1.1, xmin, ymin, zmin,
                         dx, dy, dz, NX, NY, NZ, NG, NG, NG, *orders)
    assert rel(HG(R) - r0, rr - r0) < DEP_REL


def test_orders_without_weights_are_rejected(dev, pxr):
    """The reference's shape tables stop at order 3 (and the gather's start at
    order 1): such orders are argument errors, and nothing is written."""
    from picsar_amd import PxrHipError
    from picsar_amd._lib import ERR_ARG
    n = 4
    xp = [T(dev, np.full(n, 5.0)) for _ in range(3)]
    e = [T(dev, np.zeros(n)) for _ in range(6)]
    F = [TG(dev, np.ones(SHAPE, order="F")) for _ in range(6)]
    for o in ((4, 4, 4), (1, 0, 1), (3, 3, 4)):
        with pytest.raises(PxrHipError) as ei:
            pxr.geteb3d_energy_conserving(n, *xp, *e, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG,
                                          *o, *F, False, True, 8, 0)
        assert ei.value.code == ERR_ARG
    assert all(not H(t).any() for t in e)
    J = [TG(dev, np.zeros(SHAPE, order="F")) for _ in range(3)]
    u = [T(dev, np.zeros(n)) for _ in range(3)]
    one = T(dev, np.ones(n))
    with pytest.raises(PxrHipError) as ei:
        pxr.depose_jxjyjz(*J, n, *xp, *u, one, one, 1.0, 0, 0, 0, 0.5, 1, 1, 1, NX, NY, NZ, NG,
                          NG, NG, 4, 1, 1, 0)
    assert ei.value.code == ERR_ARG
    R = TG(dev, np.zeros(SHAPE, order="F"))
    with pytest.raises(PxrHipError) as ei:
        pxr.pxr_depose_rho_n(R, n, *xp, one, 1.0, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG,
                             1, 2, 5)
    assert ei.value.code == ERR_ARG
    assert not any(HG(t).any() for t in J) and not HG(R).any()
