"""The binned form of the leaf deposition (pxr_leafbin.hip): depose_jxjyjz
with the Esirkepov current and pxr_depose_rho_n, for large particle sets,
count the particles into blocks of nodes and sum each block's
contributions in LDS before one global add per node.  Called through the C
ABI like any caller of the leaf routines, with the runtime's switch-over
(`leaf.bin_min`) set so that both forms run on the same inputs.

Bars: against the oracle 1e-12 relative in norm (the contributions are the
same arithmetic; only the order of the FP64 sums differs), against the
global-atomic form 1e-13; particles outside the arrays are skipped and
reported by both forms alike.
"""
import numpy as np
import pytest

from tests.test_gpu_leaf import NG, NX, NY, NZ, SHAPE, H, HG, T, TG, particles, rel

pytestmark = pytest.mark.gpu


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


@pytest.fixture
def binned(dev):
    """set leaf.bin_min for one test, then restore it"""
    from picsar_amd import _lib
    import ctypes
    old = ctypes.c_int64()
    _lib.call("pxrhip_get_tuning", b"leaf.bin_min", ctypes.byref(old))

    def set_min(v):
        _lib.call("pxrhip_set_tuning", b"leaf.bin_min", int(v))
    yield set_min
    set_min(old.value)


def plasma(rng, n, sort=False, guard=0.0):
    """n particles with velocities up to 0.8 c; optionally ordered by cell as
    a tiled code holds them; `guard` cells of the guard region included."""
    dx, dy, dz = 0.9, 1.0, 1.2
    xmin, ymin, zmin = 0.3, -1.0, 2.0
    x, y, z = particles(rng, n, lo=-guard, margin=1.0 - guard)
    if sort:
        o = np.lexsort((np.floor(x), np.floor(y), np.floor(z)))
        x, y, z = x[o], y[o], z[o]
    u = [rng.uniform(-1.5, 1.5, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 2.0, n)
    geo = (xmin, ymin, zmin, dx, dy, dz)
    return (xmin + x * dx, ymin + y * dy, zmin + z * dz), u, g, w, geo


def gpu_current(pxr, dev, P, u, g, w, geo, order, J0, q=-1.1, dt=0.45):
    xmin, ymin, zmin, dx, dy, dz = geo
    J = [TG(dev, j) for j in J0]
    pxr.depose_jxjyjz(*J, len(w), *[T(dev, a) for a in (*P, *u, g, w)], q, xmin, ymin, zmin,
                      dt, dx, dy, dz, NX, NY, NZ, NG, NG, NG, order, order, order, 0)
    return [HG(j) for j in J]


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("sort", [False, True])
def test_binned_current_against_oracle_and_atomic_form(dev, pxr, orc, binned, order, sort):
    rng = np.random.default_rng(700 + order + 10 * sort)
    n = 120000
    P, u, g, w, geo = plasma(rng, n, sort=sort, guard=1.0)
    xmin, ymin, zmin, dx, dy, dz = geo
    q, dt = -1.1, 0.45
    J0 = [np.asfortranarray(rng.standard_normal(SHAPE) * 0.1) for _ in range(3)]
    Jr = [j.copy(order="F") for j in J0]
    assert orc.depose_esirkepov(*Jr, *P, *u, g, w, q, xmin, ymin, zmin, dt, dx, dy, dz, NX, NY,
                                NZ, NG, NG, NG, order) == 0
    binned(1)
    Jb = gpu_current(pxr, dev, P, u, g, w, geo, order, J0, q, dt)
    binned(0)
    Ja = gpu_current(pxr, dev, P, u, g, w, geo, order, J0, q, dt)
    for c in range(3):
        assert rel(Jb[c] - J0[c], Jr[c] - J0[c]) < 1e-12, (c, rel(Jb[c] - J0[c], Jr[c] - J0[c]))
        assert rel(Jb[c] - J0[c], Ja[c] - J0[c]) < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("lpw,l4", [(True, False), (False, False), (True, True)])
def test_binned_rho_against_oracle(dev, pxr, orc, binned, order, lpw, l4):
    rng = np.random.default_rng(800 + order)
    n = 100000
    P, u, g, w, geo = plasma(rng, n, sort=True, guard=1.0)
    xmin, ymin, zmin, dx, dy, dz = geo
    xp, yp, zp = P
    if l4:  # both signs of x and y, folded by the routine
        xmin, ymin = 0.0, 0.0
        xp = rng.uniform(-(NX - 1.0), NX - 1.0, n) * dx
        yp = rng.uniform(-(NY - 1.0), NY - 1.0, n) * dy
    r0 = np.asfortranarray(rng.standard_normal(SHAPE) * 0.1)
    rr = r0.copy(order="F")
    assert orc.depose_rho_n(rr, xp, yp, zp, w, -1.1, xmin, ymin, zmin, dx, dy, dz, NX, NY, NZ,
                            NG, NG, NG, order, order, order, lpw, l4) == 0
    binned(1)
    R = TG(dev, r0)
    ww = w if lpw else w[:1]
    pxr.pxr_depose_rho_n(R, n, *[T(dev, a) for a in (xp, yp, zp, ww)], -1.1, xmin, ymin, zmin,
                         dx, dy, dz, NX, NY, NZ, NG, NG, NG, order, order, order, lpw, l4)
    assert rel(HG(R) - r0, rr - r0) < 1e-12


def test_binned_forms_skip_and_report_particles_outside(dev, pxr, orc, binned):
    """Particles whose stencil leaves the arrays are skipped and reported
    (PXRHIP_ERR_BOUNDS) as by the global-atomic form; the others deposit."""
    from picsar_amd import PxrHipError
    from picsar_amd._lib import ERR_BOUNDS
    rng = np.random.default_rng(900)
    n = 50000
    P, u, g, w, geo = plasma(rng, n)
    xmin, ymin, zmin, dx, dy, dz = geo
    P = [a.copy() for a in P]
    P[0][:7] = xmin + (NX + NG + 5) * dx     # beyond the guards
    P[2][7:9] = zmin - (NG + 3) * dz
    keep = np.ones(n, dtype=bool)
    keep[:9] = False
    J0 = [np.zeros(SHAPE, order="F") for _ in range(3)]
    Jr = [np.zeros(SHAPE, order="F") for _ in range(3)]
    assert orc.depose_esirkepov(*Jr, *[a[keep] for a in P], *[a[keep] for a in u], g[keep],
                                w[keep], -1.1, xmin, ymin, zmin, 0.45, dx, dy, dz, NX, NY, NZ,
                                NG, NG, NG, 3) == 0
    binned(1)
    J = [TG(dev, j) for j in J0]
    with pytest.raises(PxrHipError) as ei:
        pxr.depose_jxjyjz(*J, n, *[T(dev, a) for a in (*P, *u, g, w)], -1.1, xmin, ymin, zmin,
                          0.45, dx, dy, dz, NX, NY, NZ, NG, NG, NG, 3, 3, 3, 0)
    assert ei.value.code == ERR_BOUNDS and "9 particle" in str(ei.value)
    for c in range(3):
        assert rel(HG(J[c]), Jr[c]) < 1e-12
    R = TG(dev, np.zeros(SHAPE, order="F"))
    rr = np.zeros(SHAPE, order="F")
    assert orc.depose_rho_n(rr, *[a[keep] for a in P], w[keep], 1.0, xmin, ymin, zmin, dx, dy,
                            dz, NX, NY, NZ, NG, NG, NG, 2, 2, 2) == 0
    with pytest.raises(PxrHipError) as ei:
        pxr.pxr_depose_rho_n(R, n, *[T(dev, a) for a in (*P, w)], 1.0, xmin, ymin, zmin, dx, dy,
                             dz, NX, NY, NZ, NG, NG, NG, 2, 2, 2)
    assert ei.value.code == ERR_BOUNDS
    assert rel(HG(R), rr) < 1e-12


def test_binned_current_many_work_items(dev, pxr, orc, binned):
    """Two million particles on a 64^3 grid, crowded into a corner so that
    blocks hold tens of thousands of particles (many work items of one
    block) next to empty ones; against the oracle."""
    from picsar_amd import _lib
    import ctypes
    rng = np.random.default_rng(901)
    n, m, ng = 2_000_000, 64, 3
    shape = (m + 1 + 2 * ng,) * 3
    x = np.concatenate([rng.uniform(0, 12, n // 2), rng.uniform(0, m, n - n // 2)])
    y = np.concatenate([rng.uniform(0, 12, n // 2), rng.uniform(0, m, n - n // 2)])
    z = np.concatenate([rng.uniform(0, 12, n // 2), rng.uniform(0, m, n - n // 2)])
    u = [rng.uniform(-1.0, 1.0, n) for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 2.0, n)
    Jr = [np.zeros(shape, order="F") for _ in range(3)]
    assert orc.depose_esirkepov(*Jr, x, y, z, *u, g, w, 0.7, 0, 0, 0, 0.5, 1, 1, 1, m, m, m, ng,
                                ng, ng, 3) == 0
    binned(1)
    old = ctypes.c_int64()
    _lib.call("pxrhip_get_tuning", b"leaf.bin_chunk", ctypes.byref(old))
    try:
        for chunk in (4096, 1000):
            _lib.call("pxrhip_set_tuning", b"leaf.bin_chunk", chunk)
            J = [TG(dev, np.zeros(shape, order="F")) for _ in range(3)]
            pxr.depose_jxjyjz(*J, n, *[T(dev, a) for a in (x, y, z, *u, g, w)], 0.7, 0, 0, 0, 0.5,
                              1, 1, 1, m, m, m, ng, ng, ng, 3, 3, 3, 0)
            for c in range(3):
                assert rel(HG(J[c]), Jr[c]) < 1e-12, (chunk, c)
    finally:
        _lib.call("pxrhip_set_tuning", b"leaf.bin_chunk", old.value)
