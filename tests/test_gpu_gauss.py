"""Gauss's law through the step, on the GPU.

The reference's acceptance check for its solver and deposition is that
div E*eps0 - rho stays put (Python_tests, |div E - rho| < 1e-5): the Esirkepov
current satisfies the discrete continuity equation with the charge density of
the same shape order, and the Yee E update changes div E by -dt div J.  Here
rho (pxrdepose_rho_on_grid + charge_bcs) and div E (calc_field_div) are taken
on the GPU before and after several whole steps, and the residual
div E - rho must not move beyond round-off on the nodes 0..n-1 of each axis.
(Node n of rho is left out: the periodic fold of one rank adds the node-n
deposit there twice, field_boundaries.F90:431-441, as in the reference.)

The direct deposition does not conserve charge: the same check moves by many
orders of magnitude more, which shows that the check can fail.
"""
import numpy as np
import pytest

from tests.pic_oracle import default_dt, thermal_plasma

pytestmark = pytest.mark.gpu

N = (16, 14, 12)
D = (1.0, 0.9, 1.1)
X0 = (-3.0, 0.5, 2.0)
PPC = 4


@pytest.fixture(scope="module")
def dev():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return _lib


def neutral_plasma(order, depo, pusher="vay", antenna=False, seed=3, precision="fp64"):
    """Electrons and ions at the same positions with the same weights (rho
    = 0 up to rounding), E = 0 and a random periodic B."""
    from picsar_amd.picsarpy import picsar as pxr
    from picsar_amd.sim import Simulation
    sim = Simulation(*N, *D, *X0, 2, 2, 2, order=order, pusher=pusher, depo=depo, norder=2,
                     dt=default_dt(*D), precision=precision)
    rng = np.random.default_rng(seed)
    g = sim.ng
    shape = tuple(n + 1 + 2 * g for n in N)
    for k in ("bx", "by", "bz"):
        a = np.asfortranarray(rng.standard_normal(shape) * 0.05)
        pxr.field_bc(a, g, g, g, *N)
        sim.fields[k].upload(a)
    x, y, z, ux, uy, uz, w = thermal_plasma(rng, *N, PPC, *D, *X0, 0.1)
    sim.add_species("e", -1.0, 1.0, x, y, z, ux, uy, uz, w)
    ui = [rng.standard_normal(len(x)) * 0.01 for _ in range(3)]
    sim.add_species("i", 1.0, 1836.0, x, y, z, *ui, w)
    if antenna:
        sim.add_laser_antenna(spot=(X0[0] + 8.0, X0[1] + 6.1, X0[2] + 5.3), vector=(0, 0, 1),
                              polvector1=(1, 0, 0), lambda_laser=8.0, laser_w0=4.0,
                              laser_ctau=6.0, t_peak=1.5, emax1=0.05, emax2=0.02, charge=1.0)
    return sim


def residual(sim):
    """div E - rho on the nodes 0..n-1 of each axis, and |div E| there."""
    norm = sim.gauss_residual()
    g = sim.ng
    cut = tuple(slice(g, g + n) for n in N)
    dv = sim.diag_grid("dive").numpy()[cut]
    rho = sim.diag_grid("rho").numpy()[cut]
    return dv - rho, dv, norm


# charge density scale: PPC particles of weight ~1 per cell
SCALE = PPC / (D[0] * D[1] * D[2]) * np.sqrt(np.prod(N))


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("antenna", [False, True])
def test_gauss_law_is_kept_by_the_esirkepov_step(dev, order, antenna):
    sim = neutral_plasma(order, "esirkepov", antenna=antenna)
    r0, _, n0 = residual(sim)
    if not antenna:
        assert np.linalg.norm(r0) < 1e-13 * SCALE  # neutral and E = 0
    sim.step(8)
    r1, dv, n1 = residual(sim)
    assert np.linalg.norm(dv) > 1e-4 * SCALE  # the fields did evolve
    drift = np.linalg.norm(r1 - r0) / SCALE
    assert drift < 1e-12, drift
    assert np.isfinite(n0) and np.isfinite(n1)


def test_gauss_law_check_fails_for_the_direct_current(dev):
    sim = neutral_plasma(3, "direct")
    r0, _, _ = residual(sim)
    sim.step(8)
    r1, _, _ = residual(sim)
    assert np.linalg.norm(r1 - r0) / SCALE > 1e-6


def test_managers_compute_rho_every_step(dev):
    """Managers.step calls calc_diags as submain.F90:180 does; rho is then
    that of the present particles and the residual stays put."""
    from picsar_amd.managers import Managers
    sim = neutral_plasma(2, "esirkepov", pusher="boris")
    r0, _, _ = residual(sim)
    mgr = Managers(sim)
    mgr.step(3)
    rho = sim.diag_grid("rho").numpy().copy()
    mgr.calc_diags()
    np.testing.assert_allclose(sim.diag_grid("rho").numpy(), rho, rtol=0, atol=1e-13 * PPC)
    mgr.calc_field_div()
    g = sim.ng
    cut = tuple(slice(g, g + n) for n in N)
    r1 = sim.diag_grid("dive").numpy()[cut] - sim.diag_grid("rho").numpy()[cut]
    assert np.linalg.norm(r1 - r0) / SCALE < 1e-12
    assert mgr.get_norm_divErho() == pytest.approx(sim.get_norm_divErho())


@pytest.mark.parametrize("order", [1, 3])
def test_gauss_law_in_mixed_precision(dev, order):
    """A mixed-precision species deposits an FP32 current per particle from
    FP64 positions; rho (FP64, from the positions and the widened FP32
    weights) then keeps div E - rho to FP32 round-off of the current, not
    FP64: the bound is 1e-5 of the charge-density scale after eight steps
    (measured: about 1e-8), against 1e-12 in FP64."""
    sim = neutral_plasma(order, "esirkepov", precision="mixed")
    r0, _, _ = residual(sim)
    assert np.linalg.norm(r0) < 1e-13 * SCALE
    sim.step(8)
    r1, dv, _ = residual(sim)
    assert np.linalg.norm(dv) > 1e-4 * SCALE
    assert np.linalg.norm(r1 - r0) / SCALE < 1e-5


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tiled_rho_matches_the_leaf_routine(dev, order, precision):
    """pxrhip_tiles_depose_rho (LDS blocks per work item) against
    pxr_depose_rho_n over the same particles, binned and then after a push
    that moved them out of their bins (deposited from global memory)."""
    import ctypes as C
    from picsar_amd import _lib
    from picsar_amd.devmem import DeviceArray, DeviceGrid
    sim = neutral_plasma(order, "esirkepov", precision=precision)
    g = sim.ng

    def leaf():
        r = DeviceGrid(*N, g)
        for sp in sim.species:
            a = sp.arrays()
            dev_arr = [DeviceArray.from_numpy(np.ascontiguousarray(a[k], dtype=np.float64))
                       for k in ("x", "y", "z", "w")]
            _lib.call("pxrhip_depose_rho_n", r.ptr, sp.npart, *(d.ptr for d in dev_arr), sp.q,
                      sim.xmin, sim.ymin, sim.zmin, sim.dx, sim.dy, sim.dz, *N, g, g, g, order,
                      order, order, 1, 0)
        return r.numpy()

    def tiled():
        r = DeviceGrid(*N, g)
        for sp in sim.species:
            s = sp.struct(sp.cur)
            _lib.call("pxrhip_tiles_depose_rho", C.byref(sim.grid), C.byref(s), r.ptr, order)
        _lib.call("pxrhip_tiles_check")
        return r.numpy()

    # electrons only, so that rho does not cancel
    sim.species[1].ldodepos = False
    sim.species = sim.species[:1]
    for stage in ("binned", "moved"):
        want, got = leaf(), tiled()
        scale = np.abs(want).max()
        assert scale > 0
        assert np.abs(got - want).max() <= 1e-13 * scale, (stage, np.abs(got - want).max() / scale)
        sim.gather_push_deposit()  # positions move; the bins are not rebuilt
        sim.check()
