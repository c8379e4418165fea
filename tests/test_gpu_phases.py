"""The phases of the tiled kernel on their own, against the oracle:

  * pxrhip_tiles_gather_push -- field_gathering_plus_particle_pusher
    (particle_pusher_manager_3d.F90:431-672) on a binned species;
  * pxrhip_tiles_depose_currents -- pxrdepose_currents_on_grid_jxjyjz
    (current_deposition_manager_3d.F90:224-505) of a binned species from
    its particles as they stand;
  * frozen species (lfreeze: not pushed) and species that do not deposit
    (ldodepos = .FALSE.), modules.F90:567-570, inside whole steps.
"""
import numpy as np
import pytest

from tests.pic_oracle import OracleState, default_dt, thermal_plasma
from tests.test_gpu_step import build, check_tiles, compare_fields, compare_particles, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)
    return None


@pytest.mark.parametrize("order,depo", [(1, "esirkepov"), (2, "esirkepov"), (3, "esirkepov"),
                                        (1, "direct"), (3, "direct")])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_deposit_phase_matches_the_leaf_deposition(torch, orc, order, depo, precision):
    """Deposit-only launch on a binned species whose momenta carry it up to
    0.6 cell per step: J equals the oracle's depose_jxjyjz of the same
    particles (1e-12 in FP64; FP32 round-off of the per-particle current in
    mixed precision)."""
    idsc = 2.0 ** -15 if precision == "mixed" else 1e-9
    sim, ref, tng = build(torch, orc, order, True, "vay", depo, uth=(0.5, 0.05),
                          precision=precision, id_scale=idsc, ppc=2)
    f = sim.fields
    for k in ("jx", "jy", "jz"):
        f[k].fill(0.0)
        ref.f[k][...] = 0.0
    for sp in sim.species:
        sim.depose_currents_species(sp)
    sim.check()
    for sp, rs in zip(sim.species, ref.species):
        a = sp.arrays()  # the GPU's particles, as binned, in FP64
        for k in ("x", "y", "z", "ux", "uy", "uz", "gaminv", "w"):
            rs[k] = a[k].astype(np.float64)
    ref.deposit(zero_j=True)
    tol = 1e-12 if precision == "fp64" else 1e-6
    for k in ("jx", "jy", "jz"):
        assert rel(tng(f[k]), ref.f[k]) <= tol, (k, rel(tng(f[k]), ref.f[k]))


@pytest.mark.parametrize("order,lo,pusher", [(1, True, "boris"), (2, False, "vay"),
                                             (3, True, "vay"), (3, False, "boris")])
def test_push_phase_matches_the_oracle(torch, orc, order, lo, pusher):
    """Gather-and-push launch without deposition: positions and momenta as
    the oracle's gather + push (1e-12), and J untouched."""
    sim, ref, tng = build(torch, orc, order, lo, pusher, "esirkepov")
    f = sim.fields
    for k in ("jx", "jy", "jz"):
        f[k].fill(0.0)
    for sp, rs in zip(sim.species, ref.species):
        a = sp.arrays()
        for k in ("x", "y", "z", "ux", "uy", "uz", "gaminv", "w"):
            rs[k] = a[k].copy()
    sim.gather_push()
    sim.check()
    ref.gather_push()
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)
    for k in ("jx", "jy", "jz"):
        assert np.abs(tng(f[k])).max() == 0.0


@pytest.mark.parametrize("lfreeze,ldodepos", [(True, True), (True, False), (False, False)])
@pytest.mark.parametrize("split", [False, True])
def test_frozen_and_non_depositing_species(torch, orc, lfreeze, ldodepos, split):
    """A second species that is frozen, does not deposit, or both, next to
    mobile electrons, over whole steps (fused, and in the reference's order
    when an absorbing face splits the step): fields and particles as the
    oracle's, which skips the push of a frozen species and the deposition
    of one that does not deposit.  Its momenta are not zero, so a frozen
    species that deposits puts a current on the grid."""
    from picsar_amd.sim import Simulation, _fd_weights
    nx, ny, nz = 16, 12, 12
    dt = default_dt(1.0, 1.0, 1.0)
    pb = ("absorbing",) + ("periodic",) * 5 if split else None
    sim = Simulation(nx, ny, nz, ntx=2, nty=2, ntz=2, order=3, pusher="vay",
                     depo="esirkepov", dt=dt, pbound=pb)
    assert sim.split_step == split
    ref = OracleState(orc, nx, ny, nz, sim.ng, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, dt, 3, True, "vay",
                      "esirkepov")
    ref.pbound = tuple(sim.grid.pbound)
    ref.nt = (2, 2, 2)
    rng = np.random.default_rng(4)
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        a = np.asfortranarray(rng.standard_normal(ref.f[k].shape) * 0.05)
        orc.field_bc_periodic(a, nx, ny, nz, sim.ng, sim.ng, sim.ng)
        ref.f[k][...] = a
        sim.fields[k].upload(a)
    pe = thermal_plasma(rng, nx, ny, nz, 2, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.3)
    pi = thermal_plasma(rng, nx, ny, nz, 2, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.2, id_offset=10 ** 5)
    ref.add_species(-1.0, 1.0, *pe)
    sim.add_species("e", -1.0, 1.0, *pe, gaminv=ref.species[-1]["gaminv"])
    ref.add_species(1.0, 100.0, *pi, lfreeze=lfreeze, ldodepos=ldodepos)
    sim.add_species("i", 1.0, 100.0, *pi, gaminv=ref.species[-1]["gaminv"], lfreeze=lfreeze,
                    ldodepos=ldodepos)
    step = ref.step_reference_order if split else ref.step
    sim.step(1)
    step()
    check_tiles(sim)
    compare_fields(sim, ref, lambda g: g.numpy(), 1e-12)
    compare_particles(sim, ref, 1e-12 * nx, 1e-12)
    if lfreeze:
        a = sim.species[1].arrays()
        ia = np.argsort(a["w"])
        ib = np.argsort(pi[6])
        np.testing.assert_array_equal(a["x"][ia], pi[0][ib])
        np.testing.assert_array_equal(a["ux"][ia], pi[3][ib])
    sim.step(2)
    step()
    step()
    compare_fields(sim, ref, lambda g: g.numpy(), 1e-9)
