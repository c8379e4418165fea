"""Edge cases of the GPU path: empty and ragged inputs.

The leaf routines with np = 0 change nothing and raise nothing, as the
reference's loops run zero times (field_gathering_manager_3d.F90:418 returns
at np = 0).  The tiled step is compared with the oracle step at 1e-12 with a
species that has no particles, with every particle in one tile of eight, and
with a single particle.
"""
import numpy as np
import pytest

from tests.pic_oracle import OracleState, default_dt, thermal_plasma
from tests.test_gpu_step import check_tiles, compare_fields, compare_particles

pytestmark = pytest.mark.gpu

N = (16, 14, 12)
D = (1.0, 0.9, 1.1)
X0 = (-3.0, 0.5, 2.0)


@pytest.fixture(scope="module")
def dev():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return _lib


def test_leaf_routines_accept_zero_particles(dev, orc):
    from picsar_amd.picsarpy import picsar as pxr
    nx, ny, nz, g = 6, 5, 4, 3
    shp = (nx + 1 + 2 * g, ny + 1 + 2 * g, nz + 1 + 2 * g)
    rng = np.random.default_rng(2)
    F = [np.asfortranarray(rng.standard_normal(shp)) for _ in range(6)]
    P = [np.array([0.5]) for _ in range(8)]  # one element each; np = 0 reads none
    E = [np.array([7.0]) for _ in range(6)]
    for order in (1, 2, 3):
        pxr.geteb3d_energy_conserving(0, *P[:3], *E, 0, 0, 0, 1, 1, 1, nx, ny, nz, g, g, g,
                                      order, order, order, *F, False, True, 8, 0)
    assert all(e[0] == 7.0 for e in E)
    pxr.pxr_boris_push_u_3d(0, *P[3:7], *E, -1.0, 1.0, 0.1)
    pxr.pxr_ebcancelpush3d(0, *P[3:7], *E, -1.0, 1.0, 0.1, 0)
    pxr.pxr_pushxyz(0, *P[:7], 0.1)
    pxr.pxr_set_gamma(0, *P[3:7])
    assert all(p[0] == 0.5 for p in P)
    J = [np.asfortranarray(np.zeros(shp)) for _ in range(3)]
    for algo in (0, 3):
        for order in (1, 2, 3):
            pxr.depose_jxjyjz(*J, 0, *P, -1.0, 0, 0, 0, 0.1, 1, 1, 1, nx, ny, nz, g, g, g,
                              order, order, order, algo)
    rho = np.asfortranarray(np.zeros(shp))
    pxr.pxr_depose_rho_n(rho, 0, *P[:3], P[7], -1.0, 0, 0, 0, 1, 1, 1, nx, ny, nz, g, g, g, 3,
                         3, 3)
    assert not any(a.any() for a in J) and not rho.any()
    pxr.pxr_particle_bin_sorting(0, *P[:7], P[7], 1, 0, 0, 0, 4, 4, 4, 1, 1, 1)
    assert all(p[0] == 0.5 for p in P)


def make(orc, species, ntiles=(2, 2, 2), order=3, depo="esirkepov", seed=1):
    """A Simulation and its oracle with random periodic fields and the given
    species [(q, m, (x, y, z, ux, uy, uz, w)), ...]."""
    from picsar_amd.sim import Simulation
    dt = default_dt(*D)
    sim = Simulation(*N, *D, *X0, *ntiles, order=order, pusher="vay", depo=depo, dt=dt)
    ref = OracleState(orc, *N, sim.ng, *D, *X0, dt, order, True, "vay", depo)
    ref.pbound = tuple(sim.grid.pbound)
    ref.nt = tuple(ntiles)
    rng = np.random.default_rng(seed)
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        a = np.asfortranarray(rng.standard_normal(ref.f[k].shape) * 0.05)
        orc.field_bc_periodic(a, *N, sim.ng, sim.ng, sim.ng)
        ref.f[k][...] = a
        sim.fields[k].upload(a)
    for i, (q, m, p) in enumerate(species):
        ref.add_species(q, m, *p)
        sim.add_species(f"s{i}", q, m, *p, gaminv=ref.species[-1]["gaminv"])
    return sim, ref


def plasma(rng, ppc, uth, id_offset=0):
    return thermal_plasma(rng, *N, ppc, *D, *X0, uth, id_offset=id_offset)


def run_and_compare(sim, ref, steps=2):
    for _ in range(steps):
        sim.step(1)
        ref.step()
    check_tiles(sim)
    compare_fields(sim, ref, lambda g: g.numpy(), 1e-12)
    L = max(n * d for n, d in zip(N, D))
    compare_particles(sim, ref, 1e-12 * L, 1e-12)


@pytest.mark.parametrize("depo", ["esirkepov", "direct"])
def test_step_with_an_empty_species(dev, orc, depo):
    rng = np.random.default_rng(3)
    e = plasma(rng, 4, 0.2)
    empty = tuple(np.zeros(0) for _ in range(7))
    sim, ref = make(orc, [(-1.0, 1.0, e), (1.0, 1836.0, empty)], depo=depo)
    assert sim.species[1].npart == 0
    run_and_compare(sim, ref)
    assert sim.species[1].npart == 0
    assert sim.kinetic_energy() > 0


def test_step_with_every_particle_in_one_tile(dev, orc):
    """Seven of the eight tiles hold nothing; the eighth holds a dense
    clump, some of which leaves it during the steps."""
    rng = np.random.default_rng(4)
    n = 3000
    # cells 1.5 to 4.5 of each axis: inside the first tile's membership
    pos = [X0[d] + D[d] * rng.uniform(1.5, 4.5, n) for d in range(3)]
    u = [rng.standard_normal(n) * 0.4 for _ in range(3)]
    w = 1.0 + np.arange(n) * 1e-9
    sim, ref = make(orc, [(-1.0, 1.0, (*pos, *u, w))])
    np_tiles = sim.species[0].active.tile_np.numpy()
    assert (np_tiles > 0).sum() == 1
    run_and_compare(sim, ref, steps=3)


@pytest.mark.parametrize("order", [1, 3])
def test_step_with_a_single_particle(dev, orc, order):
    p = tuple(np.array([v]) for v in (X0[0] + 7.3, X0[1] + 6.2, X0[2] + 5.9, 0.3, -0.2, 0.5,
                                      1.0))
    sim, ref = make(orc, [(-1.0, 1.0, p)], order=order)
    run_and_compare(sim, ref, steps=3)
    assert sim.species[0].npart == 1
