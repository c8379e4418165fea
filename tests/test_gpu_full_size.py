"""The fused step at full size: the 128^3-cell thermal plasma, 16 particles
per cell of electrons and of ions (67,108,864 particles), order-3 gather,
Vay, order-3 Esirkepov, Yee, 8^3 tiles of 16^3 nodes, FP64 and mixed
precision.

The oracle cannot run at this size, so the step is checked through
properties that hold at any size:
- the particle count is kept (periodic boundaries);
- every particle lies in the tile whose membership holds it, the tile
  formula of particle_boundaries.F90:397 evaluated on the host, and the
  tiles' sorting bins are packed;
- the discrete Gauss law: div E - rho keeps its initial value to FP64
  round-off (FP32 round-off of the per-particle current in mixed precision),
  as Esirkepov's current satisfies the continuity equation of the
  same-order charge (esirkepov_3d.F90:3263);
- the total energy (kinetic at the half step, field at the whole step)
  changes by less than 1e-6 of itself over the steps.
"""
import numpy as np
import pytest

from tests.test_gpu_step import check_tiles

pytestmark = pytest.mark.gpu

N = 128
PPC = 16
NT = 8
STEPS = 10
OMEGA_PE = 0.02  # electron plasma frequency (c = dx = 1), as in the benchmark


@pytest.fixture(scope="module")
def dev():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return _lib


def thermal(precision):
    from picsar_amd.sim import Simulation
    sim = Simulation(N, N, N, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, NT, NT, NT, order=3,
                     l_lower_order_in_v=True, pusher="vay", depo="esirkepov", norder=2,
                     cfl=0.95, precision=precision)
    rng = np.random.default_rng(2024)
    npart = N ** 3 * PPC
    for name, q, m, uth in (("electrons", -1.0, 1.0, 0.01), ("ions", 1.0, 1836.0, 0.001)):
        pos = [rng.random(npart) * N for _ in range(3)]
        mom = [rng.standard_normal(npart) * uth for _ in range(3)]
        sim.add_species(name, q, m, *pos, *mom, np.full(npart, OMEGA_PE ** 2 / PPC))
        del pos, mom
    return sim


def gauss(sim):
    sim.gauss_residual()
    g = sim.ng
    cut = (slice(g, g + N),) * 3
    dv = sim.diag_grid("dive").numpy()[cut]
    rho = sim.diag_grid("rho").numpy()[cut]
    return dv - rho, float(np.linalg.norm(rho))


def check_membership(sim):
    """Tile of every particle from its position (particle_boundaries.F90:397:
    floor((x - xmin + dx/2) / (n0 dx)), the last tile taking the rest)."""
    n0 = (N + 1) // NT
    for sp in sim.species:
        b = sp.active
        npt = b.tile_np.numpy()
        owner = np.repeat(np.arange(sim.ntiles), npt)
        t = np.zeros(sp.npart, dtype=np.int64)
        for d, key in enumerate(("x", "y", "z")):
            x = b.arrays[key].numpy(sp.npart)
            it = np.minimum(np.floor((x + 0.5) / n0).astype(np.int64), NT - 1)
            t += it * NT ** d
            del x, it
        bad = np.count_nonzero(t != owner)
        assert bad == 0, f"{sp.name}: {bad} particles outside their tile"


@pytest.mark.parametrize("precision,gauss_tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_full_size_step_keeps_invariants(dev, precision, gauss_tol):
    sim = thermal(precision)
    n0 = sim.total_particles()
    assert n0 == 2 * N ** 3 * PPC
    r0, rho0 = gauss(sim)
    e0 = sim.kinetic_energy() + sim.field_energy()
    sim.step(STEPS)
    assert sim.total_particles() == n0
    check_tiles(sim)
    check_membership(sim)
    r1, _ = gauss(sim)
    drift = float(np.linalg.norm(r1 - r0)) / rho0
    assert drift < gauss_tol, drift
    e1 = sim.kinetic_energy() + sim.field_energy()
    assert sim.field_energy() > 0.0  # the fields did evolve
    assert abs(e1 - e0) / e0 < 1e-6, (e0, e1)
