"""The fused step at full size, through properties that hold at any size
(the oracle cannot run at these sizes):
- the 128^3-cell thermal plasma, 16 particles per cell of electrons and of
  ions (67,108,864 particles) in 8^3 tiles of 16^3 nodes;
- the 256^3-cell thermal plasma, 8 per cell of each (268,435,456), in 32^3
  tiles of 8^3 nodes (the configuration of the headline measurement);
- the Gaussian electron bunch (sigma 20 cells, peak 64 ppc, gamma 10 along
  z) in a 1 ppc background on 256^3 cells in 8^3-node tiles, whose tiles
  hold from a few hundred to tens of thousands of particles;
all with the order-3 gather, Vay, order-3 Esirkepov and Yee, in FP64 and in
mixed precision.  Checked after the steps:
- the particle count is kept (periodic boundaries);
- every particle lies in the tile whose membership holds it, the tile
  formula of particle_boundaries.F90:397 evaluated on the host, and the
  tiles' sorting bins are packed;
- the discrete Gauss law: div E - rho keeps its initial value to FP64
  round-off (FP32 round-off of the per-particle current in mixed precision),
  as Esirkepov's current satisfies the continuity equation of the
  same-order charge (esirkepov_3d.F90:3263);
- for the thermal plasmas, the total energy (kinetic at the half step,
  field at the whole step) changes by less than 1e-6 of itself.
"""
import numpy as np
import pytest

from tests.test_gpu_step import check_tiles

pytestmark = pytest.mark.gpu

STEPS = 10
OMEGA_PE = 0.02  # electron plasma frequency (c = dx = 1), as in the benchmark


@pytest.fixture(scope="module")
def dev():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return _lib


def make_sim(n, nt, precision):
    from picsar_amd.sim import Simulation
    return Simulation(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, nt, nt, nt, order=3,
                      l_lower_order_in_v=True, pusher="vay", depo="esirkepov", norder=2,
                      cfl=0.95, precision=precision)


def thermal(n, ppc, nt, precision):
    sim = make_sim(n, nt, precision)
    rng = np.random.default_rng(2024)
    npart = n ** 3 * ppc
    for name, q, m, uth in (("electrons", -1.0, 1.0, 0.01), ("ions", 1.0, 1836.0, 0.001)):
        pos = [rng.random(npart) * n for _ in range(3)]
        mom = [rng.standard_normal(npart) * uth for _ in range(3)]
        sim.add_species(name, q, m, *pos, *mom, np.full(npart, OMEGA_PE ** 2 / ppc))
        del pos, mom
        print(f"{name}: {npart} particles on the device", flush=True)
    return sim


def bunch(n, nt, precision):
    """The benchmark's bunch workload: a gamma = 10 electron bunch (sigma 20
    cells, peak 64 ppc) in a uniform 1 ppc electron background."""
    sim = make_sim(n, nt, precision)
    rng = np.random.default_rng(12345)
    sig, peak, gamma = 20.0, 64.0, 10.0
    nb = int(round(peak * (2 * np.pi) ** 1.5 * sig ** 3))
    c = n / 2.0
    pos = [np.clip(rng.normal(c, sig, nb), 0.0, np.nextafter(n, 0)) for _ in range(3)]
    uz0 = np.sqrt(gamma ** 2 - 1.0)
    mom = [rng.standard_normal(nb) * 0.01, rng.standard_normal(nb) * 0.01,
           uz0 + rng.standard_normal(nb) * 0.01 * uz0]
    sim.add_species("bunch", -1.0, 1.0, *pos, *mom, np.full(nb, OMEGA_PE ** 2 / peak))
    del pos, mom
    npart = n ** 3
    pos = [rng.random(npart) * n for _ in range(3)]
    mom = [rng.standard_normal(npart) * 0.01 for _ in range(3)]
    sim.add_species("background", -1.0, 1.0, *pos, *mom, np.full(npart, OMEGA_PE ** 2))
    return sim


def gauss(sim):
    sim.gauss_residual()
    g = sim.ng
    cut = (slice(g, g + sim.nx), slice(g, g + sim.ny), slice(g, g + sim.nz))
    dv = sim.diag_grid("dive").numpy()[cut]
    rho = sim.diag_grid("rho").numpy()[cut]
    return dv - rho, float(np.linalg.norm(rho))


def check_membership(sim, n, nt):
    """Tile of every particle from its position (particle_boundaries.F90:397:
    floor((x - xmin + dx/2) / (n0 dx)), the last tile taking the rest)."""
    n0 = (n + 1) // nt
    for sp in sim.species:
        b = sp.active
        npt = b.tile_np.numpy()
        owner = np.repeat(np.arange(sim.ntiles), npt)
        t = np.zeros(sp.npart, dtype=np.int64)
        for d, key in enumerate(("x", "y", "z")):
            x = b.arrays[key].numpy(sp.npart)
            it = np.minimum(np.floor((x + 0.5) / n0).astype(np.int64), nt - 1)
            t += it * nt ** d
            del x, it
        bad = np.count_nonzero(t != owner)
        assert bad == 0, f"{sp.name}: {bad} particles outside their tile"
        del owner, t


WORKLOADS = {
    "thermal128": (lambda p: thermal(128, 16, 8, p), 128, 8),
    "thermal256": (lambda p: thermal(256, 8, 32, p), 256, 32),
    "bunch256": (lambda p: bunch(256, 32, p), 256, 32),
}


@pytest.mark.timeout(600)  # building 268M particles on the host takes a minute
@pytest.mark.parametrize("workload", list(WORKLOADS))
@pytest.mark.parametrize("precision,gauss_tol", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_full_size_step_keeps_invariants(dev, workload, precision, gauss_tol):
    make, n, nt = WORKLOADS[workload]
    sim = make(precision)
    n0 = sim.total_particles()
    r0, rho0 = gauss(sim)
    e0 = sim.kinetic_energy() + sim.field_energy()
    sim.step(STEPS)
    print(f"{workload} {precision}: {STEPS} steps done", flush=True)
    assert sim.total_particles() == n0
    check_tiles(sim)
    check_membership(sim, n, nt)
    r1, _ = gauss(sim)
    drift = float(np.linalg.norm(r1 - r0)) / rho0
    assert drift < gauss_tol, drift
    assert sim.field_energy() > 0.0  # the fields did evolve
    if workload.startswith("thermal"):
        e1 = sim.kinetic_energy() + sim.field_energy()
        assert abs(e1 - e0) / e0 < 1e-6, (e0, e1)
