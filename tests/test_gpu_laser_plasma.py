"""The laser-driven underdense plasma, scaled down, against the oracle.

The large single-GPU configuration (512 x 512 x 1024 cells) combines every
feature of the step that the smaller configurations leave out: the
order-4 FDTD (yee.F90:46-133, 752-833 with FD_weights, submain.F90:927-984),
a laser antenna in the plane z = 0 (laser_pusher_manager_3d.F90:15-271,
tiling.F90:1737-1919), electrons whose density ramps up linearly, immobile
ions (lfreeze, ldodepos = .FALSE., modules.F90:567-570), the Vay pusher,
order-3 Esirkepov deposition, 8^3-node tiles, and mixed precision.  Here the
same set-up runs on 32 x 32 x 64 cells (ramp over 8 cells to 4 ppc,
lambda = 8 cells, waist 8 cells, a0 = 2):

  * FP64: E, B, J, positions and momenta at 1e-12 (relative) after one
    step, fields at 1e-9 after three;
  * mixed precision, against the same FP64 oracle started from the same
    FP64 values (so the FP32 storage of the momenta and weights is part of
    the error): J, E and B within 1e-5 after one step and 1e-4 after three.
"""
import numpy as np
import pytest

from tests.pic_oracle import OracleState
from tests.test_gpu_step import _antenna_oracle_dict, compare_particles, rel

pytestmark = pytest.mark.gpu

NX, NY, NZ = 32, 32, 64
LAMBDA, PPC, Z0, RAMP = 8.0, 4, 4, 8


@pytest.fixture(scope="module")
def torch():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)
    return None


def electrons(rng):
    """4 ppc from plane Z0 on, density rising linearly over RAMP cells to
    0.1 of the critical density; weights times (1 + id * 1e-9) carry ids."""
    k0 = 2 * np.pi / LAMBDA
    n_plateau = 0.1 * k0 ** 2  # omega^2 = n_c with m = q = eps0 = c = 1
    cells = np.arange(NX * NY * (NZ - Z0))
    c = np.stack([cells % NX, (cells // NX) % NY, Z0 + cells // (NX * NY)])
    c = np.repeat(c, PPC, axis=1).astype(np.float64)
    m = c.shape[1]
    x, y, z = (c[a] + rng.random(m) for a in range(3))
    dens = n_plateau * np.clip((z - Z0) / RAMP, 0.0, 1.0)
    u = [rng.standard_normal(m) * 0.02 for _ in range(3)]
    w = dens / PPC * (1.0 + np.arange(m) * 1e-9)
    keep = w > 0
    return [a[keep] for a in (x, y, z, *u, w)]


def build(torch, orc, precision):
    from picsar_amd.laser import emax_from_a0
    from picsar_amd.sim import Simulation, _fd_weights
    sim = Simulation(NX, NY, NZ, ntx=(NX + 1) // 8, nty=(NY + 1) // 8, ntz=(NZ + 1) // 8,
                     order=3, l_lower_order_in_v=True, pusher="vay", depo="esirkepov", norder=4,
                     precision=precision)
    ref = OracleState(orc, NX, NY, NZ, sim.ng, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, sim.dt, 3, True,
                      "vay", "esirkepov", 4, _fd_weights(4))
    ref.nt = (sim.grid.ntx, sim.grid.nty, sim.grid.ntz)
    rng = np.random.default_rng(21)
    pe = electrons(rng)
    assert len(np.unique(pe[6])) == len(pe[6])
    ref.add_species(-1.0, 1.0, *pe)
    sim.add_species("electrons", -1.0, 1.0, *pe, gaminv=ref.species[-1]["gaminv"])
    # immobile ions: one per cell of the plasma, at rest, neither pushed nor depositing
    ni = NX * NY * (NZ - Z0)
    pi = [rng.random(ni) * NX, rng.random(ni) * NY, Z0 + rng.random(ni) * (NZ - Z0)]
    zero = np.zeros(ni)
    wi = 0.1 * (2 * np.pi / LAMBDA) ** 2 * np.clip((pi[2] - Z0) / RAMP, 0.0, 1.0)
    wi = wi * (1.0 + np.arange(ni) * 1e-9)  # ids
    assert len(np.unique(wi)) == ni
    ref.add_species(1.0, 1836.0, *pi, zero, zero, zero, wi, lfreeze=True, ldodepos=False)
    sim.add_species("ions", 1.0, 1836.0, *pi, zero, zero, zero, wi, lfreeze=True,
                    ldodepos=False)
    ant = sim.add_laser_antenna(spot=(NX / 2.0, NY / 2.0, 0.0), vector=(0, 0, 1),
                                polvector1=(1, 0, 0), lambda_laser=LAMBDA, laser_w0=8.0,
                                laser_ctau=2 * LAMBDA, t_peak=4.0,
                                emax1=emax_from_a0(2.0, LAMBDA), charge=-1.0)
    assert ant.npart == NX * NY
    ref.add_antenna(_antenna_oracle_dict(ant))
    return sim, ref


def field_errors(sim, ref):
    return {k: rel(sim.fields[k].numpy(), ref.f[k])
            for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")}


def test_laser_plasma_fp64(torch, orc):
    sim, ref = build(torch, orc, "fp64")
    sim.step(1)
    ref.step()
    err = field_errors(sim, ref)
    assert max(err.values()) <= 1e-12, err
    assert np.linalg.norm(sim.fields["jx"].numpy()) > 0  # the antenna drives jx
    L = float(max(NX, NY, NZ))
    compare_particles(sim, ref, 1e-12 * L, 1e-12)
    ions = sim.species[1].arrays()  # frozen: untouched
    assert (ions["ux"] == 0).all() and (ions["z"] >= Z0).all()
    sim.step(2)
    ref.step()
    ref.step()
    err = field_errors(sim, ref)
    assert max(err.values()) <= 1e-9, err


def test_laser_plasma_mixed_precision(torch, orc):
    sim, ref = build(torch, orc, "mixed")
    sim.step(1)
    ref.step()
    err = field_errors(sim, ref)
    assert max(err.values()) <= 1e-5, err
    sim.step(2)
    ref.step()
    ref.step()
    err = field_errors(sim, ref)
    assert max(err.values()) <= 1e-4, err
