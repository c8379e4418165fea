"""The reference's Python driver, run unchanged against ``picsar_python``.

examples/example_scripts_python/test.py:14-175 configures PICSAR through the
module variables of ``picsar_python.picsarpy.picsar``, sets it up
(default_init, set_particle_species_properties, mpi_minimal_init,
mpi_initialise, set_tile_split, init_tile_arrays,
py_add_particles_to_species, init_stencil_coefficients) and steps with the
managers.  The sequence below is that script's, in Python 3, with its two
stale calls written as the reference's sources now have them:
``push_particles`` is ``field_gathering_plus_particle_pusher``
(particle_pusher_manager_3d.F90:50) and py_add_particles_to_species takes
``npiid`` and a pid array (tiling.F90:1382).  The data are made smaller and
non-trivial: 24^3 cells of size 1, seeded thermal momenta, weights carrying
ids, and random initial fields assigned to the module's field arrays.  The
fields and particles are compared with the oracle's step in the reference's
order after one step (1e-12) and three (1e-9).
"""
import numpy as np
import pytest

from tests.pic_oracle import OracleState
from tests.test_gpu_step import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)
    return None


def run_driver(precision, nsteps_first=1):
    from picsar_python import picsarpy as pxrpy

    pxr = pxrpy.picsar
    pxr.mpi_close()  # a fresh rank for every run of the driver

    #### Input parameters (replace input file)
    pxr.default_init()
    pxr.topology = 0
    pxr.nprocx = 1
    pxr.nprocy = 1
    pxr.nprocz = 1
    ## Simulation box size and duration
    pxr.nx_global_grid = 25
    pxr.ny_global_grid = 25
    pxr.nz_global_grid = 25
    pxr.xmin = -12.0
    pxr.xmax = 12.0
    pxr.ymin = -12.0
    pxr.ymax = 12.0
    pxr.zmin = -12.0
    pxr.zmax = 12.0
    pxr.dtcoef = 0.7
    ## Maxwell solver
    pxr.norderx = 2
    pxr.nordery = 2
    pxr.norderz = 2
    ## Current deposition order
    pxr.nox = 1
    pxr.noy = 1
    pxr.noz = 1
    ## Number of guard cells
    pxr.nxguards = 2
    pxr.nyguards = 2
    pxr.nzguards = 2
    pxr.nxjguards = 2
    pxr.nyjguards = 2
    pxr.nzjguards = 2
    ## Tiling parameters
    pxr.ntilex = 3
    pxr.ntiley = 3
    pxr.ntilez = 3
    ## Sorting
    pxr.sorting_activated = 1
    pxr.sorting_dx = 1.
    pxr.sorting_dy = 1.
    pxr.sorting_dz = 1.
    pxr.sorting_shiftx = 0.
    pxr.sorting_shifty = 0.
    pxr.sorting_shiftz = 0.
    pxr.precision = precision
    pxr.nc = 1e-3  # plasma density (omega_pe^2 in these units)

    #### Init particle distribution
    name = ["electron", "proton"]
    charge = [-1., 1.]
    mass = [1., 1836.]
    nppcell = [2, 2]
    xmin = [-1.3, -1.3]; ymin = [-1.3, -1.3]; zmin = [-1.3, -1.3]
    xmax = [1.3, 1.3]; ymax = [1.3, 1.3]; zmax = [1.3, 1.3]
    vdriftx = [0., 0.]; vdrifty = [0., 0.]; vdriftz = [0., 0.]
    vthx = [0., 0.]; vthy = [0., 0.]; vthz = [0., 0.]
    sorting_period = [5, 5]
    sorting_start = [0, 0]
    pxr.nspecies = len(name)
    for i in range(0, pxr.nspecies):
        pxr.set_particle_species_properties(i + 1, name[i], mass[i], charge[i], nppcell[i],
                                            xmin[i], ymin[i], zmin[i], xmax[i], ymax[i],
                                            zmax[i], vdriftx[i], vdrifty[i], vdriftz[i],
                                            vthx[i], vthy[i], vthz[i], sorting_period[i],
                                            sorting_start[i])
    #### MPI INIT
    pxr.mpi_minimal_init(-1)
    pxr.mpi_initialise()
    mpirank = pxr.rank
    assert mpirank == 0

    rng = np.random.default_rng(8)
    uth = [0.05, 0.002]
    loaded = {}

    def density_profile_global(x, y, z):
        r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
        Rd = (pxr.xmax - pxr.xmin)
        return np.where(r > Rd, 0., 1.)

    def py_load_particles():
        pid0 = 0
        for ispecies in range(0, pxr.nspecies):
            x, y, z = np.mgrid[0:pxr.nx, 0:pxr.ny, 0:pxr.nz]
            x = x.flatten()
            y = y.flatten()
            z = z.flatten()
            x = pxr.x_grid_min_local + x * pxr.dx
            y = pxr.y_grid_min_local + y * pxr.dy
            z = pxr.z_grid_min_local + z * pxr.dz
            dens = density_profile_global(x, y, z)
            x = np.extract(dens > 0., x)
            y = np.extract(dens > 0., y)
            z = np.extract(dens > 0., z)
            dens = np.extract(dens > 0., dens)
            partw = dens * pxr.nc * pxr.dx * pxr.dy * pxr.dz / (nppcell[ispecies])
            for i in range(0, nppcell[ispecies]):
                n = np.size(x)
                ux, uy, uz = (rng.standard_normal(n) * uth[ispecies] for _ in range(3))
                gi = 1.0 / np.sqrt(1.0 + ux ** 2 + uy ** 2 + uz ** 2)
                w = partw * (1.0 + (pid0 + np.arange(n)) * 1e-9)  # ids
                pid0 += n
                px = x + pxr.dx * (i + 0.3) / nppcell[ispecies]
                py = y + pxr.dy * (i + 0.6) / nppcell[ispecies]
                pz = z + pxr.dz * (i + 0.8) / nppcell[ispecies]
                pxr.py_add_particles_to_species(ispecies + 1, n, 1, px, py, pz, ux, uy, uz, gi,
                                                w[:, None])
                loaded.setdefault(ispecies, []).append((px, py, pz, ux, uy, uz, w))

    def initallpy():
        pxr.dt = pxr.dtcoef / (pxr.clight * np.sqrt(1.0 / pxr.dx ** 2 + 1.0 / pxr.dy ** 2 +
                                                    1.0 / pxr.dz ** 2))
        pxr.sorting_dx *= pxr.dx
        pxr.sorting_dy *= pxr.dy
        pxr.sorting_dz *= pxr.dz
        pxr.set_tile_split()
        pxr.init_tile_arrays()
        py_load_particles()
        pxr.init_stencil_coefficients()

    def steppy(nt):
        for i in range(0, nt):
            # Push particles
            pxr.field_gathering_plus_particle_pusher()
            pxr.particle_bcs()
            # Sorting
            pxr.particle_sorting_sub()
            # Deposit currents on the grid
            pxr.pxrdepose_currents_on_grid_jxjyjz()
            pxr.current_bcs()
            # Push EM fields
            pxr.push_bfield()
            pxr.bfield_bcs()
            pxr.push_efield()
            pxr.efield_bcs()
            pxr.push_bfield()
            pxr.bfield_bcs()
            # Compute diags
            pxr.calc_diags()

    initallpy()
    return pxr, steppy, loaded, rng


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_driver_runs_against_picsar_python(torch, orc, precision):
    pxr, steppy, loaded, rng = run_driver(precision)
    # random fields, assigned to the module's arrays as a Forthon driver does
    n = (pxr.nx, pxr.ny, pxr.nz)
    ng = pxr.nxguards
    ref = OracleState(orc, *n, ng, pxr.dx, pxr.dy, pxr.dz, pxr.xmin, pxr.ymin, pxr.zmin,
                      pxr.dt, 1, True, "boris", "esirkepov")
    ref.nt = (pxr.ntilex, pxr.ntiley, pxr.ntilez)
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        a = np.asfortranarray(rng.standard_normal(ref.f[k].shape) * 0.05)
        orc.field_bc_periodic(a, *n, ng, ng, ng)
        ref.f[k][...] = a
        setattr(pxr, k, a)
        np.testing.assert_array_equal(getattr(pxr, k), a)
    for isp, q, m in ((0, -1.0, 1.0), (1, 1.0, 1836.0)):
        cat = [np.concatenate([c[j] for c in loaded[isp]]) for j in range(7)]
        ref.add_species(q, m, *cat)
    tol = 1e-12 if precision == "fp64" else 1e-5
    steppy(1)
    ref.step_reference_order()
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        assert rel(getattr(pxr, k), ref.f[k]) <= tol, (k, rel(getattr(pxr, k), ref.f[k]))
    if precision == "fp64":
        for isp in (0, 1):
            a = pxr.species_arrays(isp + 1)
            rs = ref.species[isp]
            ia, ib = np.argsort(a["w"]), np.argsort(rs["w"])
            np.testing.assert_array_equal(a["w"][ia], rs["w"][ib])
            for k in ("ux", "uy", "uz", "x", "y", "z"):
                assert rel(a[k][ia], rs[k][ib]) <= 1e-12, k
    assert pxr.rho.shape == ref.f["ex"].shape  # calc_diags deposited rho
    steppy(2)
    ref.step_reference_order()
    ref.step_reference_order()
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        assert rel(getattr(pxr, k), ref.f[k]) <= (1e-9 if precision == "fp64" else 1e-4), k
    pxr.mpi_close()
