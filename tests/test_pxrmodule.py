"""Host-side logic of the picsar_python drop-in (picsar_amd.pxrmodule): the
defaults of default_init (control_file.F90:77-239), the grid arithmetic of
mpi_initialise (mpi_routines.F90:209, 1505-1540), the tile-split checks and
the errors for what one rank on the GPU does not build.  No GPU is used."""
import numpy as np
import pytest


def fresh():
    from picsar_amd.pxrmodule import PicsarModule
    p = PicsarModule()
    p.default_init()
    return p


def test_import_name_is_the_references():
    from picsar_python import picsarpy
    from picsar_amd.pxrmodule import PicsarModule
    assert isinstance(picsarpy.picsar, PicsarModule)


def test_default_init_values():
    p = fresh()
    assert (p.nox, p.noy, p.noz) == (1, 1, 1)
    assert (p.norderx, p.nordery, p.norderz) == (2, 2, 2)
    assert p.nxguards == max(p.nox, 2) and p.nzjguards == 2
    assert p.partcom == 0 and p.particle_pusher == 0 and p.currdepo == 0
    assert p.dtcoef == 0.7 and p.l_lower_order_in_v
    assert p.nc == p.nlab * p.g0 == 1e23
    assert (p.sorting_dx, p.sorting_shiftx) == (1.0, 0.0)
    assert p.clight == 1.0 and p.eps0 == 1.0  # normalised units of the -DLIBRARY build


def test_mpi_initialise_grid():
    p = fresh()
    p.nx_global_grid, p.ny_global_grid, p.nz_global_grid = 101, 51, 33
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -5e-6, 5e-6, 0.0, 1.0, -2.0, 6.0
    p.mpi_minimal_init(-1)
    p.mpi_initialise()
    assert (p.nx, p.ny, p.nz) == (100, 50, 32)
    assert (p.nx_global, p.nx_grid) == (100, 101)
    assert p.dx == (5e-6 - -5e-6) / 100.0 and p.dz == 8.0 / 32.0
    assert p.x_grid_min_local == p.xmin and p.z_max_local == p.zmax
    assert p.rank == 0


def test_one_rank_only():
    p = fresh()
    p.nprocx = 2
    with pytest.raises(NotImplementedError):
        p.mpi_minimal_init(-1)


def test_tile_split_needs_four_nodes_per_tile():
    p = fresh()
    p.nx_global_grid = p.ny_global_grid = p.nz_global_grid = 17
    p.xmax = p.ymax = p.zmax = 16.0
    p.mpi_initialise()
    p.ntilex = 4
    p.set_tile_split()  # 17 nodes, 4 tiles of 4
    p.ntilex = 5
    with pytest.raises(ValueError):
        p.set_tile_split()


def test_species_properties_and_staging():
    p = fresh()
    p.set_particle_species_properties(1, "electron", 1.0, -1.0, 4, -1, -1, -1, 1, 1, 1,
                                      0, 0, 0, 0.1, 0.1, 0.1, 5, 0)
    p.set_particle_species_properties(2, "ion", 1836.0, 1.0, 4, -1, -1, -1, 1, 1, 1,
                                      0, 0, 0, 0.1, 0.1, 0.1, 5, 0, False)
    assert p.nspecies == 2
    assert p._species_props[1]["ldodepos"] and not p._species_props[2]["ldodepos"]
    x = np.arange(3.0)
    pid = np.stack([np.full(3, 2.0), np.arange(3.0) + 10], axis=1)  # weight, particle ids
    p.py_add_particles_to_species(1, 3, 2, x, x, x, x, x, x, np.ones(3), pid)
    pid2 = np.stack([np.full(3, 5.0), np.arange(3.0) + 20], axis=1)
    p.py_add_particles_to_species(1, 3, 2, x, x, x, x, x, x, np.ones(3), pid2)
    st = p._staged[1]
    assert len(st) == 2
    np.testing.assert_array_equal(st[0]["w"], 2.0)
    np.testing.assert_array_equal(st[1]["w"], 5.0)
    # the columns beyond the weight are kept, to travel with their particles
    np.testing.assert_array_equal(st[1]["pid"][:, 0], [20.0, 21.0, 22.0])
    # one species keeps one number of pid columns
    with pytest.raises(ValueError):
        p.py_add_particles_to_species(1, 3, x, x, x, x, x, x, np.ones(3), np.full(3, 5.0))
    # the older form (weights only) on a species of its own
    p.py_add_particles_to_species(2, 3, x, x, x, x, x, x, np.ones(3), np.full(3, 7.0))
    assert "pid" not in p._staged[2][0]
    with pytest.raises(TypeError):
        p.py_add_particles_to_species(1, 3, x, x)


def test_stencil_coefficients():
    p = fresh()
    p.norderx = p.nordery = p.norderz = 4
    p.init_stencil_coefficients()
    np.testing.assert_allclose(p.xcoeffs, [9.0 / 8.0, -1.0 / 24.0], rtol=1e-15)


def test_leaf_routines_are_reachable():
    p = fresh()
    assert callable(p.geteb3d_energy_conserving)
    assert callable(p.depose_jxjyjz)
    with pytest.raises(AttributeError):
        p.no_such_routine
