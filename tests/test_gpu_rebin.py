"""The tiled re-binning (pxrhip_tiles_rebin_sort) is a stable counting sort.

After particle_bcs, particles are assigned to tiles by the tile index of
particle_boundaries.F90:397 and ordered inside each tile by the sorting key
of pxr_particle_bin_sorting (sorting.F90:294-304), which sorts stably
(:333-343).  Here the GPU's output order is compared, bit for bit, with
numpy's stable argsort of the input order by (tile, cell), the bins being
computed on the host from the reference's formulas, and two runs from the
same input are compared with each other.
"""
import math

import numpy as np
import pytest

from tests.test_gpu_step import build

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)
    return None


def host_bins(sim, x, y, z):
    """Global bin of every (in-box) position: tile index by
    particle_boundaries.F90:397, cell of the tile's sorting grid by
    sorting.F90:278-304 (origin x_tile_min + shift, extent up to
    x_tile_max + sdx), tiles numbered x fastest."""
    g = sim.grid
    x0 = (sim.xmin, sim.ymin, sim.zmin)
    d = (sim.dx, sim.dy, sim.dz)
    n = (sim.nx, sim.ny, sim.nz)
    nt = (g.ntx, g.nty, g.ntz)
    sd = (g.sorting_dx, g.sorting_dy, g.sorting_dz)
    P = (x, y, z)
    tmin, tmax, tidx, n0s = [], [], [], []
    for a in range(3):
        n0 = (n[a] + 1) // nt[a]
        lo, hi = [], []
        for i in range(nt[a]):
            gmin = x0[a] + (i * n0) * d[a]
            gmax = gmin + (n0 - 1) * d[a]
            lo.append(x0[a] if i == 0 else gmin - d[a] / 2.0)
            hi.append(x0[a] + n[a] * d[a] if i == nt[a] - 1 else gmax + d[a] / 2.0)
        tmin.append(lo)
        tmax.append(hi)
        n0s.append(n0 if nt[a] > 1 else n[a] + 1)
        t = np.floor(((P[a] - x0[a]) + d[a] / 2.0) / (n0s[a] * d[a])).astype(np.int64)
        tidx.append(np.clip(t, 0, nt[a] - 1))
    # sorting grid of every tile and the first bin of each
    ntiles = nt[0] * nt[1] * nt[2]
    n3 = np.zeros((ntiles, 3), dtype=np.int64)
    base = np.zeros(ntiles + 1, dtype=np.int64)
    org = np.zeros((ntiles, 3))
    for iz in range(nt[2]):
        for iy in range(nt[1]):
            for ix in range(nt[0]):
                t = ix + nt[0] * (iy + nt[1] * iz)
                for a, ii in enumerate((ix, iy, iz)):
                    lo = tmin[a][ii]
                    hi = tmax[a][ii] + sd[a]
                    n3[t, a] = math.ceil((hi - lo) * (1.0 / sd[a]))
                    org[t, a] = lo
                base[t + 1] = base[t] + n3[t].prod()
    t = tidx[0] + nt[0] * (tidx[1] + nt[1] * tidx[2])
    c = [np.floor((P[a] - org[t, a]) * (1.0 / sd[a])).astype(np.int64) for a in range(3)]
    for a in range(3):
        assert (c[a] >= 0).all() and (c[a] < n3[t, a]).all()
    cell = c[2] * n3[t, 0] * n3[t, 1] + c[1] * n3[t, 0] + c[0]
    return base[t] + cell, base[-1]


def wrap(sim, x, y, z):
    out = []
    for v, x0, L in ((x, sim.xmin, sim.nx * sim.dx), (y, sim.ymin, sim.ny * sim.dy),
                     (z, sim.zmin, sim.nz * sim.dz)):
        v = v.copy()
        v[v < x0] += L
        v[v >= x0 + L] -= L
        out.append(v)
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rebin_is_the_stable_sort_by_tile_and_cell(torch, orc, precision):
    idsc = 2.0 ** -15 if precision == "mixed" else 1e-9
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", uth=(0.6, 0.1),
                          precision=precision, id_scale=idsc, ppc=3)
    for sp in sim.species:
        sim.gather_push_species(sp)  # the particles move, some out of the box
    for sp in sim.species:
        src = sp.arrays()  # input order of the re-binning
        xw, yw, zw = wrap(sim, src["x"], src["y"], src["z"])
        bins, nbins = host_bins(sim, xw, yw, zw)
        order = np.argsort(bins, kind="stable")
        sim._rebin(sp)
        dst = sp.arrays()
        np.testing.assert_array_equal(dst["x"], xw[order])
        np.testing.assert_array_equal(dst["y"], yw[order])
        np.testing.assert_array_equal(dst["z"], zw[order])
        for k in ("ux", "uy", "uz", "gaminv", "w"):
            np.testing.assert_array_equal(dst[k], src[k][order])
        bs = sp.active.bin_start.numpy()
        np.testing.assert_array_equal(bs, np.concatenate(
            [[0], np.cumsum(np.bincount(bins, minlength=nbins))]))
    sim.check()


def test_rebin_is_reproducible(torch, orc):
    """Two simulations built from the same seed and stepped the same way
    hold bitwise identical particle arrays after every step (the current
    differs in its last bits between runs: FP64 atomics)."""
    sims = [build(torch, orc, 2, True, "boris", "esirkepov", uth=(0.4, 0.05), ppc=4)[0]
            for _ in range(2)]
    for s in sims:
        for sp in s.species:
            s.gather_push_species(sp)
            s._rebin(sp)
    for spa, spb in zip(sims[0].species, sims[1].species):
        a, b = spa.arrays(), spb.arrays()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        np.testing.assert_array_equal(spa.active.bin_start.numpy(), spb.active.bin_start.numpy())


def test_arrival_order_rebin_keeps_the_cells(torch, orc):
    """With rebin.stable = 0 the particles of a cell may come in any order,
    but every particle is in the same cell as with the stable sort."""
    from picsar_amd import _lib
    sims = [build(torch, orc, 3, True, "vay", "esirkepov", uth=(0.6, 0.1), ppc=3)[0]
            for _ in range(2)]
    try:
        for i, s in enumerate(sims):
            _lib.call("pxrhip_set_tuning", b"rebin.stable", 1 - i)
            for sp in s.species:
                s.gather_push_species(sp)
                s._rebin(sp)
    finally:
        _lib.call("pxrhip_set_tuning", b"rebin.stable", 1)
    for spa, spb in zip(sims[0].species, sims[1].species):
        np.testing.assert_array_equal(spa.active.bin_start.numpy(), spb.active.bin_start.numpy())
        a, b = spa.arrays(), spb.arrays()
        bs = spa.active.bin_start.numpy()
        cell = np.repeat(np.arange(len(bs) - 1), np.diff(bs))
        ka = np.lexsort((a["w"], cell))
        kb = np.lexsort((b["w"], cell))
        for k in a:
            np.testing.assert_array_equal(a[k][ka], b[k][kb], err_msg=k)


def test_tile_index_on_and_next_to_tile_faces(torch, orc):
    """Positions on the tile faces (x0 + (i n0 - 1/2) dx) and 1 to 4 ulps to
    either side, where the tile index of particle_boundaries.F90:397 turns
    on the rounding of its division, with random positions on the other
    axes: the re-binning puts each particle in the tile and cell the
    reference's formula gives (the division done on the host by numpy)."""
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", ppc=1,
                          d=(1.0, 0.9, 1.1))
    rng = np.random.default_rng(3)
    g = sim.grid
    x0 = (sim.xmin, sim.ymin, sim.zmin)
    d = (sim.dx, sim.dy, sim.dz)
    n = (sim.nx, sim.ny, sim.nz)
    nt = (g.ntx, g.nty, g.ntz)
    cols = []
    for a in range(3):
        n0 = (n[a] + 1) // nt[a]
        faces = [x0[a] + (i * n0 - 0.5) * d[a] for i in range(1, nt[a])]
        vals = []
        for f in faces:
            v = f
            vals.append(v)
            for direction in (np.inf, -np.inf):
                v = f
                for _ in range(4):
                    v = np.nextafter(v, direction)
                    vals.append(v)
        vals = np.array(vals)
        k = len(vals)
        P = [x0[b] + rng.uniform(0, n[b], k) * d[b] for b in range(3)]
        P[a] = vals
        cols.append(P)
    x, y, z = (np.concatenate([c[b] for c in cols]) for b in range(3))
    m = len(x)
    u = np.zeros(m)
    w = 1.0 + np.arange(m) * 1e-9
    sp = sim.add_species("faces", -1.0, 1.0, x, y, z, u, u, u, w)
    bins, nbins = host_bins(sim, x, y, z)
    order = np.argsort(bins, kind="stable")
    dst = sp.arrays()
    np.testing.assert_array_equal(dst["x"], x[order])
    np.testing.assert_array_equal(dst["y"], y[order])
    np.testing.assert_array_equal(dst["z"], z[order])
    bs = sp.active.bin_start.numpy()
    np.testing.assert_array_equal(bs, np.concatenate(
        [[0], np.cumsum(np.bincount(bins, minlength=nbins))]))
    sim.check()
