"""GPU parity of the tiled step (fused gather+push+deposit, current guards,
FDTD, periodic re-binning and per-tile sort) against the CPU oracle step.

One step is compared at 1e-12 relative (norm) on E, B, J and on every
particle (matched through the id carried by its weight).  Several steps are
compared at a looser bar because FP64 atomics make the GPU sum order differ
from the oracle's and round-off then grows with the plasma dynamics.
"""
import numpy as np
import pytest

from tests.pic_oracle import OracleState, default_dt, thermal_plasma

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return None


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(
        np.linalg.norm(np.asarray(b).ravel()), 1e-300)


def build(torch, orc, order, lo, pusher, depo, norder=2, ntiles=(2, 2, 2), n=(16, 14, 12),
          ppc=4, uth=(0.2, 0.02), chunk=0, field_amp=0.05, seed=1, d=(1.0, 0.9, 1.1),
          precision="fp64", id_scale=1e-9, pbound=None, partcom=0):
    from picsar_amd.sim import Simulation, _fd_weights
    nx, ny, nz = n
    dx, dy, dz = d
    xmin, ymin, zmin = -3.0, 0.5, 2.0
    dt = default_dt(dx, dy, dz)
    sim = Simulation(nx, ny, nz, dx, dy, dz, xmin, ymin, zmin, *ntiles, order=order,
                     l_lower_order_in_v=lo, pusher=pusher, depo=depo, norder=norder, dt=dt,
                     chunk=chunk, precision=precision, pbound=pbound, partcom=partcom)
    ref = OracleState(orc, nx, ny, nz, sim.ng, dx, dy, dz, xmin, ymin, zmin, dt, order, lo,
                      pusher, depo, norder, _fd_weights(norder) if norder > 2 else None)
    ref.pbound = tuple(sim.grid.pbound)
    ref.partcom = partcom
    ref.nt = tuple(ntiles)
    rng = np.random.default_rng(seed)
    shape = ref.f["ex"].shape
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        a = np.asfortranarray(rng.standard_normal(shape) * field_amp)
        orc.field_bc_periodic(a, nx, ny, nz, sim.ng, sim.ng, sim.ng)
        ref.f[k][...] = a
        sim.fields[k].upload(a)
    off = 0
    for (q, m, u) in ((-1.0, 1.0, uth[0]), (1.0, 1836.0, uth[1])):
        p = thermal_plasma(rng, nx, ny, nz, ppc, dx, dy, dz, xmin, ymin, zmin, u, id_offset=off,
                           id_scale=id_scale)
        off += len(p[0])
        ref.add_species(q, m, *p)
        sim.add_species("e" if q < 0 else "i", q, m, *p, gaminv=ref.species[-1]["gaminv"])
    return sim, ref, (lambda g: g.numpy())


def compare_particles(sim, ref, tol_x, tol_u):
    for sp, rs in zip(sim.species, ref.species):
        a = sp.arrays()
        if len(rs["w"]) == 0:
            assert len(a["w"]) == 0
            continue
        ia = np.argsort(a["w"])
        ib = np.argsort(rs["w"])
        np.testing.assert_array_equal(a["w"][ia], rs["w"][ib])
        for k in ("x", "y", "z"):
            assert np.abs(a[k][ia] - rs[k][ib]).max() <= tol_x, k
        for k in ("ux", "uy", "uz", "gaminv"):
            assert rel(a[k][ia], rs[k][ib]) <= tol_u, (k, rel(a[k][ia], rs[k][ib]))


def compare_fields(sim, ref, to_numpy_grid, tol):
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        got = to_numpy_grid(sim.fields[k])
        r = rel(got, ref.f[k])
        assert r <= tol, (k, r)


def compare_physical_fields(sim, ref, to_numpy_grid, tol):
    """E and B on every node; J on the nodes 0..n-1 of each axis."""
    g = sim.ng
    cut = (slice(g, g + sim.nx), slice(g, g + sim.ny), slice(g, g + sim.nz))
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        got = to_numpy_grid(sim.fields[k])
        want = ref.f[k]
        if k.startswith("j"):
            got, want = got[cut], want[cut]
        r = rel(got, want)
        assert r <= tol, (k, r)


def check_tiles(sim):
    """Every particle sits in the tile that holds it (membership of
    set_tile_split_for_species) and the tiles are packed and cell-sorted."""
    for sp in sim.species:
        b = sp.active
        off = b.tile_off.numpy()
        npt = b.tile_np.numpy()
        assert off[0] == 0 and off[-1] == sp.npart
        np.testing.assert_array_equal(off[1:] - off[:-1], npt)
        x = b.arrays["x"].numpy(sp.npart)
        L = sim.nx * sim.dx
        assert (x >= sim.xmin).all() and (x < sim.xmin + L + 1e-12).all()
        bs = b.bin_start.numpy()
        assert bs[0] == 0 and bs[-1] == sp.npart and (np.diff(bs) >= 0).all()


CASES = [
    # order, lower order in v, pusher, deposition, FDTD order
    (1, True, "boris", "direct", 2),
    (1, False, "vay", "esirkepov", 2),
    (2, True, "vay", "esirkepov", 2),
    (2, False, "boris", "direct", 2),
    (3, True, "vay", "esirkepov", 2),
    (3, False, "boris", "esirkepov", 4),
    (3, True, "vay", "direct", 4),
]


@pytest.mark.parametrize("order,lo,pusher,depo,norder", CASES)
def test_one_step(torch, orc, order, lo, pusher, depo, norder):
    sim, ref, tng = build(torch, orc, order, lo, pusher, depo, norder)
    check_tiles(sim)
    sim.step(1)
    ref.step()
    check_tiles(sim)
    compare_fields(sim, ref, tng, 1e-12)
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)


@pytest.mark.parametrize("order,depo", [(1, "esirkepov"), (2, "direct"), (3, "esirkepov"),
                                         (3, "direct")])
def test_relativistic_particles(torch, orc, order, depo):
    """Electrons at u ~ 3 (v up to 0.99 c) cross a cell boundary on most
    steps, as C4's gamma = 10 bunch does, so the wide Esirkepov box and the
    re-binning across tiles are the common case: one step at 1e-12."""
    sim, ref, tng = build(torch, orc, order, True, "vay", depo, uth=(3.0, 0.3), seed=11)
    sim.step(1)
    ref.step()
    check_tiles(sim)
    compare_fields(sim, ref, tng, 1e-12)
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)


def test_rebin_at_tile_boundaries(torch):
    """Particles at rest exactly on, and a few ulps either side of, every
    tile boundary (the real-number boundary xmin + (k*n0 - 1/2)*dx of
    particle_boundaries.F90:397) land in the tile the reference's formula
    floor(((x - xmin) + dx/2) / (n0*dx)) gives, evaluated here in IEEE
    double by numpy, and keep their positions bit for bit."""
    from picsar_amd.sim import Simulation
    n, nt = (21, 18, 16), (3, 2, 4)
    d = (0.7, 1.3, 0.9)
    x0 = (-1.1, 0.3, 5.0)
    n0 = [(n[a] + 1) // nt[a] for a in range(3)]
    rng = np.random.default_rng(5)
    cols = []
    for a in range(3):
        vals = []
        for k in range(1, nt[a]):
            v = x0[a] + (k * n0[a] - 0.5) * d[a]
            for _ in range(3):
                v = np.nextafter(v, -np.inf)
            for _ in range(7):
                vals.append(v)
                v = np.nextafter(v, np.inf)
        cols.append(np.repeat(np.array(vals), 8))
    m = max(len(c) for c in cols)
    npart = 3 * m
    pos = [x0[a] + rng.random(npart) * n[a] * d[a] for a in range(3)]
    for a in range(3):  # block a: axis a on a boundary, the others random
        pos[a][a * m:a * m + len(cols[a])] = cols[a]
    w = (np.arange(npart) + 1.0) * 1e-9  # ids
    sim = Simulation(*n, *d, *x0, *nt, order=2, pusher="vay", depo="esirkepov")
    z = np.zeros(npart)
    sim.add_species("e", -1.0, 1.0, *pos, z, z, z, w)
    sim.step(1)
    check_tiles(sim)
    sp = sim.species[0]
    a = sp.arrays()
    off = sp.active.tile_off.numpy()
    tile = np.repeat(np.arange(len(off) - 1), np.diff(off))
    o = np.argsort(a["w"])
    it = []
    for ax, k in enumerate(("x", "y", "z")):
        np.testing.assert_array_equal(a[k][o], pos[ax])  # at rest: not moved
        t = np.floor(((pos[ax] - x0[ax]) + d[ax] / 2.0) / (float(n0[ax]) * d[ax]))
        it.append(np.clip(t, 0, nt[ax] - 1).astype(np.int64))
    np.testing.assert_array_equal(tile[o], it[0] + nt[0] * (it[1] + nt[1] * it[2]))


def _antenna_oracle_dict(a):
    """The oracle's copy of a GPU laser antenna (its initial state)."""
    h = a.arrays()
    return dict(x=h["x"], y=h["y"], z=h["z"],
                pid=np.asfortranarray(np.stack([h["w"], h["xx"], h["yy"]], axis=1)),
                window=a.time_window, emax=a.emax, emax1=a.emax1, emax2=a.emax2, pv1=a.pv1,
                pv2=a.pv2, k0=a.k0, q_z=a.q_z, charge=a.charge, tau=a.laser_tau,
                t_peak=a.t_peak, torder=a.temporal_order, polangle=a.polangle)


@pytest.mark.parametrize("depo", ["esirkepov", "direct"])
def test_laser_antennas_in_the_step(torch, orc, depo):
    """A plasma with two laser antennas (a Gaussian one normal to z, a
    Hanning-windowed one on a tilted plane): each step pushes the antenna
    particles by the laser profile at it*dt - dt/2 and deposits their current
    with the plasma's (submain.F90:113-135).  One step at 1e-12, three at
    1e-9 against the oracle."""
    sim, ref, tng = build(torch, orc, 3, True, "vay", depo, uth=(0.05, 0.005))
    x0, y0, z0 = sim.xmin, sim.ymin, sim.zmin
    ants = [sim.add_laser_antenna(spot=(x0 + 8.0, y0 + 6.1, z0 + 5.3), vector=(0, 0, 1),
                                  polvector1=(1, 0, 0), lambda_laser=8.0, laser_w0=4.0,
                                  laser_ctau=6.0, t_peak=1.5, emax1=0.05, emax2=0.02,
                                  polangle=0.25, charge=1.0),
            sim.add_laser_antenna(spot=(x0 + 7.0, y0 + 6.0, z0 + 6.0), vector=(0, 0.6, 0.8),
                                  polvector1=(1, 0, 0), lambda_laser=6.0, laser_w0=3.0,
                                  laser_ctau=5.0, t_peak=2.0, emax1=0.03, time_window=1,
                                  charge=-1.0)]
    for a in ants:
        assert a.npart > 0
        ref.add_antenna(_antenna_oracle_dict(a))
    sim.step(1)
    ref.step()
    compare_fields(sim, ref, tng, 1e-12)
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)
    for a, ra in zip(ants, ref.antennas):
        h = a.arrays()
        for k in ("x", "y", "z", "ux", "uy", "uz"):
            assert np.abs(h[k] - ra[k]).max() <= 1e-14 * max(np.abs(ra[k]).max(), 1.0), k
    assert np.linalg.norm(tng(sim.fields["jx"])) > 0
    sim.step(2)
    ref.step()
    ref.step()
    compare_fields(sim, ref, tng, 1e-9)


@pytest.mark.parametrize("partcom", [0, 1])
@pytest.mark.parametrize("pbound", [
    ("absorbing",) * 6,
    ("reflecting", "reflecting", "absorbing", "absorbing", "periodic", "periodic"),
    ("periodic", "periodic", "periodic", "periodic", "reinjecting", "reinjecting"),
    ("reinjecting", "absorbing", "reflecting", "periodic", "periodic", "reflecting"),
    ("reflecting",) * 6,
])
def test_particle_boundaries(torch, orc, pbound, partcom):
    """Absorbing, reflecting and reinjecting domain boundaries with fast
    particles (u_th = 0.6 c), under the reference's default exchange
    (partcom 0, particle_bcs_tiles_and_mpi_3d, where a reflecting face is
    periodic) and under partcom 1 (particle_bcs_mpi_blocking, which
    reflects).  The step runs in the reference's order (push, boundaries,
    deposition from the particles as the boundaries left them) and matches
    the oracle in that order: the particles kept, their positions and
    momenta and the fields after one step (1e-12) and three (1e-9)."""
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", uth=(0.6, 0.05),
                          pbound=pbound, partcom=partcom)
    eff = [("periodic" if (b == "reflecting" and partcom == 0) else b) for b in pbound]
    assert sim.split_step == any(b != "periodic" for b in eff)
    n0 = sim.total_particles()
    sim.step(1)
    ref.step_reference_order()
    check_tiles(sim)
    if sim.split_step:
        compare_fields(sim, ref, tng, 1e-12)
    else:
        # every face periodic in effect: the fused step deposits before the
        # wrap; E, B and J on the nodes 0..n-1 are the reference order's
        compare_physical_fields(sim, ref, tng, 1e-12)
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)
    if "absorbing" in pbound:
        assert sim.total_particles() < n0
    else:
        assert sim.total_particles() == n0
    for _ in range(2):
        sim.step(1)
        ref.step_reference_order()
    check_tiles(sim)
    (compare_fields if sim.split_step else compare_physical_fields)(sim, ref, tng, 1e-9)
    assert sim.total_particles() == sum(len(s["x"]) for s in ref.species)


def test_fused_periodic_step_is_the_reference_order(torch, orc):
    """With periodic faces the fused step deposits each particle before its
    wrap; the reference deposits after it (submain.F90:99-140).  The two
    give the same E and B everywhere and the same J on the nodes 0..n-1 of
    each axis; they differ only in J's node n and guards, which hold what
    the periodic fold leaves there and which the E update does not use
    (efield_bcs then copies node 0 to node n).  Checked with fast
    particles (u_th = 0.6 c) against the oracle in the reference's order."""
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", uth=(0.6, 0.05))
    assert not sim.split_step
    sim.step(1)
    ref.step_reference_order()
    compare_physical_fields(sim, ref, tng, 1e-12)
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)


@pytest.mark.parametrize("order,pusher,depo", [(3, "vay", "esirkepov"), (2, "boris", "direct"),
                                               (1, "boris", "esirkepov")])
def test_managers_in_reference_order(torch, orc, order, pusher, depo):
    """The zero-argument managers (picsar_amd.managers) called in
    submain.F90's order -- gather + push, particle boundaries and sort, then
    the deposition from the wrapped positions -- with an antenna and
    reflecting x faces: one step at 1e-12, three at 1e-9 against the oracle
    in the same order."""
    from picsar_amd.managers import Managers
    pb = ("reflecting", "reflecting", "periodic", "periodic", "periodic", "periodic")
    sim, ref, tng = build(torch, orc, order, True, pusher, depo, pbound=pb, partcom=1)
    a = sim.add_laser_antenna(spot=(sim.xmin + 8.0, sim.ymin + 6.1, sim.zmin + 5.3),
                              vector=(0, 0, 1), polvector1=(1, 0, 0), lambda_laser=8.0,
                              laser_w0=4.0, laser_ctau=6.0, t_peak=1.5, emax1=0.05)
    ref.add_antenna(_antenna_oracle_dict(a))
    mgr = Managers(sim)
    mgr.step(1)
    ref.step_reference_order()
    check_tiles(sim)
    compare_fields(sim, ref, tng, 1e-12)
    L = max(sim.nx * sim.dx, sim.ny * sim.dy, sim.nz * sim.dz)
    compare_particles(sim, ref, 1e-12 * L, 1e-12)
    mgr.step(2)
    ref.step_reference_order()
    ref.step_reference_order()
    compare_fields(sim, ref, tng, 1e-9)


def test_managers_run_mixed_precision(torch, orc):
    """The managers drive a mixed-precision species through the tiled
    kernels, in the reference's order, with reflecting faces (partcom 1):
    fields and momenta after one step within FP32 round-off of the FP64
    oracle in that order."""
    from picsar_amd.managers import Managers
    pb = ("reflecting", "reflecting", "periodic", "periodic", "absorbing", "absorbing")
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", precision="mixed",
                          id_scale=2.0 ** -15, ppc=2, pbound=pb, partcom=1, uth=(0.4, 0.05))
    for s in ref.species:  # the oracle starts from the values the GPU holds
        for k in ("ux", "uy", "uz", "w"):
            s[k] = s[k].astype(np.float32).astype(np.float64)
        s["gaminv"] = 1.0 / np.sqrt(1.0 + s["ux"] ** 2 + s["uy"] ** 2 + s["uz"] ** 2)
    Managers(sim).step(1)
    ref.step_reference_order()
    check_tiles(sim)
    for k in ("jx", "jy", "jz"):
        assert rel(tng(sim.fields[k]), ref.f[k]) <= 1e-5, k
    assert sim.total_particles() == sum(len(s["x"]) for s in ref.species)
    compare_particles(sim, ref, 1e-5, 1e-5)


def test_absorbing_boundaries_refuse_graph_capture(torch, orc):
    sim, ref, tng = build(torch, orc, 1, True, "boris", "esirkepov", pbound=("absorbing",) * 6)
    sim.step(1)
    with pytest.raises(RuntimeError):
        sim.capture_steps(2)


def test_laser_antenna_refuses_graph_capture(torch, orc):
    sim, ref, tng = build(torch, orc, 1, True, "boris", "esirkepov")
    sim.add_laser_antenna(spot=(sim.xmin + 4.0, sim.ymin + 4.0, sim.zmin + 4.0),
                          vector=(0, 0, 1), polvector1=(1, 0, 0), lambda_laser=8.0,
                          laser_w0=4.0, laser_ctau=6.0, t_peak=2.0, emax1=0.05)
    sim.step(1)
    with pytest.raises(RuntimeError):
        sim.capture_steps(2)


@pytest.mark.parametrize("chunk", [37, 100000])
def test_chunking_does_not_change_results(torch, orc, chunk):
    """Tiny chunks (runs cut inside rows, many chunks per tile) and one chunk
    per sorting plane give the oracle's result."""
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", chunk=chunk)
    sim.step(1)
    ref.step()
    compare_fields(sim, ref, tng, 1e-12)
    compare_particles(sim, ref, 1e-12 * sim.nx * sim.dx, 1e-12)


@pytest.mark.parametrize("tune", [
    # no field block fits: gather and deposit from global memory
    {"fused.wg_per_cu": 2, "fused.lds_bytes": 4096},
    {"fused.wg_per_cu": 2, "fused.lds_bytes": 79616},
    {"fused.wg_per_cu": 3, "fused.lds_bytes": 52224},
    {"fused.j_lds_bytes": 0},         # fields in LDS, every current block deposited globally
    {"fused.j_lds_bytes": 6000},      # some current blocks in LDS, the others globally
    {"fused.spread": 0},
    {"fused.jpad": 0},
    {"fused.sstride": 5},             # lanes 5 particles apart in super-blocks of 320
])
@pytest.mark.parametrize("depo", ["esirkepov", "direct"])
def test_overflow_and_variants_match(torch, orc, tune, depo):
    """Work items whose blocks exceed the LDS budget are finished by the
    overflow kernel; the result is the oracle's either way."""
    from picsar_amd import _lib
    import ctypes as C
    saved = {}
    for k, v in tune.items():
        old = C.c_int64(0)
        _lib.call("pxrhip_get_tuning", k.encode(), C.byref(old))
        saved[k] = old.value
        _lib.call("pxrhip_set_tuning", k.encode(), v)
    try:
        sim, ref, tng = build(torch, orc, 3, True, "vay", depo)
        sim.step(1)
        ref.step()
    finally:
        for k, v in saved.items():
            _lib.call("pxrhip_set_tuning", k.encode(), v)
    compare_fields(sim, ref, tng, 1e-12)
    compare_particles(sim, ref, 1e-12 * sim.nx * sim.dx, 1e-12)


def _last_launch():
    from picsar_amd import _lib
    import ctypes as C
    out = []
    for k in ("last.wg_per_cu", "last.lds_bytes"):
        v = C.c_int64(0)
        _lib.call("pxrhip_get_tuning", k.encode(), C.byref(v))
        out.append(v.value)
    return tuple(out)


@pytest.mark.parametrize("depo,n,ntiles,want", [
    ("esirkepov", (16, 14, 12), (2, 2, 2), (3, 52224)),   # narrow tiles: three workgroups
    ("esirkepov", (40, 12, 12), (1, 2, 2), (2, 79616)),   # a 40-cell-wide tile: two
    ("direct", (16, 14, 12), (2, 2, 2), (2, 79616)),      # order-3 direct FP64: two
])
def test_automatic_launch_choice(torch, orc, depo, n, ntiles, want):
    """fused.wg_per_cu = 0 picks three workgroups per CU when a four-row
    work item of the widest tile fits 52224 bytes, else two; the step is the
    oracle's either way."""
    sim, ref, tng = build(torch, orc, 3, True, "vay", depo, n=n, ntiles=ntiles)
    sim.step(1)
    ref.step()
    assert _last_launch() == want
    compare_fields(sim, ref, tng, 1e-12)
    compare_particles(sim, ref, 1e-12 * sim.nx * sim.dx, 1e-12)


@pytest.mark.parametrize("depo,margin,precision", [("esirkepov", 0, "fp64"),
                                                    ("esirkepov", 60, "fp64"),
                                                    ("direct", 0, "fp64"),
                                                    ("esirkepov", 0, "mixed")])
def test_stale_sorting_uses_global_paths(torch, orc, depo, margin, precision):
    """Two gather+push+deposit passes without re-binning in between: in the
    second, particles have left the sorting cells their work item was built
    from, so stencils leave the LDS blocks and those particles gather and
    deposit through global memory.  Fast particles (u_th = 0.6 c) make this
    frequent.  The result is still the oracle's (mixed precision: J at 1e-5
    of the oracle's, which is given the FP32-rounded momenta and weights)."""
    from picsar_amd import _lib
    import ctypes as C
    old = C.c_int64(0)
    _lib.call("pxrhip_get_tuning", b"fused.margin_pct", C.byref(old))
    _lib.call("pxrhip_set_tuning", b"fused.margin_pct", margin)
    try:
        kw = dict(precision="mixed", id_scale=2.0 ** -15, ppc=2) if precision == "mixed" else {}
        sim, ref, tng = build(torch, orc, 3, True, "vay", depo, uth=(0.6, 0.05), **kw)
        if precision == "mixed":
            for rs in ref.species:
                for k in ("ux", "uy", "uz", "w"):
                    rs[k][...] = rs[k].astype(np.float32).astype(np.float64)
                rs["gaminv"][...] = 1.0 / np.sqrt(1 + rs["ux"] ** 2 + rs["uy"] ** 2 +
                                                  rs["uz"] ** 2)
        sim.gather_push_deposit()
        for sp in sim.species:
            sim.gather_push_deposit_species(sp)  # J accumulates
        sim.check()
    finally:
        _lib.call("pxrhip_set_tuning", b"fused.margin_pct", old.value)
    ref.gather_push_deposit(zero_j=True)
    ref.gather_push_deposit(zero_j=False)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for k in ("jx", "jy", "jz"):
        assert rel(tng(sim.fields[k]), ref.f[k]) <= tol, k
    if precision == "fp64":
        compare_particles(sim, ref, 1e-12 * sim.nx * sim.dx, 1e-12)


def test_single_tile_and_uneven_tiles(torch, orc):
    # 1 tile (the reference's ntx == 1 branch) and a split whose last tile
    # takes the remainder (17 nodes over 3 tiles: 5, 5, 7)
    for nt in ((1, 1, 1), (3, 2, 3)):
        sim, ref, tng = build(torch, orc, 2, True, "boris", "esirkepov", ntiles=nt,
                              n=(16, 16, 16))
        sim.step(1)
        ref.step()
        check_tiles(sim)
        compare_fields(sim, ref, tng, 1e-12)
        compare_particles(sim, ref, 1e-12 * 16, 1e-12)


def test_ten_steps(torch, orc):
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", ppc=2)
    for _ in range(10):
        sim.step(1)
        ref.step()
    check_tiles(sim)
    compare_fields(sim, ref, tng, 1e-9)
    compare_particles(sim, ref, 1e-9, 1e-9)


def test_c1_config_hundred_steps(torch, orc):
    """The CPU-runnable configuration: 32^3 periodic thermal plasma, dx = 1,
    dt = 0.95 CFL, electrons and ions at 8 ppc (u_th 0.01 c and 0.001 c),
    Boris, order-1 direct deposition, Yee, one tile, 100 steps, FP64,
    against the oracle's step.  FP64 atomics reorder sums on the GPU, so the
    comparison after 100 coupled steps is at 1e-9."""
    from picsar_amd.sim import Simulation
    n, ppc = 32, 8
    dt = default_dt(1.0, 1.0, 1.0)
    sim = Simulation(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1, 1, 1, order=1,
                     l_lower_order_in_v=True, pusher="boris", depo="direct", norder=2, dt=dt)
    ref = OracleState(orc, n, n, n, sim.ng, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, dt, 1, True, "boris",
                      "direct")
    rng = np.random.default_rng(2024)
    off = 0
    for (q, m, u) in ((-1.0, 1.0, 0.01), (1.0, 1836.0, 0.001)):
        p = list(thermal_plasma(rng, n, n, n, ppc, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, u,
                                id_offset=off))
        off += len(p[0])
        p[6] = p[6] * (0.02 ** 2 / ppc)  # electron plasma frequency 0.02
        ref.add_species(q, m, *p)
        sim.add_species("e" if q < 0 else "i", q, m, *p, gaminv=ref.species[-1]["gaminv"])
    for _ in range(100):
        sim.step(1, check=False)
        ref.step()
    sim.check()
    check_tiles(sim)
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        assert rel(sim.fields[k].numpy(), ref.f[k]) <= 1e-9, k
    for sp, rs in zip(sim.species, ref.species):
        a = sp.arrays()
        ia, ib = np.argsort(a["w"]), np.argsort(rs["w"])
        np.testing.assert_array_equal(a["w"][ia], rs["w"][ib])
        for k in ("x", "y", "z"):
            assert np.abs(a[k][ia] - rs[k][ib]).max() <= 1e-9 * n, k
        for k in ("ux", "uy", "uz"):
            assert rel(a[k][ia], rs[k][ib]) <= 1e-9, k


def test_graph_replays_match_eager_steps(torch, orc):
    """Two whole steps captured in a HIP graph and replayed twice, with eager
    work and synchronisation between the replays, give the eager steps'
    state after each replay, and the oracle's after five steps."""
    sim_e, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", ppc=2)
    sim_g, _, _ = build(torch, orc, 3, True, "vay", "esirkepov", ppc=2)
    sim_e.step(1)
    sim_g.step(1)
    ref.step()
    graph = sim_g.capture_steps(2)
    for replay in range(2):
        sim_e.step(2)
        sim_g.step_graph(graph)
        for _ in range(2):
            ref.step()
        sim_g.kinetic_energy()  # eager work and a synchronisation between replays
        for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
            a, b = sim_g.fields[k].numpy(), sim_e.fields[k].numpy()
            assert rel(a, b) <= 1e-12, (replay, k, rel(a, b))
        check_tiles(sim_g)
    assert sim_g.it == sim_e.it == 5
    compare_fields(sim_g, ref, tng, 1e-11)
    compare_particles(sim_g, ref, 1e-11, 1e-11)
    del graph


def test_graph_refuses_replay_after_library_memory_moved(torch, orc):
    """A graph records the library's scratch buffers and tile table.  Once
    another simulation with a different grid (and a larger species) has run,
    those are reallocated, and launching the old graph raises instead of
    replaying on freed memory; a graph captured afresh runs."""
    from picsar_amd._lib import PxrHipError
    sim, ref, tng = build(torch, orc, 2, True, "boris", "esirkepov", ppc=2)
    sim.step(1)
    graph = sim.capture_steps(2)
    sim.step_graph(graph)
    other, _, _ = build(torch, orc, 2, True, "boris", "esirkepov", ppc=6, n=(18, 16, 14),
                        ntiles=(2, 2, 3))
    other.step(1)
    it = sim.it
    with pytest.raises(PxrHipError):
        graph.launch()
    assert sim.it == it
    del graph
    sim.step(1)  # scratch and tile table for this grid again
    g2 = sim.capture_steps(2)
    sim.step_graph(g2)
    check_tiles(sim)
    del g2


def test_autotune_keeps_the_physics(torch, orc):
    """The autotuner's trial steps are real steps: after trying two launch
    configurations (3 steps each) the simulation is the oracle's after 6
    steps, and the faster configuration is left applied."""
    from picsar_amd.autotune import autotune, get_tuning, set_tuning
    before = get_tuning()
    sim, ref, tng = build(torch, orc, 3, True, "vay", "esirkepov", ppc=2)
    cands = [{"fused.chunk": 512, "fused.spread": 1}, {"fused.chunk": 4096, "fused.spread": 0}]
    try:
        res = autotune(sim, cands, steps=2)
        assert res["best"] in cands and len(res["trials"]) == 2
        assert get_tuning(("fused.chunk",))["fused.chunk"] == res["best"]["fused.chunk"]
    finally:
        set_tuning(before)
    for _ in range(6):
        ref.step()
    assert sim.it == 6
    check_tiles(sim)
    compare_fields(sim, ref, tng, 1e-9)
    compare_particles(sim, ref, 1e-9, 1e-9)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_relativistic_bunch_on_uneven_tiles(torch, orc, precision):
    """A dense gamma = 10 electron bunch (moving half a cell per step along
    z, so that most particles cross a cell face) in a sparse background:
    tiles hold from a handful to thousands of particles.  Three steps
    against the oracle (FP64 at 1e-12 per step compounded to 1e-11; mixed
    precision at 1e-5)."""
    from picsar_amd.sim import Simulation
    n = (24, 24, 32)
    dt = default_dt(1.0, 1.0, 1.0)
    sim = Simulation(*n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 3, 3, 4, order=3, l_lower_order_in_v=True,
                     pusher="vay", depo="esirkepov", norder=2, dt=dt, precision=precision)
    ref = OracleState(orc, *n, sim.ng, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, dt, 3, True, "vay",
                      "esirkepov")
    rng = np.random.default_rng(77)
    nb, nbg = 20000, 24 * 24 * 32
    scale = 2.0 ** -15 if precision == "mixed" else 1e-9
    pos = [np.clip(rng.normal(c, 2.5, nb), 0.0, np.nextafter(m, 0)) for c, m in
           ((12.0, 24.0), (12.0, 24.0), (16.0, 32.0))]
    mom = [rng.standard_normal(nb) * 0.05, rng.standard_normal(nb) * 0.05,
           np.sqrt(99.0) + rng.standard_normal(nb) * 0.1]
    bg = [rng.random(nbg) * m for m in n]
    bmom = [rng.standard_normal(nbg) * 0.01 for _ in range(3)]
    for name, p, u, off in (("bunch", pos, mom, 0), ("background", bg, bmom, nb)):
        w = 1e-3 * (1.0 + (np.arange(len(p[0])) + off) * scale)
        if precision == "mixed":
            u = [a.astype(np.float32).astype(np.float64) for a in u]
            w = w.astype(np.float32).astype(np.float64)
        ref.add_species(-1.0, 1.0, *p, *u, w)
        sim.add_species(name, -1.0, 1.0, *p, *u, w, gaminv=ref.species[-1]["gaminv"])
    counts = sim.species[0].active.tile_np.numpy()
    assert counts.max() > 50 * max(counts.min(), 1)
    for _ in range(3):
        sim.step(1)
        ref.step()
    check_tiles(sim)
    tol = 1e-11 if precision == "fp64" else 1e-5
    for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        assert rel(sim.fields[k].numpy(), ref.f[k]) <= tol, (k, rel(sim.fields[k].numpy(), ref.f[k]))
    for sp, rs in zip(sim.species, ref.species):
        a = sp.arrays()
        ia, ib = np.argsort(a["w"]), np.argsort(rs["w"])
        for k in ("ux", "uy", "uz"):
            assert rel(a[k][ia].astype(np.float64), rs[k][ib]) <= tol, k


@pytest.mark.parametrize("norder", [2, 4, 6])
def test_default_time_step_is_stable(torch, norder):
    """The default dt is cfl (0.95) times the stability limit of the stencil:
    random fields in vacuum keep their energy over 300 steps.  At the Yee
    limit an order-4 stencil is unstable (checked by dt = 0.95 of the Yee
    limit: the grid-scale modes grow by orders of magnitude)."""
    import math
    from picsar_amd.sim import Simulation, _fd_weights

    def run(dt):
        sim = Simulation(24, 20, 16, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 1, 1, 1, order=1,
                         norder=norder, dt=dt)
        rng = np.random.default_rng(4)
        for k in ("ex", "ey", "ez", "bx", "by", "bz"):
            a = np.asfortranarray(rng.standard_normal(sim.fields[k].shape))
            sim.fields[k].upload(a)
        sim.efield_bcs()
        sim.bfield_bcs()
        e0 = sim.field_energy()
        for _ in range(300):
            sim.field_step()
        return sim, sim.field_energy() / e0

    sim, growth = run(None)
    yee = 1.0 / math.sqrt(3.0)
    wsum = 1.0 if norder == 2 else float(np.abs(_fd_weights(norder)).sum())
    assert sim.dt == pytest.approx(0.95 * yee / wsum, rel=1e-14)
    assert growth < 2.0, growth
    if norder == 4:
        _, growth = run(0.95 * yee)
        assert growth > 1e3, growth
