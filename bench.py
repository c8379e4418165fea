#!/usr/bin/env python3
"""Throughput of the PIC hot path on one MI355X.

Default workload (the precision configuration, the one the metric is quoted
on): a periodic thermal plasma on a 256^3 grid, dx = dy = dz = 1, dt = 0.95
CFL, electrons and ions at 8 particles per cell each (268,435,456
macro-particles), Maxwellian momenta (u_th = 0.01 c and 0.001 c), weights
giving an electron plasma frequency 0.02, order-3 energy-conserving gather,
Vay pusher, order-3 Esirkepov deposition, 2nd-order Yee FDTD, 8^3-node tiles.
It is built twice from the same seed and the same arrays: in mixed precision
(FP64 positions; FP32 momenta, gamma, weights, gather, push and per-particle
current; FP64 LDS and global current and FDTD) -- the headline `value` -- and
in FP64, timed beside it.  Synthetic data generated from a fixed seed and
resident on the GPU before timing.

A timed step is the whole PIC step: fused gather+push+deposit for both
species, current guard fold, FDTD (B half, E, B half with guard exchanges),
periodic re-binning and per-tile sort.  `value` is particle-steps/s over that
whole step; the gather+push+deposit stage alone is reported in `stages`.

The error of the mixed-precision step against the FP64 one is reported
under `precision_error`: J and momenta after the first step (the particles
of both runs are still in the same order then), J and the kinetic energy
after the timed steps, and the total-energy drift of each run over 1000
steps.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --n 128 --ppc 16 --tiles 8 --precision fp64 --no-compare   # 128^3 FP64
    python bench.py --workload bunch --n 256 --tiles 32 --precision mixed       # uneven tiles
    python bench.py --workload laser --n 512 --precision mixed                  # laser plasma
    python bench.py --workload leaf --n 128 --ppc 16 --order 3                  # leaf deposition

`--workload leaf` times the drop-in leaf routines depose_jxjyjz (Esirkepov)
and pxr_depose_rho_n on n^3 * ppc particles of one species handed over as
plain device arrays, in cell order and in random order, in their global-
atomic form and their binned form (particles counted into blocks of nodes
and summed in LDS), and prints one JSON line of those timings.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import copy
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4   # MI355X peak engine clock
NUM_CU = 256


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


BYTES_PER_PARTICLE_FP64 = 120  # read x,y,z,ux,uy,uz,gaminv,w; write x..gaminv
BYTES_PER_PARTICLE_MIXED = 84  # positions FP64 (24 + 24), u, gaminv, w FP32 (20 + 16)
BYTES_PER_CELL = 72            # read E, B (48 B); write J (24 B)
OMEGA_PE = 0.02                # electron plasma frequency (c = dx = 1)

# committed counter summaries of the fused launch, by workload and precision
PROFILES = {
    ("thermal", 256, 8, 32, "mixed"): "fused_c3_mixed",
    ("thermal", 256, 8, 32, "fp64"): "fused_c3_fp64",
    ("thermal", 128, 16, 8, "fp64"): "fused_c2",
    ("thermal", 128, 16, 8, "mixed"): "fused_c2_mixed",
}


def profile_name(args):
    """Stem of the committed counter summaries that hold this run's fused
    launch, for the default launch parameters only."""
    if args.chunk or args.tune or args.autotune or \
            (args.order, args.pusher, args.depo) != (3, "vay", "esirkepov"):
        return None
    return PROFILES.get((args.workload, args.n, args.ppc, args.tiles, args.precision))


def profile_value(name, key):
    """`key = value` line of profiles/<name>.md, or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name + ".md")):
            if line.startswith(key + " ="):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def laser_dims(args):
    """Cells of the laser workload: 512 x 512 x 1024 at --n 512."""
    return args.n, args.n, 2 * args.n


def workload_name(args):
    prec = "FP64" if args.precision == "fp64" else "mixed precision"
    tiles = f"{(args.n + 1) // args.tiles}^3-node tiles"
    if args.workload == "laser":
        nx, ny, nz = laser_dims(args)
        return (f"laser-driven underdense plasma {nx}x{ny}x{nz}, electrons ramped over 100 cells "
                f"to 4 ppc, immobile ions at 1 ppc (frozen, not depositing), "
                f"Gaussian laser a0=2, lambda 32 cells, waist 64 cells, antenna at z=0, "
                f"order-4 FDTD, order-3 gather, Vay, order-3 Esirkepov, "
                f"{args.tile_nodes or 8}^3-node tiles, {prec}")
    meth = (f"order-{args.order} gather, {args.pusher.capitalize()}, order-{args.order} "
            f"{'Esirkepov' if args.depo == 'esirkepov' else 'direct'}")
    if args.workload == "bunch":
        return (f"Gaussian electron bunch (sigma 20 cells, peak 64 ppc, gamma 10) in a 1 ppc "
                f"background, {args.n}^3 periodic, {meth}, Yee, {tiles}, {prec}")
    return (f"thermal plasma {args.n}^3 periodic, {args.ppc} ppc electrons + ions, {meth}, "
            f"Yee, {tiles}, {prec}")


def build_laser_workload(args):
    """A laser-driven underdense plasma: order-4 FDTD; electrons at 4 ppc
    with a weight that ramps linearly over 100 cells (from z = 16) to a
    plateau at 0.1 critical density; immobile ions at 1 ppc with the
    electrons' charge density (frozen and not depositing: the reference's
    lfreeze and ldodepos); a Gaussian laser (a0 = 2, lambda = 32 cells,
    waist 64 cells) launched by an antenna in the plane z = 0.  The
    particles are generated and uploaded slab by slab."""
    from picsar_amd.laser import emax_from_a0
    from picsar_amd.sim import Simulation
    nx, ny, nz = laser_dims(args)
    tn = args.tile_nodes or 8
    sim = Simulation(nx, ny, nz, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, (nx + 1) // tn, (ny + 1) // tn,
                     (nz + 1) // tn, order=3, l_lower_order_in_v=True, pusher="vay",
                     depo="esirkepov", norder=4, cfl=0.95, chunk=args.chunk,
                     precision=args.precision)
    lam, ppc, z0, ramp = 32.0, 4, 16, 100
    k0 = 2 * np.pi / lam
    n_plateau = 0.1 * k0 ** 2       # 0.1 n_c with omega^2 = n_c (m = q = eps0 = 1)
    rng = np.random.default_rng(12345)

    def slabs(ppc_, mom, step=8):
        for k0_ in range(z0, nz, step):
            if (k0_ - z0) % (16 * step) == 0:
                log(f"generating particles: z planes {k0_} of {nz}")
            k1 = min(k0_ + step, nz)
            cells = np.arange(nx * ny * (k1 - k0_))
            kk = k0_ + cells // (nx * ny)
            jj = (cells // nx) % ny
            ii = cells % nx
            c = np.repeat(np.stack([ii, jj, kk]), ppc_, axis=1).astype(np.float64)
            m = c.shape[1]
            x = c[0] + rng.random(m)
            y = c[1] + rng.random(m)
            z = c[2] + rng.random(m)
            dens = n_plateau * np.clip((z - z0) / ramp, 0.0, 1.0)
            u = [rng.standard_normal(m) * mom for _ in range(3)]
            yield dict(x=x, y=y, z=z, ux=u[0], uy=u[1], uz=u[2], w=dens / ppc_)

    sim.add_species_chunks("electrons", -1.0, 1.0, nx * ny * ppc * (nz - z0), slabs(ppc, 0.001))
    sim.add_species_chunks("ions", 1.0, 1836.0, nx * ny * (nz - z0), slabs(1, 0.0),
                           lfreeze=True, ldodepos=False)
    sim.add_laser_antenna(spot=(nx / 2.0, ny / 2.0, 0.0), vector=(0, 0, 1), polvector1=(1, 0, 0),
                          lambda_laser=lam, laser_w0=64.0, laser_ctau=2 * lam, t_peak=4 * lam,
                          emax1=emax_from_a0(2.0, lam), charge=-1.0)
    return sim


def make_sim(args, precision):
    from picsar_amd.sim import Simulation
    n = args.n
    return Simulation(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, args.tiles, args.tiles, args.tiles,
                      order=args.order, l_lower_order_in_v=True, pusher=args.pusher,
                      depo=args.depo, norder=2, cfl=0.95, chunk=args.chunk, precision=precision)


def build_workloads(args, precisions):
    """The workload in each of `precisions`, every one from the same arrays."""
    if args.workload == "laser":
        assert len(precisions) == 1
        return {precisions[0]: build_laser_workload(args)}
    sims = {p: make_sim(args, p) for p in precisions}
    n = args.n
    rng = np.random.default_rng(12345)
    if args.workload == "bunch":
        # a Gaussian electron bunch (sigma 20 cells, peak 64 ppc, gamma = 10
        # along z) in a uniform 1 ppc electron background: tiles hold from a
        # few hundred to tens of thousands of particles
        sig, peak, gamma = 20.0, 64.0, 10.0
        nb = int(round(peak * (2 * np.pi) ** 1.5 * sig ** 3))
        c = n / 2.0
        pos = [np.clip(rng.normal(c, sig, nb), 0.0, np.nextafter(n, 0)) for _ in range(3)]
        uz0 = np.sqrt(gamma ** 2 - 1.0)
        mom = [rng.standard_normal(nb) * 0.01, rng.standard_normal(nb) * 0.01,
               uz0 + rng.standard_normal(nb) * 0.01 * uz0]
        for s in sims.values():
            s.add_species("bunch", -1.0, 1.0, *pos, *mom, np.full(nb, OMEGA_PE ** 2 / peak))
        del pos, mom
        npart = n ** 3
        pos = [rng.random(npart) * n for _ in range(3)]
        mom = [rng.standard_normal(npart) * 0.01 for _ in range(3)]
        for s in sims.values():
            s.add_species("background", -1.0, 1.0, *pos, *mom, np.full(npart, OMEGA_PE ** 2))
        return sims
    npart = n ** 3 * args.ppc
    for name, q, m, uth in (("electrons", -1.0, 1.0, 0.01), ("ions", 1.0, 1836.0, 0.001)):
        pos = [rng.random(npart) * n for _ in range(3)]
        mom = [rng.standard_normal(npart) * uth for _ in range(3)]
        w = np.full(npart, OMEGA_PE ** 2 / args.ppc)
        for s in sims.values():
            s.add_species(name, q, m, *pos, *mom, w)
        del pos, mom, w
    return sims


def build_workload(args):
    return build_workloads(args, [args.precision])[args.precision]


def gauss_state(sim):
    """div E - rho on the nodes 0..n-1 of each axis (host copy) and the norm
    of rho there, or None where the discrete Gauss law is not that of the
    solver (FDTD order above 2)."""
    if sim.norder != 2:
        return None
    sim.gauss_residual()
    g = sim.ng
    cut = (slice(g, g + sim.nx), slice(g, g + sim.ny), slice(g, g + sim.nz))
    dv = sim.diag_grid("dive").numpy()[cut]
    rho = sim.diag_grid("rho").numpy()[cut]
    return dv - rho, float(np.linalg.norm(rho))


def choose_tiles(args):
    """Tile size by measurement (--autotune-tiles): the workload is built with
    8^3- and 16^3-node tiles (and the size given on the command line), run
    for one step and then timed over three whole steps; the fastest is kept
    in args.  One candidate is resident at a time."""
    import gc
    from picsar_amd.devmem import synchronize
    laser = args.workload == "laser"
    if laser:
        cands = sorted({8, 16, args.tile_nodes or 8})
    else:
        cands = sorted({(args.n + 1) // 8, (args.n + 1) // 16, args.tiles})
    times = {}
    for c in cands:
        a = copy.copy(args)
        if laser:
            a.tile_nodes = c
        else:
            a.tiles = c
        sim = build_workload(a)
        sim.step(1, check=False)
        synchronize()
        t0 = time.perf_counter()
        sim.step(3, check=False)
        synchronize()
        sim.check()
        times[c] = (time.perf_counter() - t0) / 3 * 1e3
        log(f"autotune tiles: {workload_name(a)}: {times[c]:.2f} ms/step")
        del sim
        gc.collect()
    best = min(times, key=times.get)
    if laser:
        args.tile_nodes = best
    else:
        args.tiles = best
    key = "tile_nodes" if laser else "tiles_per_axis"
    return {key: best, "ms_per_step": {str(k): round(v, 3) for k, v in times.items()}}


def elapsed(a, b):
    import ctypes
    from picsar_amd import _lib
    v = ctypes.c_double(0.0)
    _lib.call("pxrhip_timer_elapsed", a, b, ctypes.byref(v))
    return v.value


def timed_steps(sim, args, label):
    """W warmup steps, then K steps with HIP events around every stage and
    every fused launch (recorded on the library's stream)."""
    from picsar_amd import _lib
    from picsar_amd.devmem import synchronize
    import ctypes
    for _ in range(args.warmup):
        sim.step(1, check=False)
    sim.check()
    synchronize()
    log(f"{label}: {args.warmup} warmup steps done")
    ke0, fe0 = sim.kinetic_energy(), sim.field_energy()
    K = args.steps
    nsp = len(sim.species)
    # timer slots per step: 0 start, 1..2*nsp fused launches, then stage ends
    per = 2 * nsp + 5
    t0 = time.perf_counter()
    for s in range(K):
        b = s * per
        _lib.call("pxrhip_timer_record", b)
        for k in ("jx", "jy", "jz"):
            sim.fields[k].fill(0.0)
        for i, sp in enumerate(sim.species):
            _lib.call("pxrhip_timer_record", b + 1 + 2 * i)
            sim.gather_push_deposit_species(sp)
            _lib.call("pxrhip_timer_record", b + 2 + 2 * i)
        sim.laser_push_deposit()
        _lib.call("pxrhip_timer_record", b + 2 * nsp + 1)
        sim.current_bcs()
        _lib.call("pxrhip_timer_record", b + 2 * nsp + 2)
        sim.field_step()
        _lib.call("pxrhip_timer_record", b + 2 * nsp + 3)
        sim.particle_bcs_and_sorting()
        _lib.call("pxrhip_timer_record", b + 2 * nsp + 4)
        sim.it += 1
    synchronize()
    wall = time.perf_counter() - t0
    sim.check()
    log(f"{label}: {K} timed steps done in {wall:.2f} s")
    stage = {"gather_push_deposit": 0.0, "current_bcs": 0.0, "fdtd": 0.0, "rebin_sort": 0.0}
    launches = {i: [] for i in range(nsp)}
    total_ms = 0.0
    for s in range(K):
        b = s * per
        stage["gather_push_deposit"] += elapsed(b, b + 2 * nsp + 1)
        stage["current_bcs"] += elapsed(b + 2 * nsp + 1, b + 2 * nsp + 2)
        stage["fdtd"] += elapsed(b + 2 * nsp + 2, b + 2 * nsp + 3)
        stage["rebin_sort"] += elapsed(b + 2 * nsp + 3, b + 2 * nsp + 4)
        total_ms += elapsed(b, b + 2 * nsp + 4)
        for i in range(nsp):
            launches[i].append(elapsed(b + 1 + 2 * i, b + 2 + 2 * i))
    ni, no = ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.call("pxrhip_tiles_stats", ctypes.byref(ni), ctypes.byref(no))
    log(f"{label}: last launch: {ni.value} work items, {no.value} finished from global memory")
    # the dominant kernel: the fused launches of the species that are pushed
    # and deposit, each over its own particles
    pushed = [i for i, sp in enumerate(sim.species) if not sp.lfreeze and sp.ldodepos]
    lms = [float(np.mean(launches[i])) for i in pushed]
    bpp = BYTES_PER_PARTICLE_MIXED if sim.precision == "mixed" else BYTES_PER_PARTICLE_FP64
    ncells = sim.nx * sim.ny * sim.nz
    bytes_launch = float(np.mean([sim.species[i].npart * bpp + ncells * BYTES_PER_CELL
                                  for i in pushed]))
    per_launch_ms = float(np.mean(lms))
    N = sim.pushed_particles()  # frozen species are neither gathered, pushed nor deposited
    ke, fe = sim.kinetic_energy(), sim.field_energy()
    return dict(N=N, wall=wall, gpu_ms=total_ms, K=K, stage_ms={k: v / K for k, v in stage.items()},
                per_launch_ms=per_launch_ms, bytes_launch=bytes_launch,
                achieved=bytes_launch / (per_launch_ms * 1e-3) / 1e9,
                value=N * K / (total_ms * 1e-3), ke0=ke0, fe0=fe0, ke=ke, fe=fe,
                items=ni.value, overflow=no.value)


def graph_steps(sim, K):
    """The same steps replayed from a HIP graph of two whole steps."""
    from picsar_amd import _lib
    from picsar_amd.devmem import synchronize
    g = sim.capture_steps(2)
    sim.step_graph(g, 1, check=False)
    nrep = max(1, K // 2)
    t0g = 60000
    _lib.call("pxrhip_timer_record", t0g)
    sim.step_graph(g, nrep, check=False)
    _lib.call("pxrhip_timer_record", t0g + 1)
    synchronize()
    sim.check()
    gms = elapsed(t0g, t0g + 1) / (2 * nrep)
    del g
    log(f"graph: {2 * nrep} steps replayed at {gms:.3f} ms/step")
    return {"steps": 2 * nrep, "ms_per_step": gms}


def run_to(sim, total, label):
    """Untimed steps until sim.it == total (a progress line every half minute)."""
    from picsar_amd.devmem import synchronize
    extra = max(0, total - sim.it)
    t_log = time.perf_counter()
    for i in range(extra):
        sim.step(1, check=False)
        if time.perf_counter() - t_log > 30.0:
            synchronize()
            log(f"{label}: untimed steps: {i + 1} of {extra}")
            t_log = time.perf_counter()
    sim.check()
    synchronize()
    if extra:
        log(f"{label}: {extra} more untimed steps, {sim.it} in all")


def energy_record(sim, e_init, label):
    ke, fe = sim.kinetic_energy(), sim.field_energy()
    return {"steps": sim.it, "kinetic_initial": e_init[0], "field_initial": e_init[1],
            "kinetic_final": ke, "field_final": fe,
            "relative_drift": (ke + fe - e_init[0] - e_init[1]) / (e_init[0] + e_init[1])}


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def current_error(sa, sb):
    """Relative L2 difference of J of sa against sb, all three components."""
    num = den = 0.0
    for k in ("jx", "jy", "jz"):
        a, b = sa.fields[k].numpy(), sb.fields[k].numpy()
        num += float(np.sum((a - b) ** 2))
        den += float(np.sum(b ** 2))
    return (num / max(den, 1e-300)) ** 0.5


def first_step_errors(smx, s64, sample):
    """Step 1 of both runs, stage by stage: after the fused gather + push +
    deposit the particles of both runs are still in the same order (same
    initial arrays, same stable binning), so momenta are compared particle
    by particle on the first `sample` of each species; J after the guard
    fold."""
    out = {}
    for s in (smx, s64):
        s.gather_push_deposit()
        s.current_bcs()
    num = den = 0.0
    for spm, sp6 in zip(smx.species, s64.species):
        n = min(sample, spm.npart)
        for k in ("ux", "uy", "uz"):
            a = spm.active.arrays[k].numpy(n).astype(np.float64)
            b = sp6.active.arrays[k].numpy(n)
            num += float(np.sum((a - b) ** 2))
            den += float(np.sum(b ** 2))
    out["momentum_rel_step1"] = (num / den) ** 0.5
    out["momentum_sample"] = (f"the first {sample} particles of each species (cell-sorted order) "
                              f"after the first push")
    out["j_rel_step1"] = current_error(smx, s64)
    for s in (smx, s64):
        s.field_step()
        s.particle_bcs_and_sorting()
        s.it += 1
        s.check()
    return out


# ---------------------------------------------------------------------------
# CPU baseline: the reference's FP64 kernels on the host cores, tile by tile
# ---------------------------------------------------------------------------
def cpu_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on
    (its affinity; one GPU's share of the host on the GPU boxes)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_baseline(args, port=False, seconds=None):
    """The reference's compiled kernels, built with its production flags
    (-O3 -fopenmp), on the host cores ("reference"), or the C restatement of
    the oracle ("port") when they are absent or when ``port`` asks for it.
    The sample is whole tiles of the timed workload: the tile size and
    particles per cell of the run, electrons and ions."""
    from oracle import load_oracle, load_ref_driver, load_reference
    ref = None if port else load_reference(production=True)
    drv = None if port else load_ref_driver()
    kind = "reference" if ref is not None else "port"
    lib = ref if ref is not None else load_oracle()
    tnodes = (args.n + 1) // args.tiles     # nodes per tile of the workload
    nc, ng, ppc = tnodes - 1, 3, args.ppc
    shape = (nc + 1 + 2 * ng,) * 3
    dt = 0.95 / np.sqrt(3.0)
    threads = cpu_threads()
    rng = np.random.default_rng(7)
    fields = [np.asfortranarray(rng.standard_normal(shape) * 1e-3) for _ in range(6)]
    ntile = tnodes ** 3 * ppc
    pool = []
    for i in range(threads):
        r = np.random.default_rng(1000 + i)
        sp = []
        for q, m, uth in ((-1.0, 1.0, 0.01), (1.0, 1836.0, 0.001)):
            x, y, z = (r.uniform(-0.5, nc + 0.5, ntile) for _ in range(3))
            ux, uy, uz = (r.standard_normal(ntile) * uth for _ in range(3))
            g = 1.0 / np.sqrt(1 + ux * ux + uy * uy + uz * uz)
            sp.append((q, m, (x, y, z, ux, uy, uz, g, np.ones(ntile))))
        pool.append(sp)

    def tile_work(i):
        """The reference's per-tile work for both species of one tile:
        gather (field_gathering_plus_particle_pusher_3_3_3's arithmetic),
        Vay push, pushxyz and order-3 Esirkepov into the tile's J block."""
        done = 0
        for q, m, arrs in pool[i % threads]:
            x, y, z, ux, uy, uz, g, w = (a.copy() for a in arrs)
            n = len(x)
            e = [np.zeros(n) for _ in range(6)]
            J = [np.zeros(shape, order="F") for _ in range(3)]
            if kind == "reference":
                lib.geteb3d_energy_conserving(x, y, z, *e, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, nc, nc,
                                              nc, ng, ng, ng, 3, *fields, True, False, algo=0,
                                              lvect=64)
                lib.ebcancelpush3d(ux, uy, uz, g, *e, q, m, dt, 0)
                lib.pushxyz(x, y, z, ux, uy, uz, g, dt)
                lib.depose_jxjyjz(*J, x, y, z, ux, uy, uz, g, w, q, 0.0, 0.0, 0.0, dt, 1.0, 1.0,
                                  1.0, nc, nc, nc, ng, ng, ng, 3, algo=0)
            else:
                lib.geteb3d_energy_conserving(x, y, z, *e, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, nc, nc,
                                              nc, ng, ng, ng, 3, 3, 3, *fields, True, False)
                lib.ebcancelpush3d(ux, uy, uz, g, *e, q, m, dt, 0)
                lib.pushxyz(x, y, z, ux, uy, uz, g, dt)
                lib.depose_esirkepov(*J, x, y, z, ux, uy, uz, g, w, q, 0.0, 0.0, 0.0, dt, 1.0, 1.0,
                                     1.0, nc, nc, nc, ng, ng, ng, 3)
            done += n
        return done

    budget = args.cpu_seconds if seconds is None else seconds
    if drv is not None and ref is not None:
        # the reference's kernels under an OpenMP loop over tiles (no Python
        # in the timed loop), one tile per thread at a time as its managers
        pool_arr = np.stack([np.stack([np.stack(arrs) for _, _, arrs in sp]) for sp in pool])
        total, wall = drv.time_tiles(threads, budget, nc, ng, fields, pool_arr,
                                     [sp[0] for sp in pool[0]], [sp[1] for sp in pool[0]], dt)
        loop = "an OpenMP loop over tiles (oracle/ref_tile_driver.c)"
    else:
        total, t0 = 0, time.perf_counter()
        i = 0
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            while time.perf_counter() - t0 < budget:
                futs = [ex.submit(tile_work, i + k) for k in range(threads)]
                i += threads
                for f in futs:
                    total += f.result()
        wall = time.perf_counter() - t0
        loop = "Python threads over tiles"
    total = int(total)
    build = ("the reference's Fortran kernels built with amdflang -O3 -fopenmp -DLIBRARY "
             "(its Makefile's optimisation flags)") if kind == "reference" else \
        "the oracle's C restatement, gcc -O2"
    return {"value": total / wall, "unit": "particle-steps/s", "cores": threads, "kind": kind,
            "nproc": os.cpu_count(), "build": build,
            "sample": f"{total} particle-steps in {wall:.1f} s: whole {tnodes}^3-node tiles of "
                      f"the workload ({ppc} ppc electrons + ions), order-3 gather + Vay + "
                      f"pushxyz + order-3 Esirkepov into the tile J block, {threads} host "
                      f"threads, {loop} (FP64)"}


def roofline_record(r, args, prec):
    a = copy.copy(args)
    a.precision = prec
    name = profile_name(a)
    rec = {"bound": "hbm", "achieved": r["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": r["achieved"] / HBM_PEAK_GBS,
           "traffic": profile_value(name + "_hbm_traffic", "traffic_bytes_per_launch")
           if name else None,
           "traffic_source": f"profiles/{name}_hbm_traffic.md (rocprofv3 FETCH_SIZE, "
                             f"WRITE_SIZE, copy-calibrated)" if name else None,
           "kernel": "k_fused (gather+push+deposit, one species per launch)",
           "launch_ms": r["per_launch_ms"], "bytes_per_launch": r["bytes_launch"]}
    if name:
        # compute-side fractions of the same kernel from its counter summary:
        # VALU instructions issued and LDS instructions' active cycles per CU
        # cycle of the launch (one wave instruction per cycle per CU is the
        # VALU issue peak; FP64 FMA at that rate is the 78.6 TFLOP/s peak)
        for key in ("valu_frac", "lds_frac", "valu_insts_per_particle",
                    "lds_insts_per_particle"):
            rec[key] = profile_value(name + "_counters", key)
        rec["counters_source"] = f"profiles/{name}_counters.md"
    return rec


def leaf_bench(args):
    """Leaf deposition through the C ABI: depose_jxjyjz (current_depo_algo
    0, Esirkepov) and pxr_depose_rho_n of args.order on n^3 * ppc particles
    (u ~ 0.05 c), cell-ordered and in random order, with the runtime's
    switch-over to the binned form set to never (global atomics) and to its
    default.  Each call is timed with HIP events on the library stream; the
    best of args.steps calls is reported."""
    import ctypes
    from picsar_amd import _lib
    from picsar_amd.devmem import DeviceArray, DeviceGrid
    from picsar_amd.picsarpy import picsar as pxr
    pxr.l_nodalgrid = False
    n, ng, o = args.n, 3, args.order
    rng = np.random.default_rng(12345)
    npart = n ** 3 * args.ppc
    x, y, z = (rng.random(npart) * n for _ in range(3))
    u = [rng.standard_normal(npart) * 0.05 for _ in range(3)]
    g = 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = np.full(npart, 1e-3)
    J = [DeviceGrid(n, n, n, ng) for _ in range(3)]
    R = DeviceGrid(n, n, n, ng)
    default_min = ctypes.c_int64(0)
    _lib.call("pxrhip_get_tuning", b"leaf.bin_min", ctypes.byref(default_min))

    def timed(fn):
        best = float("inf")
        for _ in range(max(1, args.steps)):
            _lib.call("pxrhip_timer_record", 60000)
            fn()
            _lib.call("pxrhip_timer_record", 60001)
            best = min(best, elapsed(60000, 60001))
        return best

    out = {}
    orders = {"cell_ordered": np.lexsort((np.floor(x), np.floor(y), np.floor(z))),
              "random": np.arange(npart)}
    for name, perm in orders.items():
        P = [DeviceArray.from_numpy(np.ascontiguousarray(a[perm])) for a in (x, y, z, *u, g, w)]
        for form, v in (("global_atomics", 0), ("binned", default_min.value)):
            _lib.call("pxrhip_set_tuning", b"leaf.bin_min", int(v))
            tj = timed(lambda: pxr.depose_jxjyjz(*J, npart, *P, -1.0, 0, 0, 0, 0.5, 1, 1, 1, n, n,
                                                n, ng, ng, ng, o, o, o, 0))
            tr = timed(lambda: pxr.pxr_depose_rho_n(R, npart, P[0], P[1], P[2], P[7], -1.0, 0, 0,
                                                    0, 1, 1, 1, n, n, n, ng, ng, ng, o, o, o))
            out[f"{name}/{form}"] = {"depose_jxjyjz_ms": tj, "depose_rho_n_ms": tr}
            log(f"leaf {name} {form}: J {tj:.2f} ms, rho {tr:.2f} ms")
        del P
    _lib.call("pxrhip_set_tuning", b"leaf.bin_min", int(default_min.value))
    best = out["cell_ordered/binned"]["depose_jxjyjz_ms"]
    print(json.dumps({
        "metric": "leaf Esirkepov current deposition, particles/s",
        "value": npart / (best * 1e-3), "unit": "particles/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic (numpy seed 12345)",
        "config": {"workload": f"leaf depose_jxjyjz and pxr_depose_rho_n, order {o}, {n}^3 "
                               f"cells, {npart} particles of one species (u ~ 0.05 c), "
                               f"{ng} guards", "calls_per_timing": max(1, args.steps),
                   "timing": "best call, HIP events on the library stream"},
        "ms": out}))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--tiles", type=int, default=32, help="tiles per axis")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--tile-nodes", type=int, default=0,
                    help="nodes per tile along each axis of the laser workload (default 8)")
    ap.add_argument("--workload", choices=("thermal", "bunch", "laser", "leaf"), default="thermal",
                    help="thermal: periodic thermal plasma (electrons + ions); bunch: Gaussian "
                         "gamma=10 electron bunch in a 1 ppc background (uneven tiles); laser: "
                         "laser-driven underdense plasma on n x n x 2n cells (512: the large run)")
    ap.add_argument("--order", type=int, choices=(1, 2, 3), default=3,
                    help="shape order of the gather and the deposition (thermal and bunch)")
    ap.add_argument("--pusher", choices=("vay", "boris"), default="vay")
    ap.add_argument("--depo", choices=("esirkepov", "direct"), default="esirkepov")
    ap.add_argument("--autotune-tiles", action="store_true",
                    help="choose the tile size by timing 8^3- and 16^3-node tiles first")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch parameters with picsar_amd.autotune before timing")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed",
                    help="fp64: every particle array FP64; mixed: FP64 positions, FP32 momenta, "
                         "gamma and weights, FP32 gather/push/per-particle current, FP64 J and "
                         "FDTD")
    ap.add_argument("--no-compare", action="store_true",
                    help="time only --precision (by default a mixed-precision run is timed "
                         "beside an FP64 run of the same arrays, and their difference reported)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="total steps (warmup + timed + untimed) over which the energy drift "
                         "is reported")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true",
                    help="after the eager timed steps (which give the stage and launch timings), "
                         "replay the same number of steps from a HIP graph of two steps and "
                         "report that as the value")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch parameter of the tiled kernels (pxrhip_set_tuning)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed region, copy 1 GiB with one plain device copy "
                         "kernel (the reference point of HBM traffic counters)")
    ap.add_argument("--sample", type=int, default=1 << 24,
                    help="particles per species whose momenta are compared after step 1")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the path runs "
              "on one MI355X and is not sharded over several GPUs", file=sys.stderr)
        sys.exit(2)
    from picsar_amd import _lib
    import ctypes
    _lib.call("pxrhip_init", 0)
    for kv in args.tune:
        k, v = kv.split("=")
        _lib.call("pxrhip_set_tuning", k.encode(), int(v))
    if args.workload == "leaf":
        leaf_bench(args)
        return
    tiles_tuned = choose_tiles(args) if args.autotune_tiles else None
    compare = (not args.no_compare and args.workload != "laser" and not args.graph)
    precs = [args.precision] + ([p for p in ("fp64", "mixed") if p != args.precision]
                                if compare else [])
    sims = build_workloads(args, precs)
    main_sim = sims[args.precision]
    log(f"workload ready: {main_sim.total_particles()} particles "
        f"({main_sim.pushed_particles()} pushed), "
        f"{main_sim.nx}x{main_sim.ny}x{main_sim.nz} cells, precisions {precs}")
    tuned = None
    if args.autotune:
        from picsar_amd.autotune import autotune
        tuned = autotune(main_sim)
        log(f"autotune: {tuned['best']} ({tuned['ms_per_step']:.2f} ms/step gather+push+deposit)")
    e_init = {p: (s.kinetic_energy(), s.field_energy()) for p, s in sims.items()}
    gauss0 = {p: gauss_state(s) for p, s in sims.items()} if args.drift_steps > 0 else {}
    perr = None
    if compare:
        perr = first_step_errors(sims["mixed"], sims["fp64"], args.sample)
        log(f"step 1: J error {perr['j_rel_step1']:.2e}, momentum error "
            f"{perr['momentum_rel_step1']:.2e} (mixed against FP64)")
    runs = {}
    for p in precs:
        runs[p] = timed_steps(sims[p], args, p)
    graph = graph_steps(main_sim, args.steps) if args.graph else None
    if compare:
        perr["steps"] = sims["mixed"].it
        perr["j_rel"] = current_error(sims["mixed"], sims["fp64"])
        k6 = runs["fp64"]["ke"]
        perr["kinetic_energy_rel"] = abs(runs["mixed"]["ke"] - k6) / k6
    drift = {}
    gauss = {}
    for p in precs:
        run_to(sims[p], args.drift_steps, p)
        drift[p] = energy_record(sims[p], e_init[p], p)
        if gauss0.get(p) is not None:
            r1, rho1 = gauss_state(sims[p])
            gauss[p] = {"steps": sims[p].it, "rho_norm_initial": gauss0[p][1],
                        "rho_norm_final": rho1,
                        "residual_change_rel": float(np.linalg.norm(r1 - gauss0[p][0])) /
                        gauss0[p][1]}
    if compare and args.drift_steps > 0:
        perr["j_rel_final"] = current_error(sims["mixed"], sims["fp64"])
        perr["final_steps"] = sims["mixed"].it
        perr["energy_drift_fp64"] = drift["fp64"]["relative_drift"]
        perr["energy_drift_mixed"] = drift["mixed"]["relative_drift"]
        perr["note"] = ("J and momentum differences of the mixed-precision run against the FP64 "
                        "run from the same arrays; after many steps the two runs separate as any "
                        "two plasma runs with different round-off do, so the later differences "
                        "bound the error rather than measure it")
    tuning = {}
    for k in ("fused.lds_bytes", "fused.wg_per_cu", "fused.spread", "fused.chunk",
              "fused.j_lds_bytes", "fused.jpad", "fused.margin_pct", "fused.sstride",
              "mixed.jacc", "fused.jacc", "mixed.lds_bytes", "mixed.wg_per_cu",
              "last.wg_per_cu",
              "last.lds_bytes", "rebin.stable"):
        v = ctypes.c_int64(0)
        _lib.call("pxrhip_get_tuning", k.encode(), ctypes.byref(v))
        tuning[k] = v.value
    r = runs[args.precision]
    value = r["value"]
    if graph is not None:  # --graph: the replayed steps are the measurement
        graph["eager_ms_per_step"] = r["gpu_ms"] / r["K"]
        value = r["N"] / (graph["ms_per_step"] * 1e-3)
    gpd = r["N"] / (r["stage_ms"]["gather_push_deposit"] * 1e-3)
    cells = main_sim.nx * main_sim.ny * main_sim.nz
    out = {
        "metric": "particle-steps/s (gather+push+deposit), FP32-mixed vs FP64, % of HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["gpu_ms"] / r["K"] if graph is None else graph["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (gather, push, per-particle "
                                                          "current) + f64 (positions, J, FDTD)",
        "data": "synthetic (numpy seed 12345, resident on the GPU before timing)",
        "config": {"workload": workload_name(args), "particles": r["N"],
                   "particles_frozen": main_sim.total_particles() - r["N"], "cells": cells,
                   "tuning": tuning, "autotuned": tuned is not None,
                   "tiles_autotuned": tiles_tuned,
                   "step": "HIP graph of two steps, replayed" if graph else "eager"},
        "graph": graph,
        "stages": {"ms_per_step": r["stage_ms"],
                   "gather_push_deposit_particle_steps_per_s": gpd,
                   "wall_s_timed_region": r["wall"],
                   "work_items_last_launch": r["items"],
                   "work_items_from_global_memory": r["overflow"]},
        "roofline": roofline_record(r, args, args.precision),
        "energy": dict(drift[args.precision],
                       relative_drift_timed_region=(r["ke"] + r["fe"] - r["ke0"] - r["fe0"]) /
                       (r["ke0"] + r["fe0"]),
                       note=("the laser antenna injects energy: the change is the injected "
                             "energy, not an error of the step") if args.workload == "laser"
                       else "periodic, no sources: the change is the step's error"),
        "gauss_law": gauss.get(args.precision),
    }
    if compare:
        other = [p for p in precs if p != args.precision][0]
        ro = runs[other]
        out[other] = {"value": ro["value"], "ms_per_step": ro["gpu_ms"] / ro["K"],
                      "stages_ms_per_step": ro["stage_ms"],
                      "roofline": roofline_record(ro, args, other),
                      "energy": drift[other], "gauss_law": gauss.get(other)}
        out["mixed_over_fp64"] = runs["mixed"]["value"] / runs["fp64"]["value"]
        out["precision_error"] = perr
    del sims, main_sim
    if args.calibrate_copy:
        from picsar_amd.devmem import DeviceArray, synchronize
        a = DeviceArray(1 << 27)
        b = DeviceArray(1 << 27)
        a.fill(1.0)
        synchronize()
        b.copy_from(a)
        synchronize()
        out["calibration_copy_bytes"] = 2 * a.nbytes
        del a, b
    if not args.no_cpu_baseline and args.workload == "thermal":
        log(f"timing the CPU baseline for about {args.cpu_seconds:.0f} s")
        out["cpu_baseline"] = cpu_baseline(args)
        if out["cpu_baseline"]["kind"] == "reference":
            # the oracle's C restatement of the same work, on the same cores
            out["cpu_restatement"] = cpu_baseline(args, port=True,
                                                  seconds=max(2.0, args.cpu_seconds / 2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
