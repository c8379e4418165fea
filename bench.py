#!/usr/bin/env python3
"""Throughput of the PIC hot path on one MI355X.

Workload (the plain FP64 GPU configuration): a periodic thermal plasma on a
128^3 grid, dx = dy = dz = 1, dt = 0.95 CFL, electrons and ions at 16
particles per cell each (67,108,864 macro-particles), Maxwellian momenta
(u_th = 0.01 c and 0.001 c), weights giving an electron plasma frequency
0.02 (Debye length 0.5 cell, w_pe*dt = 0.011), order-3 energy-conserving
gather, Vay pusher,
order-3 Esirkepov deposition, 2nd-order Yee FDTD, 16^3-node tiles, FP64.
Synthetic data generated on the device from a fixed seed.

A timed step is the whole PIC step: fused gather+push+deposit for both
species, current guard fold, FDTD (B half, E, B half with guard
exchanges), periodic re-binning and per-tile sort.  `value` is
particle-steps/s over that whole step; the gather+push+deposit stage alone
is reported in `stages`.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)

BYTES_PER_PARTICLE_FP64 = 120  # read x,y,z,ux,uy,uz,gaminv,w; write x..gaminv
BYTES_PER_PARTICLE_MIXED = 84  # positions FP64 (24 + 24), u, gaminv, w FP32 (20 + 16)
BYTES_PER_CELL = 72            # read E, B (48 B); write J (24 B)
OMEGA_PE = 0.02                # electron plasma frequency (c = dx = 1)


def traffic_profile(args):
    """The committed counter summary that holds this run's fused launch:
    the default workload and launch parameters, in FP64 or mixed precision."""
    if (args.n, args.ppc, args.tiles, args.chunk) != (128, 16, 8, 0) or args.tune or \
            args.workload != "thermal" or args.autotune or \
            (args.order, args.pusher, args.depo) != (3, "vay", "esirkepov"):
        return None
    return "fused_c2_hbm_traffic.md" if args.precision == "fp64" else \
        "fused_c2_mixed_hbm_traffic.md"


def profiled_traffic(args):
    """HBM bytes per fused launch from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a copy-kernel calibration),
    valid for the default workload and launch parameters only."""
    name = traffic_profile(args)
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch ="):
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
                f"to 4 ppc (immobile ions neither move nor deposit and are not carried), "
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
    plateau at 0.1 critical density; a Gaussian laser (a0 = 2, lambda = 32
    cells, waist 64 cells) launched by an antenna in the plane z = 0.  The
    electrons are generated and uploaded slab by slab."""
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
    nper = nx * ny * ppc

    def slabs(step=8):
        for k0_ in range(z0, nz, step):
            if (k0_ - z0) % (16 * step) == 0:
                log(f"generating electrons: z planes {k0_} of {nz}")
            k1 = min(k0_ + step, nz)
            cells = np.arange(nx * ny * (k1 - k0_))
            kk = k0_ + cells // (nx * ny)
            jj = (cells // nx) % ny
            ii = cells % nx
            c = np.repeat(np.stack([ii, jj, kk]), ppc, axis=1).astype(np.float64)
            m = c.shape[1]
            x = c[0] + rng.random(m)
            y = c[1] + rng.random(m)
            z = c[2] + rng.random(m)
            dens = n_plateau * np.clip((z - z0) / ramp, 0.0, 1.0)
            u = [rng.standard_normal(m) * 0.001 for _ in range(3)]
            yield dict(x=x, y=y, z=z, ux=u[0], uy=u[1], uz=u[2], w=dens / ppc)

    sim.add_species_chunks("electrons", -1.0, 1.0, nper * (nz - z0), slabs())
    sim.add_laser_antenna(spot=(nx / 2.0, ny / 2.0, 0.0), vector=(0, 0, 1), polvector1=(1, 0, 0),
                          lambda_laser=lam, laser_w0=64.0, laser_ctau=2 * lam, t_peak=4 * lam,
                          emax1=emax_from_a0(2.0, lam), charge=-1.0)
    return sim


def build_workload(args):
    from picsar_amd.sim import Simulation
    if args.workload == "laser":
        return build_laser_workload(args)
    n = args.n
    sim = Simulation(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, args.tiles, args.tiles, args.tiles,
                     order=args.order, l_lower_order_in_v=True, pusher=args.pusher,
                     depo=args.depo, norder=2,
                     cfl=0.95, chunk=args.chunk, precision=args.precision)
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
        sim.add_species("bunch", -1.0, 1.0, *pos, *mom, np.full(nb, OMEGA_PE ** 2 / peak))
        del pos, mom
        npart = n ** 3
        pos = [rng.random(npart) * n for _ in range(3)]
        mom = [rng.standard_normal(npart) * 0.01 for _ in range(3)]
        sim.add_species("background", -1.0, 1.0, *pos, *mom, np.full(npart, OMEGA_PE ** 2))
        return sim
    npart = n ** 3 * args.ppc
    for name, q, m, uth in (("electrons", -1.0, 1.0, 0.01), ("ions", 1.0, 1836.0, 0.001)):
        pos = [rng.random(npart) * n for _ in range(3)]
        mom = [rng.standard_normal(npart) * uth for _ in range(3)]
        w = np.full(npart, OMEGA_PE ** 2 / args.ppc)
        sim.add_species(name, q, m, *pos, *mom, w)
        del pos, mom, w
    return sim


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


def run_gpu(args):
    from picsar_amd import _lib
    from picsar_amd.devmem import synchronize
    import ctypes

    def elapsed(a, b):
        v = ctypes.c_double(0.0)
        _lib.call("pxrhip_timer_elapsed", a, b, ctypes.byref(v))
        return v.value

    tiles_tuned = choose_tiles(args) if args.autotune_tiles else None
    sim = build_workload(args)
    N = sim.total_particles()
    log(f"workload ready: {N} particles, {sim.nx}x{sim.ny}x{sim.nz} cells")
    tuned = None
    if args.autotune:
        from picsar_amd.autotune import autotune
        tuned = autotune(sim)
        log(f"autotune: {tuned['best']} ({tuned['ms_per_step']:.2f} ms/step gather+push+deposit)")
    ke_init, fe_init = sim.kinetic_energy(), sim.field_energy()
    # the Gauss-law check goes with the drift leg (rho of every particle by
    # the leaf deposition, outside the timed region)
    gauss0 = gauss_state(sim) if args.drift_steps > 0 else None
    for _ in range(args.warmup):
        sim.step(1, check=False)
    sim.check()
    synchronize()
    log(f"{args.warmup} warmup steps done")
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
    log(f"{K} timed steps done in {wall:.2f} s")
    stage = {"gather_push_deposit": 0.0, "current_bcs": 0.0, "fdtd": 0.0, "rebin_sort": 0.0}
    launches = []
    total_ms = 0.0
    for s in range(K):
        b = s * per
        stage["gather_push_deposit"] += elapsed(b, b + 2 * nsp + 1)
        stage["current_bcs"] += elapsed(b + 2 * nsp + 1, b + 2 * nsp + 2)
        stage["fdtd"] += elapsed(b + 2 * nsp + 2, b + 2 * nsp + 3)
        stage["rebin_sort"] += elapsed(b + 2 * nsp + 3, b + 2 * nsp + 4)
        total_ms += elapsed(b, b + 2 * nsp + 4)
        for i in range(nsp):
            launches.append(elapsed(b + 1 + 2 * i, b + 2 + 2 * i))
    per_launch_ms = float(np.mean(launches))
    ni, no = ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.call("pxrhip_tiles_stats", ctypes.byref(ni), ctypes.byref(no))
    log(f"last launch: {ni.value} work items, {no.value} finished from global memory")
    npart_launch = N / nsp
    ncells = sim.nx * sim.ny * sim.nz
    bpp = BYTES_PER_PARTICLE_MIXED if args.precision == "mixed" else BYTES_PER_PARTICLE_FP64
    bytes_launch = npart_launch * bpp + ncells * BYTES_PER_CELL
    achieved = bytes_launch / (per_launch_ms * 1e-3) / 1e9
    ke, fe = sim.kinetic_energy(), sim.field_energy()
    done = args.warmup + K
    graph = None
    if args.graph:
        # the same steps replayed from a HIP graph of two whole steps
        # (particle counts do not change in a periodic box)
        g = sim.capture_steps(2)
        sim.step_graph(g, 1, check=False)
        nrep = max(1, K // 2)
        t0g = per * K + 10
        _lib.call("pxrhip_timer_record", t0g)
        sim.step_graph(g, nrep, check=False)
        _lib.call("pxrhip_timer_record", t0g + 1)
        synchronize()
        sim.check()
        gms = elapsed(t0g, t0g + 1) / (2 * nrep)
        sim.it += 2 * (nrep + 1)
        done += 2 * (nrep + 1)
        graph = {"steps": 2 * nrep, "ms_per_step": gms}
        del g
        log(f"graph: {2 * nrep} steps replayed at {gms:.3f} ms/step")
        ke, fe = sim.kinetic_energy(), sim.field_energy()
    # untimed steps up to the energy-drift horizon (counted from the start)
    extra = max(0, args.drift_steps - done)
    t_log = time.perf_counter()
    for i in range(extra):
        sim.step(1, check=False)
        if time.perf_counter() - t_log > 30.0:  # a progress line every half minute
            synchronize()
            log(f"untimed steps: {i + 1} of {extra}")
            t_log = time.perf_counter()
    sim.check()
    synchronize()
    if extra:
        log(f"{extra} more untimed steps for the energy drift over {done + extra} steps")
    ke_d, fe_d = sim.kinetic_energy(), sim.field_energy()
    gauss = None
    if gauss0 is not None:
        r1, rho1 = gauss_state(sim)
        gauss = {"steps": done + extra, "rho_norm_initial": gauss0[1], "rho_norm_final": rho1,
                 "residual_change_rel": float(np.linalg.norm(r1 - gauss0[0])) / gauss0[1],
                 "what": "|(div E - rho)_final - (div E - rho)_initial| / |rho_initial| on the "
                         "nodes 0..n-1 of each axis"}
        del r1
    return dict(gauss=gauss, sim=sim, N=N, wall=wall, gpu_ms=total_ms, stage_ms={k: v / K for k, v in stage.items()},
                per_launch_ms=per_launch_ms, achieved=achieved, bytes_launch=bytes_launch,
                ke0=ke0, fe0=fe0, ke=ke, fe=fe, ke_init=ke_init, fe_init=fe_init,
                drift_steps=done + extra, ke_d=ke_d, fe_d=fe_d, tuned=tuned, graph=graph,
                tiles_tuned=tiles_tuned)


# ---------------------------------------------------------------------------
# CPU baseline: the reference's FP64 kernels on the host cores, tile by tile
# ---------------------------------------------------------------------------
def cpu_baseline(args, port=False, seconds=None):
    """The reference's compiled kernels on the host cores ("reference"), or
    the C restatement of the oracle ("port") when they are absent or when
    ``port`` asks for it."""
    from oracle import load_oracle, load_reference
    ref = None if port else load_reference()
    kind = "reference" if ref is not None else "port"
    lib = ref if ref is not None else load_oracle()
    nc, ng, ppc = 15, 3, args.ppc        # one 16^3-node tile of the workload
    shape = (nc + 1 + 2 * ng,) * 3
    dt = 0.95 / np.sqrt(3.0)
    threads = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(7)
    fields = [np.asfortranarray(rng.standard_normal(shape) * 1e-3) for _ in range(6)]
    ntile = 16 ** 3 * ppc
    pool = []
    for i in range(threads):
        r = np.random.default_rng(1000 + i)
        sp = []
        for q, m, uth in ((-1.0, 1.0, 0.01), (1.0, 1836.0, 0.001)):
            x, y, z = (r.uniform(-0.5, 15.5, ntile) for _ in range(3))
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
    total, t0 = 0, time.perf_counter()
    i = 0
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        while time.perf_counter() - t0 < budget:
            futs = [ex.submit(tile_work, i + k) for k in range(threads)]
            i += threads
            for f in futs:
                total += f.result()
    wall = time.perf_counter() - t0
    return {"value": total / wall, "unit": "particle-steps/s", "cores": threads, "kind": kind,
            "sample": f"{total} particle-steps in {wall:.1f} s: whole 16^3-node tiles of the "
                      f"workload ({ppc} ppc electrons + ions), order-3 gather + Vay + pushxyz + "
                      f"order-3 Esirkepov into the tile J block, {threads} host threads over "
                      f"tiles (FP64)"}


def precision_comparison(args):
    """The same step at 256^3 cells, 8 ppc electrons + ions, 8^3-node tiles
    (the fastest of the tile sizes measured at this size, in both precisions),
    order-3 Esirkepov, in FP64 and in mixed precision (20 timed steps each):
    the mixed-precision step against the FP64 one."""
    import copy
    out = {}
    for prec in ("fp64", "mixed"):
        a = copy.copy(args)
        a.n, a.ppc, a.tiles, a.precision = 256, 8, 32, prec
        a.steps, a.warmup, a.drift_steps, a.autotune, a.graph = 20, 2, 0, False, False
        log(f"comparison: {workload_name(a)}")
        r = run_gpu(a)
        del r["sim"]
        out[prec] = {"value": r["N"] * a.steps / (r["gpu_ms"] * 1e-3),
                     "ms_per_step": r["gpu_ms"] / a.steps,
                     "gather_push_deposit_ms_per_step": r["stage_ms"]["gather_push_deposit"],
                     "launch_ms": r["per_launch_ms"],
                     "roofline_frac": r["achieved"] / HBM_PEAK_GBS}
        out["workload"] = workload_name(a).replace(", mixed precision", "").replace(", FP64", "")
        out["particles"] = r["N"]
    out["unit"] = "particle-steps/s"
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="cells per axis")
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--tiles", type=int, default=8, help="tiles per axis")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--tile-nodes", type=int, default=0,
                    help="nodes per tile along each axis of the laser workload (default 8)")
    ap.add_argument("--workload", choices=("thermal", "bunch", "laser"), default="thermal",
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
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="fp64: every particle array FP64; mixed: FP64 positions, FP32 momenta, "
                         "gamma and weights, FP32 gather/push/LDS current, FP64 J and FDTD")
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
    ap.add_argument("--no-comparison", action="store_true",
                    help="skip the FP64 / mixed-precision comparison at 256^3 cells, 8 ppc, "
                         "8^3-node tiles that follows the default workload")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed region, copy 1 GiB with one plain device copy "
                         "kernel (the reference point of HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the path runs "
              "on one MI355X and is not sharded over several GPUs", file=sys.stderr)
        sys.exit(2)
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)
    tuning = {}
    for kv in args.tune:
        k, v = kv.split("=")
        _lib.call("pxrhip_set_tuning", k.encode(), int(v))
    r = run_gpu(args)
    for k in ("fused.lds_bytes", "fused.wg_per_cu", "fused.spread", "fused.chunk",
              "fused.j_lds_bytes", "fused.jpad", "fused.margin_pct", "fused.sstride",
              "mixed.jacc",
              "mixed.lds_bytes", "mixed.wg_per_cu", "last.wg_per_cu", "last.lds_bytes"):
        import ctypes
        v = ctypes.c_int64(0)
        _lib.call("pxrhip_get_tuning", k.encode(), ctypes.byref(v))
        tuning[k] = v.value
    K = args.steps
    value = r["N"] * K / (r["gpu_ms"] * 1e-3)
    if r["graph"] is not None:  # --graph: the replayed steps are the measurement
        r["graph"]["eager_ms_per_step"] = r["gpu_ms"] / K
        value = r["N"] / (r["graph"]["ms_per_step"] * 1e-3)
    gpd = r["N"] / (r["stage_ms"]["gather_push_deposit"] * 1e-3)
    out = {
        "metric": "particle-steps/s (gather+push+deposit), FP32-mixed vs FP64, % of HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": r["gpu_ms"] / K if r["graph"] is None else r["graph"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (gather, push, LDS current) + f64 "
                                                          "(positions, J, FDTD)",
        "data": "synthetic (numpy seed 12345, resident on the GPU before timing)",
        "config": {"workload": workload_name(args), "particles": r["N"], "cells": args.n ** 3,
                   "tuning": tuning, "autotuned": r["tuned"] is not None,
                   "tiles_autotuned": r["tiles_tuned"],
                   "step": "HIP graph of two steps, replayed" if r["graph"] else "eager"},
        "graph": r["graph"],
        "stages": {"ms_per_step": r["stage_ms"],
                   "gather_push_deposit_particle_steps_per_s": gpd,
                   "wall_s_timed_region": r["wall"]},
        "roofline": {"bound": "hbm", "achieved": r["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": r["achieved"] / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args),
                     "traffic_source": (f"profiles/{traffic_profile(args)} (rocprofv3 "
                                        "FETCH_SIZE, WRITE_SIZE, copy-calibrated)")
                     if traffic_profile(args) else None,
                     "kernel": "k_fused (gather+push+deposit, one species per launch)",
                     "launch_ms": r["per_launch_ms"],
                     "bytes_per_launch": r["bytes_launch"]},
        "energy": {"kinetic_initial": r["ke_init"], "field_initial": r["fe_init"],
                   "steps": r["drift_steps"], "kinetic_final": r["ke_d"],
                   "field_final": r["fe_d"],
                   "relative_drift": (r["ke_d"] + r["fe_d"] - r["ke_init"] - r["fe_init"]) /
                   (r["ke_init"] + r["fe_init"]),
                   "relative_drift_timed_region": (r["ke"] + r["fe"] - r["ke0"] - r["fe0"]) /
                   (r["ke0"] + r["fe0"]),
                   "note": ("the laser antenna injects energy: the change is the injected "
                            "energy, not an error of the step") if args.workload == "laser"
                   else "periodic, no sources: the change is the step's error"},
        "gauss_law": r["gauss"],
    }
    del r["sim"]
    if (not args.no_comparison and args.workload == "thermal" and args.n == 128
            and args.precision == "fp64" and args.order == 3 and not args.graph):
        out["precision_comparison"] = precision_comparison(args)
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
    if not args.no_cpu_baseline:
        log(f"timing the CPU baseline for about {args.cpu_seconds:.0f} s")
        out["cpu_baseline"] = cpu_baseline(args)
        if out["cpu_baseline"]["kind"] == "reference":
            # the oracle's C restatement of the same work, on the same cores
            out["cpu_restatement"] = cpu_baseline(args, port=True,
                                                  seconds=max(2.0, args.cpu_seconds / 2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
