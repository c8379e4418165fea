#!/usr/bin/env python3
"""Throughput of the PIC hot path on one MI355X.

Workload (the plain FP64 GPU configuration): a periodic thermal plasma on a
128^3 grid, dx = dy = dz = 1, dt = 0.95 CFL, electrons and ions at 16
particles per cell each (67,108,864 macro-particles), Maxwellian momenta
(u_th = 0.01 c and 0.001 c), order-3 energy-conserving gather, Vay pusher,
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
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE_FP64 = 120  # read x,y,z,ux,uy,uz,gaminv,w; write x..gaminv
BYTES_PER_CELL = 72            # read E, B (48 B); write J (24 B)


def build_workload(args):
    import torch
    from picsar_amd.sim import Simulation
    n = args.n
    sim = Simulation(n, n, n, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, args.tiles, args.tiles, args.tiles,
                     order=3, l_lower_order_in_v=True, pusher="vay", depo="esirkepov", norder=2,
                     cfl=0.95, chunk=args.chunk)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(12345)
    ncell = n ** 3
    for name, q, m, uth in (("electrons", -1.0, 1.0, 0.01), ("ions", 1.0, 1836.0, 0.001)):
        npart = ncell * args.ppc
        pos = [torch.rand(npart, generator=gen, device="cuda", dtype=torch.float64) * n
               for _ in range(3)]
        mom = [torch.randn(npart, generator=gen, device="cuda", dtype=torch.float64) * uth
               for _ in range(3)]
        w = torch.ones(npart, device="cuda", dtype=torch.float64)
        sim.add_species(name, q, m, *pos, *mom, w)
        del pos, mom, w
    torch.cuda.synchronize()
    return sim


def run_gpu(args):
    import torch
    sim = build_workload(args)
    N = sim.total_particles()
    for _ in range(args.warmup):
        sim.step(1, check=False)
    sim.check()
    torch.cuda.synchronize()
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    stage_ms = {"gather_push_deposit": 0.0, "current_bcs": 0.0, "fdtd": 0.0, "rebin_sort": 0.0}
    fused_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = [ev() for _ in range(6)]
        e[0].record()
        f = sim.fields
        for k in ("jx", "jy", "jz"):
            f[k].zero_()
        for sp in sim.species:
            a, b = ev(), ev()
            a.record()
            sim.gather_push_deposit_species(sp)
            b.record()
            fused_ms.append((a, b))
        e[1].record()
        sim.current_bcs()
        e[2].record()
        sim.field_step()
        e[3].record()
        sim.particle_bcs_and_sorting()
        e[4].record()
        sim.it += 1
        stage_ms_step = e
        stage_ms.setdefault("_events", []).append(stage_ms_step)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sim.check()
    evs = stage_ms.pop("_events")
    for e in evs:
        stage_ms["gather_push_deposit"] += e[0].elapsed_time(e[1])
        stage_ms["current_bcs"] += e[1].elapsed_time(e[2])
        stage_ms["fdtd"] += e[2].elapsed_time(e[3])
        stage_ms["rebin_sort"] += e[3].elapsed_time(e[4])
    gpu_ms = sum(stage_ms.values())
    launches = [a.elapsed_time(b) for a, b in fused_ms]
    per_launch_ms = float(np.mean(launches))
    npart_launch = N / len(sim.species)
    ncells = sim.nx * sim.ny * sim.nz
    bytes_launch = npart_launch * BYTES_PER_PARTICLE_FP64 + ncells * BYTES_PER_CELL
    achieved = bytes_launch / (per_launch_ms * 1e-3) / 1e9
    K = args.steps
    return dict(sim=sim, N=N, wall=wall, gpu_ms=gpu_ms, stage_ms={k: v / K for k, v in stage_ms.items()},
                per_launch_ms=per_launch_ms, achieved=achieved, bytes_launch=bytes_launch,
                ke=sim.kinetic_energy(), fe=sim.field_energy())


# ---------------------------------------------------------------------------
# CPU baseline: the reference's FP64 kernels on the host cores, tile by tile
# ---------------------------------------------------------------------------
def cpu_baseline(args):
    from oracle import load_oracle, load_reference
    ref = load_reference()
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

    budget = args.cpu_seconds
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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="cells per axis")
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--tiles", type=int, default=8, help="tiles per axis")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this build runs the path on one MI355X; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)
    r = run_gpu(args)
    K = args.steps
    value = r["N"] * K / (r["gpu_ms"] * 1e-3)
    gpd = r["N"] / (r["stage_ms"]["gather_push_deposit"] * 1e-3)
    out = {
        "metric": "particle-steps/s (gather+push+deposit), FP32-mixed vs FP64, % of HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": r["gpu_ms"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded on device)",
        "config": {"workload": f"thermal plasma {args.n}^3 periodic, {args.ppc} ppc electrons + "
                               f"ions, order-3 gather, Vay, order-3 Esirkepov, Yee, "
                               f"{(args.n + 1) // args.tiles}^3-node tiles, FP64",
                   "particles": r["N"], "cells": args.n ** 3},
        "stages": {"ms_per_step": r["stage_ms"],
                   "gather_push_deposit_particle_steps_per_s": gpd,
                   "wall_s_timed_region": r["wall"]},
        "roofline": {"bound": "hbm", "achieved": r["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": r["achieved"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused (gather+push+deposit, one species per launch)",
                     "launch_ms": r["per_launch_ms"],
                     "bytes_per_launch": r["bytes_launch"]},
        "energy": {"kinetic": r["ke"], "field": r["fe"]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
