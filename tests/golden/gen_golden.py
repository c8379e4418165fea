"""Write the golden vectors under tests/golden/ from the reference's own kernels.

Run in a tree where ``oracle/_ref/libpxrref.so`` has been built from the PICSAR
sources (``bash oracle/build_ref.sh``):

    python tests/golden/gen_golden.py [family ...]   # default: every family

Each ``.npz`` holds the seeded inputs of one kernel family and the outputs the
reference's compiled FP64 routines produced for them.  Inputs include the
edge cases the kernels are sensitive to: positions on nodes and half nodes,
particles in the guard region, particles crossing cell faces in both
directions, zero velocity, ties in the sort keys and keys on bin edges.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import load_reference  # noqa: E402

# one small grid shared by all fixtures: 6x5x4 cells, 3 guard nodes
NX, NY, NZ, NG = 6, 5, 4, 3
DX, DY, DZ = 0.8, 1.0, 1.25
XMIN, YMIN, ZMIN = -0.4, 1.0, 2.5
SHAPE = (NX + 1 + 2 * NG, NY + 1 + 2 * NG, NZ + 1 + 2 * NG)


def grid(rng):
    return np.asfortranarray(rng.standard_normal(SHAPE))


def positions(rng, n, lo=-0.5, hi=None):
    """Particle positions in normalised cell units inside [lo, n_cells+0.5),
    with a share of them exactly on nodes and half nodes."""
    out = []
    for ncell, x0, d in ((NX, XMIN, DX), (NY, YMIN, DY), (NZ, ZMIN, DZ)):
        u = rng.uniform(lo, ncell + 0.5 if hi is None else hi, n)
        k = n // 8
        u[:k] = rng.integers(0, ncell + 1, k).astype(np.float64)            # on nodes
        u[k:2 * k] = rng.integers(0, ncell, k).astype(np.float64) + 0.5   # half nodes
        out.append(x0 + u * d)
    return out


def gen_gather(ref, rng):
    n = 256
    xp, yp, zp = positions(rng, n)
    F = [grid(rng) for _ in range(6)]
    d = dict(xp=xp, yp=yp, zp=zp, exg=F[0], eyg=F[1], ezg=F[2], bxg=F[3], byg=F[4], bzg=F[5])
    for order in (1, 2, 3):
        for lo in (1, 0):
            for algo in (0, 1):
                e = [np.zeros(n) for _ in range(6)]
                ref.geteb3d_energy_conserving(xp, yp, zp, *e, XMIN, YMIN, ZMIN, DX, DY, DZ, NX,
                                              NY, NZ, NG, NG, NG, order, *F, bool(lo),
                                              algo=algo)
                d[f"o{order}_lo{lo}_a{algo}"] = np.stack(e)
    np.savez_compressed(os.path.join(HERE, "gather.npz"), **d)


def gen_push(ref, rng):
    n = 512
    u = [rng.standard_normal(n) * s for s in (0.3, 0.3, 3.0)]   # up to gamma ~ 10
    u[0][:8] = 0.0
    u[1][:8] = 0.0
    u[2][:8] = 0.0
    E = [rng.standard_normal(n) * 0.5 for _ in range(3)]
    B = [rng.standard_normal(n) * 2.0 for _ in range(3)]
    g0 = 1.0 / np.sqrt(1.0 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    x = [rng.uniform(-3, 3, n) for _ in range(3)]
    q, m, dt = -1.0, 1.0, 0.45
    d = dict(ux=u[0], uy=u[1], uz=u[2], gaminv=g0, ex=E[0], ey=E[1], ez=E[2], bx=B[0],
             by=B[1], bz=B[2], x=x[0], y=x[1], z=x[2], q=q, m=m, dt=dt)
    uu = [a.copy() for a in u]
    g = g0.copy()
    ref.boris_push_u_3d(*uu, g, *E, *B, q, m, dt)
    d["boris"] = np.stack(uu + [g])
    for which in (0, 1, 2):
        uu = [a.copy() for a in u]
        g = g0.copy()
        ref.ebcancelpush3d(*uu, g, *E, *B, q, m, dt, which)
        d[f"vay{which}"] = np.stack(uu + [g])
    g = np.zeros(n)
    ref.set_gamma(*u, g)
    d["set_gamma"] = g
    xx = [a.copy() for a in x]
    ref.pushxyz(*xx, *u, g0, dt)
    d["pushxyz"] = np.stack(xx)
    np.savez_compressed(os.path.join(HERE, "push.npz"), **d)


def gen_deposit(ref, rng):
    n = 256
    dt = 0.5
    xp, yp, zp = positions(rng, n, lo=0.0, hi=None)
    # velocities up to ~0.9 c, so old positions cross faces both ways
    u = [rng.uniform(-2.0, 2.0, n) for _ in range(3)]
    for a in u:
        a[:8] = 0.0
    g = 1.0 / np.sqrt(1.0 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 1.5, n)
    q = -1.3
    d = dict(xp=xp, yp=yp, zp=zp, ux=u[0], uy=u[1], uz=u[2], gaminv=g, w=w, q=q, dt=dt)
    for order in (1, 2, 3):
        for algo in (0, 3):
            j = [np.zeros(SHAPE, order="F") for _ in range(3)]
            ref.depose_jxjyjz(*j, xp, yp, zp, *u, g, w, q, XMIN, YMIN, ZMIN, dt, DX, DY, DZ,
                              NX, NY, NZ, NG, NG, NG, order, algo=algo)
            d[f"o{order}_a{algo}"] = np.stack(j)
    np.savez_compressed(os.path.join(HERE, "deposit.npz"), **d)


def gen_rho(ref, rng):
    """pxr_depose_rho_n and depose_rho_scalar_n_n_n, orders 1..3; the
    positions include nodes, half nodes and guard cells."""
    n = 256
    xp, yp, zp = positions(rng, n, lo=-1.0, hi=None)
    w = rng.uniform(0.5, 1.5, n)
    q = 0.7
    d = dict(xp=xp, yp=yp, zp=zp, w=w, q=q)
    for order in (1, 2, 3):
        for scalar in (False, True):
            rho = np.zeros(SHAPE, order="F")
            ref.depose_rho(rho, xp, yp, zp, w, q, XMIN, YMIN, ZMIN, DX, DY, DZ, NX, NY, NZ, NG,
                           NG, NG, order, scalar=scalar)
            d[f"o{order}_{'scalar' if scalar else 'n'}"] = rho
    np.savez_compressed(os.path.join(HERE, "rho.npz"), **d)


# one shape order per axis: the reference's generic routines
GENERIC_GATHER = [  # (nox, noy, noz), l_lower_order_in_v, l_nodal, l4symtry
    ((1, 2, 3), 1, 0, 0), ((3, 1, 2), 1, 0, 0), ((2, 3, 1), 0, 0, 0), ((3, 3, 3), 1, 0, 0),
    ((3, 3, 3), 0, 0, 0), ((1, 1, 1), 1, 0, 0), ((2, 2, 3), 1, 1, 0), ((3, 2, 1), 0, 1, 0),
    ((2, 3, 3), 1, 0, 1),
]
GENERIC_DEPOSIT = [(0, 1, 2), (1, 2, 3), (3, 2, 1), (3, 3, 3), (1, 1, 1), (0, 0, 0), (2, 0, 3)]
GENERIC_RHO = [  # orders, l_particles_weight, l4symtry
    ((0, 1, 2), 1, 0), ((3, 2, 1), 1, 0), ((0, 0, 0), 1, 0), ((2, 3, 1), 0, 0), ((3, 1, 3), 1, 1),
]


def gen_generic(ref, rng):
    """pxr_gete3d_n/pxr_getb3d_n_energy_conserving (field_gathering_on_3d.F90),
    pxr_depose_jxjyjz_esirkepov_n (esirkepov_3d.F90:3480) and
    pxr_depose_rho_n (charge_deposition_3d.F90:2081) with a shape order per
    axis, including order 0 where the routine has it, the nodal grid and
    the 4-fold symmetry flag."""
    n = 128
    xp, yp, zp = positions(rng, n)
    F = [grid(rng) for _ in range(6)]
    d = dict(g_xp=xp, g_yp=yp, g_zp=zp, exg=F[0], eyg=F[1], ezg=F[2], bxg=F[3], byg=F[4],
             bzg=F[5],
             gather_cases=np.array([[*o, lo, nod, l4] for o, lo, nod, l4 in GENERIC_GATHER]),
             esirk_cases=np.array(GENERIC_DEPOSIT),
             rho_cases=np.array([[*o, lpw, l4] for o, lpw, l4 in GENERIC_RHO]))
    for c, (o, lo, nod, l4) in enumerate(GENERIC_GATHER):
        e = [np.zeros(n) for _ in range(6)]
        ref.geteb3d_n(xp, yp, zp, *e, XMIN, YMIN, ZMIN, DX, DY, DZ, NX, NY, NZ, NG, NG, NG, *o,
                      *F, l4symtry=l4, l_lower_order_in_v=lo, l_nodal=nod)
        d[f"gather{c}"] = np.stack(e)
    dt = 0.5
    xp, yp, zp = positions(rng, n, lo=0.0)
    u = [rng.uniform(-2.0, 2.0, n) for _ in range(3)]
    for a in u:
        a[:8] = 0.0
    g = 1.0 / np.sqrt(1.0 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    w = rng.uniform(0.5, 1.5, n)
    d.update(d_xp=xp, d_yp=yp, d_zp=zp, ux=u[0], uy=u[1], uz=u[2], gaminv=g, w=w, q=-1.3,
             dt=dt)
    for c, o in enumerate(GENERIC_DEPOSIT):
        j = [np.zeros(SHAPE, order="F") for _ in range(3)]
        ref.depose_esirkepov_n(*j, xp, yp, zp, *u, g, w, -1.3, XMIN, YMIN, ZMIN, dt, DX, DY, DZ,
                               NX, NY, NZ, NG, NG, NG, *o)
        d[f"esirk{c}"] = np.stack(j)
    xp, yp, zp = positions(rng, n, lo=-1.0)
    d.update(r_xp=xp, r_yp=yp, r_zp=zp)
    for c, (o, lpw, l4) in enumerate(GENERIC_RHO):
        rho = np.zeros(SHAPE, order="F")
        ref.depose_rho_n(rho, xp, yp, zp, w, 0.7, XMIN, YMIN, ZMIN, DX, DY, DZ, NX, NY, NZ, NG,
                         NG, NG, *o, l_particles_weight=lpw, l4symtry=l4)
        d[f"rho{c}"] = rho
    np.savez_compressed(os.path.join(HERE, "generic.npz"), **d)


def gen_fdtd(ref, rng):
    F = [grid(rng) for _ in range(9)]
    d = {f"in{i}": f for i, f in enumerate(F)}
    lo, hi = (-1, -1, -1), (NX + 1, NY + 1, NZ + 1)
    mudt, sx, sy, sz = 0.45, 0.45 / DX, 0.45 / DY, 0.45 / DZ
    A = [f.copy(order="F") for f in F]
    ref.push_em3d_evec(lo, hi, *A, mudt, sx, sy, sz, NX, NY, NZ, NG, NG, NG)
    d["yee_e"] = np.stack(A[:3])
    A = [f.copy(order="F") for f in F]
    ref.push_em3d_bvec(lo, hi, *A[:6], 0.5 * sx, 0.5 * sy, 0.5 * sz, NX, NY, NZ, NG, NG, NG)
    d["yee_b"] = np.stack(A[3:6])
    for no in (2, 4, 6):
        c = ref.fd_weights(no)
        d[f"fdw{no}"] = c
        A = [f.copy(order="F") for f in F]
        ref.push_em3d_evec_norder(*A, mudt, sx * c, sy * c, sz * c, NX, NY, NZ, no, no, no,
                                  NG, NG, NG, 1, 1, 1)
        d[f"no{no}_e"] = np.stack(A[:3])
        A = [f.copy(order="F") for f in F]
        ref.push_em3d_bvec_norder(*A[:6], 0.5 * sx * c, 0.5 * sy * c, 0.5 * sz * c, NX, NY, NZ,
                                  no, no, no, NG, NG, NG, 1, 1, 1)
        d[f"no{no}_b"] = np.stack(A[3:6])
    d["fdw8"] = ref.fd_weights(8)
    d["fdw16"] = ref.fd_weights(16)
    np.savez_compressed(os.path.join(HERE, "fdtd.npz"), **d)


def gen_sort(ref, rng):
    # one tile of 6x5x4 nodes sorted on a unit grid shifted by half a cell
    n = 700
    lo = (-0.5, -0.5, -0.5)
    hi = (6.5, 5.5, 4.5)
    P = [rng.uniform(lo[a], hi[a], n) for a in range(3)] + [rng.standard_normal(n) for _ in range(5)]
    # ties and keys on bin edges
    P[0][:50] = np.round(P[0][:50]) - 0.5
    P[1][50:100] = 2.0
    P[2][100:150] = 1.5
    P[7] = np.arange(n, dtype=np.float64)   # pid column 1 = original index
    d = {f"in{i}": p.copy() for i, p in enumerate(P)}
    args = (lo[0], lo[1], lo[2], hi[0] + 1.0, hi[1] + 1.0, hi[2] + 1.0, 1.0, 1.0, 1.0)
    ref.particle_bin_sorting(*P, *args)
    d["out"] = np.stack(P)
    d["box"] = np.array(args)
    # a second case with a coarser sorting cell (2 cells per bin in x)
    P2 = [d[f"in{i}"].copy() for i in range(8)]
    args2 = (lo[0], lo[1], lo[2], hi[0] + 2.0, hi[1] + 1.0, hi[2] + 1.0, 2.0, 1.0, 1.0)
    ref.particle_bin_sorting(*P2, *args2)
    d["out2"] = np.stack(P2)
    d["box2"] = np.array(args2)
    np.savez_compressed(os.path.join(HERE, "sort.npz"), **d)


LASER_CASES = [
    # window, temporal_order, polangle, time, q_z (complex beam parameter)
    (0, 2, 0.0, 30.0, 1j * 64.0 ** 2 * np.pi / 32.0),
    (0, 2, 0.125, 61.5, 1j * 64.0 ** 2 * np.pi / 32.0),
    (0, 4, 0.3, 12.0, 15.0 + 1j * 400.0),
    (0, 0, 0.0, 5.0, 1j * 100.0),
    (1, 2, 0.0, 20.0, 1j * 64.0 ** 2 * np.pi / 32.0),
    (1, 2, 0.25, 47.0, -8.0 + 1j * 250.0),
    (1, 2, 0.0, 85.0, 1j * 100.0),  # after 2 t_peak: zero
]


def gen_laser(ref, rng):
    """gaussian_profile and hanning_profile of the laser antenna
    (laser_pusher_manager_3d.F90:368-478) at antenna-frame points, including
    the axis, for several windows, time orders, polarisation angles and
    complex beam parameters."""
    n = 96
    xx = rng.standard_normal(n) * 40.0
    yy = rng.standard_normal(n) * 40.0
    xx[:4] = [0.0, 1.0, -3.5, 0.0]
    yy[:4] = [0.0, 0.0, 2.0, -64.0]
    pv1 = np.array([0.6, 0.0, 0.8])
    pv2 = np.array([0.0, 1.0, 0.0])
    d = dict(xx=xx, yy=yy, pv1=pv1, pv2=pv2, emax1=2.0, emax2=0.7, k0=2 * np.pi / 32.0,
             tau=25.0, t_peak=40.0)
    for c, (win, tord, pol, t, qz) in enumerate(LASER_CASES):
        d[f"case{c}"] = np.array([win, tord, pol, t, qz.real, qz.imag])
        d[f"amp{c}"] = ref.laser_profile(win, xx, yy, d["emax1"], d["emax2"], pv1, pv2, d["k0"],
                                         qz, d["tau"], t, d["t_peak"], tord, pol)
    np.savez_compressed(os.path.join(HERE, "laser.npz"), **d)


def gen_div(ref, rng):
    """calc_field_div and calc_field_divB (diags.F90:129-197) of random
    fields on a small grid with non-cubic cells."""
    nx, ny, nz, g = 7, 5, 6, 2
    d = dict(nx=nx, ny=ny, nz=nz, ng=g, dx=0.9, dy=1.1, dz=1.3)
    shp = (nx + 1 + 2 * g, ny + 1 + 2 * g, nz + 1 + 2 * g)
    for k in ("fx", "fy", "fz"):
        d[k] = np.asfortranarray(rng.standard_normal(shp))
    for name, fn in (("dive", ref.calc_field_div), ("divb", ref.calc_field_divb)):
        out = np.asfortranarray(np.full(shp, 7.0))  # nodes outside the loops keep their value
        fn(out, d["fx"], d["fy"], d["fz"], nx, ny, nz, g, g, g, d["dx"], d["dy"], d["dz"])
        d[name] = out
    np.savez_compressed(os.path.join(HERE, "div.npz"), **d)


def main():
    ref = load_reference()
    if ref is None:
        sys.exit("oracle/_ref/libpxrref.so is missing: run bash oracle/build_ref.sh first")
    which = set(sys.argv[1:]) or {"gather", "push", "deposit", "fdtd", "sort", "rho", "laser",
                                  "div", "generic"}
    rng = np.random.default_rng(20241202)  # one stream for the first five families
    for name, fn in (("gather", gen_gather), ("push", gen_push), ("deposit", gen_deposit),
                     ("fdtd", gen_fdtd), ("sort", gen_sort)):
        if name in which:
            fn(ref, rng)
    if "rho" in which:
        gen_rho(ref, np.random.default_rng(20241203))
    if "laser" in which:
        gen_laser(ref, np.random.default_rng(20241204))
    if "div" in which:
        gen_div(ref, np.random.default_rng(20241205))
    if "generic" in which:
        gen_generic(ref, np.random.default_rng(20241206))
    meta = dict(nx=NX, ny=NY, nz=NZ, ng=NG, dx=DX, dy=DY, dz=DZ, xmin=XMIN, ymin=YMIN, zmin=ZMIN)
    np.savez(os.path.join(HERE, "grid.npz"), **meta)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
