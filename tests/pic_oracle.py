"""CPU oracle of one PIC step on one periodic rank (test infrastructure).

Composes the oracle's restated leaf routines in the reference's step order
(submain.F90:99-199) on global arrays: gather at the current positions, push,
position push, deposition of the pushed particles, periodic fold of the
current guards (current_bcs), B half step, E step, B half step with their
guard exchanges, and the periodic wrap of the positions (particle_bcs).
"""
from __future__ import annotations

import math

import numpy as np


class OracleState:
    def __init__(self, orc, nx, ny, nz, ng, dx, dy, dz, xmin, ymin, zmin, dt, order, lo,
                 pusher, depo, norder=2, coefs=None):
        self.orc = orc
        self.nx, self.ny, self.nz, self.ng = nx, ny, nz, ng
        self.dx, self.dy, self.dz = dx, dy, dz
        self.xmin, self.ymin, self.zmin = xmin, ymin, zmin
        self.dt = dt
        self.order, self.lo, self.pusher, self.depo = order, lo, pusher, depo
        self.norder = norder
        self.coefs = coefs
        shape = (nx + 1 + 2 * ng, ny + 1 + 2 * ng, nz + 1 + 2 * ng)
        self.f = {k: np.zeros(shape, order="F") for k in
                  ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")}
        self.species = []
        self.antennas = []
        self.it = 0
        # particle boundaries (x_min, x_max, y_min, y_max, z_min, z_max):
        # 0 periodic, 1 absorbing, 2 reflecting, 3 reinjecting; nt: tiles per
        # axis (reinjection keeps a particle inside the tile it was in)
        self.pbound = (0,) * 6
        self.partcom = 0
        self.nt = (1, 1, 1)

    def add_antenna(self, a):
        """a: dict of x, y, z, pid ((n, 3) Fortran order: weight, antenna-frame
        x and y) and the laser parameters of orc.laserp_pusher."""
        n = len(a["x"])
        self.antennas.append(dict(a, ux=np.zeros(n), uy=np.zeros(n), uz=np.zeros(n),
                                  gaminv=np.ones(n)))

    def add_species(self, q, m, x, y, z, ux, uy, uz, w, lfreeze=False, ldodepos=True):
        """lfreeze: not pushed (particle_pusher_manager_3d.F90:523); ldodepos:
        deposits current (current_deposition_manager_3d.F90:595)."""
        g = 1.0 / np.sqrt(1.0 + ux ** 2 + uy ** 2 + uz ** 2)
        self.species.append(dict(q=q, m=m, x=x.copy(), y=y.copy(), z=z.copy(), ux=ux.copy(),
                                 uy=uy.copy(), uz=uz.copy(), gaminv=g, w=w.copy(),
                                 lfreeze=lfreeze, ldodepos=ldodepos))

    def geom(self):
        return (self.xmin, self.ymin, self.zmin)

    def step(self):
        self.gather_push_deposit(zero_j=True)
        self.fold_currents_and_advance_fields()
        self.particle_bcs()
        self.it += 1

    def step_reference_order(self):
        """submain.F90:99-199 in the reference's order: gather + push, laser
        push, particle boundaries and sort, then the deposition from the new
        (wrapped) positions."""
        self.gather_push()
        self.push_antennas()
        self.particle_bcs()
        self.deposit(zero_j=True)
        self.deposit_antennas()
        self.fold_currents_and_advance_fields()
        self.it += 1

    def gather_push_deposit(self, zero_j=True):
        """Gather + push of every species, then deposition of the pushed
        particles into J (zeroed first when zero_j); then the antennas."""
        self.gather_push()
        self.deposit(zero_j)
        self.push_antennas()
        self.deposit_antennas()

    def gather_push(self):
        o, f = self.orc, self.f
        nx, ny, nz, ng = self.nx, self.ny, self.nz, self.ng
        d = (self.dx, self.dy, self.dz)
        dt = self.dt
        for s in self.species:
            n = len(s["x"])
            s["old"] = (s["x"].copy(), s["y"].copy(), s["z"].copy())  # for particle_bcs
            if s.get("lfreeze"):
                continue
            e = [np.zeros(n) for _ in range(6)]
            bad = o.geteb3d_energy_conserving(s["x"], s["y"], s["z"], *e, *self.geom(), *d, nx,
                                              ny, nz, ng, ng, ng, self.order, self.order,
                                              self.order, f["ex"], f["ey"], f["ez"], f["bx"],
                                              f["by"], f["bz"], self.lo, False)
            assert bad == 0
            u = (s["ux"], s["uy"], s["uz"])
            if self.pusher == "vay":
                o.ebcancelpush3d(*u, s["gaminv"], *e, s["q"], s["m"], dt, 0)
            else:
                o.boris_push_u_3d(*u, s["gaminv"], *e, s["q"], s["m"], dt)
            o.pushxyz(s["x"], s["y"], s["z"], *u, s["gaminv"], dt)

    def _depose(self, s, w, q):
        o, f = self.orc, self.f
        d = (self.dx, self.dy, self.dz)
        args = (f["jx"], f["jy"], f["jz"], s["x"], s["y"], s["z"], s["ux"], s["uy"], s["uz"],
                s["gaminv"], w, q, *self.geom(), self.dt, *d, self.nx, self.ny, self.nz, self.ng,
                self.ng, self.ng, self.order)
        bad = o.depose_esirkepov(*args) if self.depo == "esirkepov" else \
            o.depose_direct(*args, False)
        assert bad == 0

    def deposit(self, zero_j=True):
        if zero_j:
            for k in ("jx", "jy", "jz"):
                self.f[k][...] = 0.0
        for s in self.species:
            if s.get("ldodepos", True):
                self._depose(s, s["w"], s["q"])

    def push_antennas(self):
        """push_laser_particles at it*dt - dt/2 (laser_pusher_manager_3d.F90:37)."""
        dt = self.dt
        for a in self.antennas:
            self.orc.laserp_pusher(a["window"], a["pid"], a["x"], a["y"], a["z"], a["ux"],
                                   a["uy"], a["uz"], a["gaminv"], dt, a["emax"], a["emax1"],
                                   a["emax2"], a["pv1"], a["pv2"], a["k0"], a["q_z"],
                                   a["charge"], a["tau"], self.it * dt - dt / 2.0, a["t_peak"],
                                   a["torder"], a["polangle"])

    def deposit_antennas(self):
        for a in self.antennas:
            self._depose(a, np.ascontiguousarray(a["pid"][:, 0]), a["charge"])

    def fold_currents_and_advance_fields(self):
        o, f = self.orc, self.f
        nx, ny, nz, ng = self.nx, self.ny, self.nz, self.ng
        for k in ("jx", "jy", "jz"):
            o.current_bc_periodic(f[k], nx, ny, nz, ng, ng, ng)
        self._bhalf()
        self._bcs(("bx", "by", "bz"))
        self._efull()
        self._bcs(("ex", "ey", "ez"))
        self._bhalf()
        self._bcs(("bx", "by", "bz"))

    def tile_bounds(self, a):
        """Membership [tmin, tmax) of every tile along axis a
        (set_tile_split_for_species, tiling.F90:94-271)."""
        n = (self.nx, self.ny, self.nz)[a]
        d = (self.dx, self.dy, self.dz)[a]
        x0 = self.geom()[a]
        nt = self.nt[a]
        n0 = (n + 1) // nt
        out = []
        for i in range(nt):
            gmin = x0 + i * n0 * d
            gmax = gmin + (n0 - 1) * d
            lo = x0 if i == 0 else gmin - d / 2.0
            hi = x0 + n * d if i == nt - 1 else gmax + d / 2.0
            out.append((lo, hi))
        return out, n0

    def _tile_of(self, p):
        """Tile (per axis) of an in-box position (particle_boundaries.F90:397)."""
        it = []
        for a in range(3):
            d = (self.dx, self.dy, self.dz)[a]
            _, n0 = self.tile_bounds(a)
            t = int(np.floor(((p[a] - self.geom()[a]) + d / 2.0) / (n0 * d)))
            it.append(min(max(t, 0), self.nt[a] - 1))
        return it

    def particle_bcs(self):
        """Domain boundaries of the particles that left the box, axis by axis:
        periodic x +/- L; absorbing: removed; reinjecting: x = 2 x_edge - x,
        u = 0, all coordinates clamped to [tmin + d/2, tmax - d/2] of the tile
        the particle was in, no further axis; reflecting: with partcom 1 or 2
        (particle_bcs_mpi_blocking, particle_boundaries.F90:1106-1300) x +/- dx
        and the normal momentum reversed, with partcom 0, the reference's
        default (particle_bcs_tiles_and_mpi_3d, :2280-2480, whose SELECT CASE
        has no reflecting branch), the periodic wrap.  Each side is tested on
        the coordinate as it was before the boundaries ran."""
        L = (self.nx * self.dx, self.ny * self.dy, self.nz * self.dz)
        D = (self.dx, self.dy, self.dz)
        x0 = self.geom()
        bounds = [self.tile_bounds(a)[0] for a in range(3)]
        for s in self.species:
            X = [s["x"], s["y"], s["z"]]
            U = [s["ux"], s["uy"], s["uz"]]
            out = np.zeros(len(X[0]), dtype=bool)
            for a in range(3):
                out |= (X[a] < x0[a]) | (X[a] >= x0[a] + L[a])
            keep = np.ones(len(X[0]), dtype=bool)
            for i in np.nonzero(out)[0]:
                for a in range(3):
                    v = X[a][i]
                    done = False
                    for side in (0, 1):
                        if not ((v < x0[a]) if side == 0 else (v >= x0[a] + L[a])):
                            continue
                        b = self.pbound[2 * a + side]
                        if b == 2 and self.partcom == 0:
                            b = 0
                        if b == 1:
                            keep[i] = False
                            done = True
                        elif b == 2:
                            X[a][i] = v + D[a] if side == 0 else v - D[a]
                            U[a][i] = -U[a][i]
                        elif b == 3:
                            edge = x0[a] if side == 0 else x0[a] + L[a]
                            X[a][i] = 2.0 * edge - X[a][i]
                            for e in range(3):
                                U[e][i] = 0.0
                            it = self._tile_of([s["old"][e][i] for e in range(3)])
                            for e in range(3):
                                lo, hi = bounds[e][it[e]]
                                X[e][i] = max(X[e][i], lo + 0.5 * D[e])
                                X[e][i] = min(X[e][i], hi - 0.5 * D[e])
                            done = True
                        else:
                            X[a][i] = v + L[a] if side == 0 else v - L[a]
                    if done:
                        break
            if not keep.all():
                for k in ("x", "y", "z", "ux", "uy", "uz", "gaminv", "w"):
                    s[k] = s[k][keep]

    def _bcs(self, keys):
        for k in keys:
            self.orc.field_bc_periodic(self.f[k], self.nx, self.ny, self.nz, self.ng, self.ng,
                                       self.ng)

    def _bhalf(self):
        f, h = self.f, 0.5 * self.dt
        if self.norder == 2:
            lo, hi = (0, 0, 0), (self.nx, self.ny, self.nz)
            self.orc.push_em3d_bvec(lo, hi, f["ex"], f["ey"], f["ez"], f["bx"], f["by"], f["bz"],
                                    h / self.dx, h / self.dy, h / self.dz, self.nx, self.ny,
                                    self.nz, self.ng, self.ng, self.ng)
        else:
            c = self.coefs
            self.orc.push_em3d_bvec_norder(f["ex"], f["ey"], f["ez"], f["bx"], f["by"], f["bz"],
                                           h / self.dx * c, h / self.dy * c, h / self.dz * c,
                                           self.nx, self.ny, self.nz, self.norder, self.norder,
                                           self.norder, self.ng, self.ng, self.ng, 0, 0, 0)

    def _efull(self):
        f, dt = self.f, self.dt
        if self.norder == 2:
            lo, hi = (0, 0, 0), (self.nx, self.ny, self.nz)
            self.orc.push_em3d_evec(lo, hi, f["ex"], f["ey"], f["ez"], f["bx"], f["by"], f["bz"],
                                    f["jx"], f["jy"], f["jz"], dt, dt / self.dx, dt / self.dy,
                                    dt / self.dz, self.nx, self.ny, self.nz, self.ng, self.ng,
                                    self.ng)
        else:
            c = self.coefs
            self.orc.push_em3d_evec_norder(f["ex"], f["ey"], f["ez"], f["bx"], f["by"], f["bz"],
                                           f["jx"], f["jy"], f["jz"], dt, dt / self.dx * c,
                                           dt / self.dy * c, dt / self.dz * c, self.nx, self.ny,
                                           self.nz, self.norder, self.norder, self.norder,
                                           self.ng, self.ng, self.ng, 0, 0, 0)


def thermal_plasma(rng, nx, ny, nz, ppc, dx, dy, dz, xmin, ymin, zmin, uth, id_offset=0,
                   id_scale=1e-9):
    """Uniform random positions, Maxwellian momenta (normal per component);
    weights 1 + id * id_scale carry a unique id so that particles can be
    matched after re-ordering (id_scale = 2**-15 keeps them exact and
    distinct in float32 for up to 32768 particles)."""
    n = nx * ny * nz * ppc
    x = xmin + rng.uniform(0, nx, n) * dx
    y = ymin + rng.uniform(0, ny, n) * dy
    z = zmin + rng.uniform(0, nz, n) * dz
    ux, uy, uz = (rng.standard_normal(n) * uth for _ in range(3))
    w = 1.0 + (np.arange(n) + id_offset) * id_scale
    return x, y, z, ux, uy, uz, w


def default_dt(dx, dy, dz, cfl=0.95):
    return cfl / math.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2)
