"""Device-resident PIC state and the reference's per-step sequence on the GPU.

This is the manager layer of PICSAR (the zero-argument routines of
``picsar_python.picsarpy.picsar`` that act on module state: push_bfield,
push_efield, efield_bcs, bfield_bcs, current_bcs, particle_bcs,
particle_sorting_sub, field_gathering_plus_particle_pusher,
pxrdepose_currents_on_grid_jxjyjz) for one periodic rank.  Fields live in
Fortran-ordered float64 device grids with guards; each species lives in SoA
device arrays packed by tile (tile geometry of set_tile_split_for_species,
tiling.F90:94-271).

The step follows submain.F90:99-199:

    gather + push          field_gathering_plus_particle_pusher
    particle boundaries    particle_bcs            (periodic wrap, re-binning)
    sorting                particle_sorting_sub
    current deposition     pxrdepose_currents_on_grid_jxjyjz
    current guards         current_bcs
    B half step, E step, B half step with their guard exchanges

With periodic particle faces, the gather, push and deposition of a species
are one kernel that deposits the pushed particles before they are wrapped;
the periodic fold of the current guards (current_bcs) makes that identical to
depositing after the wrap, up to rounding.  Re-binning and sorting are one
counting sort that then runs after the field update, so the particles enter
the next step's gather wrapped, binned and cell-sorted, as in the reference.
With an absorbing, reinjecting or (under partcom 1, 2) reflecting face, the
step keeps the reference's order: the same kernel gathers and pushes, the
re-binning applies the boundaries and sorts, and a deposition pass of that
kernel then deposits from the particles as the boundaries left them.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from .devmem import DeviceArray, DeviceGrid
from .picsarpy import picsar as _pxr


class PxrGrid(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("nx", "ny", "nz", "ngx", "ngy", "ngz", "ntx", "nty",
                                         "ntz")] + \
               [(n, C.c_double) for n in ("xmin", "ymin", "zmin", "dx", "dy", "dz",
                                          "sorting_dx", "sorting_dy", "sorting_dz",
                                          "sorting_shiftx", "sorting_shifty",
                                          "sorting_shiftz")] + \
               [("pbound", C.c_int64 * 6), ("partcom", C.c_int64)]


MAX_PID_EXTRA = 4  # PXRHIP_MAX_PID_EXTRA


class PxrSpecies(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "z", "ux", "uy", "uz", "gaminv", "w",
                                          "tile_off", "tile_np", "bin_start")] + \
               [("capacity", C.c_int64), ("q", C.c_double), ("m", C.c_double),
                ("precision", C.c_int64), ("npid_extra", C.c_int64),
                ("pid_extra", C.c_void_p * MAX_PID_EXTRA)]


PUSHERS = {"boris": 0, "vay": 1}
# pbound values of the reference's particle boundaries (particle_boundaries.F90)
PBOUND_CODES = {"periodic": 0, "absorbing": 1, "reflecting": 2, "reinjecting": 3}
PBOUNDS_DEFAULT = ("periodic",) * 6
DEPOS = {"esirkepov": 0, "direct": 3}
COMPONENTS = ("x", "y", "z", "ux", "uy", "uz", "gaminv", "w")
PRECISIONS = {"fp64": 0, "mixed": 1}


def component_dtype(name: str, precision: str):
    """Storage type of a particle component: positions and pid columns are
    always float64; in mixed precision momenta, gaminv and weights are
    float32."""
    if precision == "mixed" and name not in ("x", "y", "z") and not name.startswith("pid"):
        return np.float32
    return np.float64


@dataclass
class _Buffers:
    arrays: Dict[str, object]
    tile_off: object
    tile_np: object
    bin_start: object


@dataclass
class Species:
    """One particle species: charge, mass and two SoA buffers (the active one
    and the target of the next re-binning)."""
    name: str
    q: float
    m: float
    capacity: int
    npart: int
    bufs: List[_Buffers] = field(default_factory=list)
    cur: int = 0
    lfreeze: bool = False   # particle_species%lfreeze: not pushed
    ldodepos: bool = True   # particle_species%ldodepos: deposits current
    precision: str = "fp64"
    npid_extra: int = 0     # pid columns beyond the weight, arrays "pid2", "pid3", ...

    @property
    def active(self) -> _Buffers:
        return self.bufs[self.cur]

    def arrays(self):
        """Host copies of the active buffer's particle arrays (npart values)."""
        return {k: v.numpy(self.npart) for k, v in self.active.arrays.items()}

    def struct(self, which: int) -> PxrSpecies:
        b = self.bufs[which]
        s = PxrSpecies()
        for k in COMPONENTS:
            setattr(s, k, b.arrays[k].ptr)
        s.npid_extra = self.npid_extra
        for c in range(self.npid_extra):
            s.pid_extra[c] = b.arrays[f"pid{c + 2}"].ptr
        s.tile_off = b.tile_off.ptr
        s.tile_np = b.tile_np.ptr
        s.bin_start = b.bin_start.ptr
        s.capacity = self.capacity
        s.q = float(self.q)
        s.m = float(self.m)
        s.precision = PRECISIONS[self.precision]
        return s


class Simulation:
    """One periodic rank: Yee fields with guards and tiled particle species.

    Units are the reference's -DLIBRARY normalised units (clight = mu0 =
    eps0 = 1); momenta are gamma*v.  ``dt`` defaults to cfl times the
    stability limit of the FDTD stencil: 1/sqrt(1/dx^2 + 1/dy^2 + 1/dz^2)
    for Yee (submain.F90:581-590 with c = 1), divided by the sum of the
    absolute stencil weights for norder > 2 (7/6 at order 4), since the
    reference's formula, stable for Yee, is not for the longer stencils at
    the same coefficient (its default dtcoef of 0.7, control_file.F90:179,
    stays below the order-4 limit of 6/7).
    """

    def __init__(self, nx, ny, nz, dx=1.0, dy=1.0, dz=1.0, xmin=0.0, ymin=0.0, zmin=0.0,
                 ntx=1, nty=1, ntz=1, order=1, l_lower_order_in_v=True, pusher="boris",
                 depo="esirkepov", norder=2, dt=None, cfl=0.95, nguard=None, sorting_dx=1.0,
                 sorting_dy=1.0, sorting_dz=1.0, chunk=0, precision="fp64", pbound=None,
                 partcom=0):
        _lib.call("pxrhip_init", -1)
        if precision not in PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.xmin, self.ymin, self.zmin = float(xmin), float(ymin), float(zmin)
        self.order = int(order)
        if self.order not in (1, 2, 3):
            raise ValueError("order must be 1, 2 or 3")
        self.l_lower_order_in_v = bool(l_lower_order_in_v)
        self.pusher = PUSHERS[pusher]
        self.depo = DEPOS[depo]
        self.norder = int(norder)
        if self.norder < 2 or self.norder % 2:
            raise ValueError("norder must be even and >= 2")
        # nxguards = max(nox, 2) (control_file.F90:111-114); the stencil of an
        # order-n FDTD needs n/2 guard nodes
        ng = max(self.order, 2, self.norder // 2) if nguard is None else int(nguard)
        self.ng = ng
        if dt is None:
            wsum = float(np.abs(_fd_weights(self.norder)).sum()) if self.norder > 2 else 1.0
            dt = cfl / (wsum * math.sqrt(1 / self.dx ** 2 + 1 / self.dy ** 2 + 1 / self.dz ** 2))
        self.dt = float(dt)
        self.chunk = int(chunk)
        self.it = 0
        for n, t in ((self.nx, ntx), (self.ny, nty), (self.nz, ntz)):
            if (n + 1) // t < 4:
                raise ValueError("tiles must be at least 4 nodes wide (tiling.F90:110-126)")
        self.grid = PxrGrid(self.nx, self.ny, self.nz, ng, ng, ng, int(ntx), int(nty), int(ntz),
                            self.xmin, self.ymin, self.zmin, self.dx, self.dy, self.dz,
                            sorting_dx * self.dx, sorting_dy * self.dy, sorting_dz * self.dz,
                            0.0, 0.0, 0.0)
        # particle boundaries at x_min, x_max, y_min, y_max, z_min, z_max:
        # "periodic", "absorbing", "reflecting", "reinjecting" (or 0..3);
        # the fields stay periodic
        pb = PBOUNDS_DEFAULT if pbound is None else tuple(pbound)
        if len(pb) != 6:
            raise ValueError("pbound needs six entries (x_min, x_max, y_min, ..., z_max)")
        codes = [PBOUND_CODES[b] if isinstance(b, str) else int(b) for b in pb]
        for c in codes:
            if c not in PBOUND_CODES.values():
                raise ValueError(f"unknown particle boundary {c}")
        self.grid.pbound[:] = codes
        # particle_bcs's exchange (particle_boundaries.F90:83-217): 0, the
        # reference's default, has no reflecting case (a reflecting face is
        # periodic there); 1 and 2 reflect (particle_bcs_mpi_blocking)
        if int(partcom) not in (0, 1, 2):
            raise ValueError("partcom must be 0, 1 or 2")
        self.grid.partcom = int(partcom)
        self.partcom = int(partcom)
        self.absorbing = PBOUND_CODES["absorbing"] in codes
        eff = [0 if (self.partcom == 0 and c == PBOUND_CODES["reflecting"]) else c for c in codes]
        # With periodic faces only, depositing the pushed particles before
        # their wrap gives the current of the reference's order once
        # current_bcs has folded the guards, so the step fuses gather, push
        # and deposition.  Any other face removes, moves or stops particles
        # before the reference deposits them: the step then runs in
        # submain.F90's order (push, particle_bcs, deposition).
        self.split_step = any(c != 0 for c in eff)
        self.ntiles = int(ntx) * int(nty) * int(ntz)
        nb = C.c_int64(0)
        _lib.call("pxrhip_tiles_nbins", C.byref(self.grid), C.byref(nb))
        self.nbins = nb.value
        self.fields = {k: DeviceGrid(self.nx, self.ny, self.nz, ng)
                       for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")}
        if self.norder > 2:
            c = _fd_weights(self.norder)
            self.coefs = c
        else:
            self.coefs = None
        self.species: List[Species] = []
        self.antennas = []
        self.diag_fields = {}  # rho, dive, divb: allocated on first use

    # ------------------------------------------------------------------
    # particles
    # ------------------------------------------------------------------
    def add_species(self, name, q, m, x, y, z, ux, uy, uz, w, gaminv=None, capacity=None,
                    lfreeze=False, ldodepos=True, pid=None) -> Species:
        """Add a species from host arrays.  gaminv defaults to
        1/sqrt(1+u^2) (pxr_set_gamma).  pid: the reference's pid columns
        beyond the weight, an (npart, k) array (k <= 4; particle ids, old
        positions ...), carried with their particles."""
        src = {k: np.asarray(v, dtype=np.float64).ravel()
               for k, v in zip(("x", "y", "z", "ux", "uy", "uz", "w"), (x, y, z, ux, uy, uz, w))}
        if gaminv is not None:
            src["gaminv"] = np.asarray(gaminv, dtype=np.float64).ravel()
        npid_extra = 0
        if pid is not None:
            pid = np.asarray(pid, dtype=np.float64).reshape(len(src["x"]), -1)
            src["pid"] = pid
            npid_extra = pid.shape[1]
        return self.add_species_chunks(name, q, m, len(src["x"]), [src], capacity=capacity,
                                       lfreeze=lfreeze, ldodepos=ldodepos, npid_extra=npid_extra)

    def add_species_chunks(self, name, q, m, npart, chunks, capacity=None, lfreeze=False,
                           ldodepos=True, npid_extra=0) -> Species:
        """Add a species of ``npart`` particles given as an iterable of dicts
        of host arrays (x, y, z, ux, uy, uz, w and optionally gaminv and, with
        npid_extra > 0, pid: an (n, npid_extra) array of the pid columns beyond
        the weight), each holding the next run of particles; large species
        can so be built slab by slab without holding every array in host
        memory at once."""
        n = int(npart)
        cap = int(capacity) if capacity is not None else max(n, 1)
        if cap < n:
            raise ValueError("capacity < number of particles")
        npid_extra = int(npid_extra)
        if not 0 <= npid_extra <= MAX_PID_EXTRA:
            raise ValueError(f"at most {MAX_PID_EXTRA} pid columns beyond the weight")
        sp = Species(name=name, q=float(q), m=float(m), capacity=cap, npart=n,
                     lfreeze=lfreeze, ldodepos=ldodepos, precision=self.precision,
                     npid_extra=npid_extra)
        comps = COMPONENTS + tuple(f"pid{c + 2}" for c in range(npid_extra))
        for _ in range(2):
            arrays = {k: DeviceArray(cap, component_dtype(k, self.precision)) for k in comps}
            b = _Buffers(arrays=arrays, tile_off=DeviceArray(self.ntiles + 1, np.int64),
                         tile_np=DeviceArray(self.ntiles, np.int64),
                         bin_start=DeviceArray(self.nbins + 1, np.int64))
            b.tile_off.fill(0)
            b.tile_np.fill(0)
            b.bin_start.fill(0)
            sp.bufs.append(b)
        # stage into buffer 1 as one unsorted run, then bin into buffer 0
        done = 0
        for ch in chunks:
            src = {k: np.ascontiguousarray(np.asarray(ch[k], dtype=np.float64).ravel())
                   for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
            if ch.get("gaminv") is None:
                src["gaminv"] = 1.0 / np.sqrt(1.0 + src["ux"] ** 2 + src["uy"] ** 2 +
                                              src["uz"] ** 2)
            else:
                src["gaminv"] = np.ascontiguousarray(np.asarray(ch["gaminv"],
                                                                dtype=np.float64).ravel())
            k0 = len(src["x"])
            if npid_extra:
                pid = np.asarray(ch["pid"], dtype=np.float64).reshape(k0, npid_extra)
                for c in range(npid_extra):
                    src[f"pid{c + 2}"] = np.ascontiguousarray(pid[:, c])
            for k, v in src.items():
                if v.size != k0:
                    raise ValueError(f"{k}: {v.size} values, expected {k0}")
            if done + k0 > n:
                raise ValueError(f"the chunks hold more than {n} particles")
            for k in comps:
                sp.bufs[1].arrays[k].upload(src[k], offset=done)
            done += k0
        if done != n:
            raise ValueError(f"the chunks hold {done} particles, expected {n}")
        off = np.zeros(self.ntiles + 1, dtype=np.int64)
        off[-1] = n
        sp.bufs[1].tile_off.upload(off)
        sp.cur = 1
        self.species.append(sp)
        self._rebin(sp)
        self.check()
        return sp

    def add_laser_antenna(self, **kw) -> "LaserAntenna":
        """A laser antenna (picsar_amd.laser.LaserAntenna): its particles are
        pushed by the laser profile and deposit current every step."""
        from .laser import LaserAntenna
        a = LaserAntenna(self, **kw)
        self.antennas.append(a)
        return a

    def laser_push_deposit(self):
        """push_laser_particles (laser_pusher_manager_3d.F90:15) and the
        antennas' current deposition."""
        self.push_antennas()
        self.deposit_antennas()

    def push_antennas(self):
        """push_laser_particles (laser_pusher_manager_3d.F90:15)."""
        for a in self.antennas:
            a.push(self)

    def deposit_antennas(self):
        """The antennas' current, added into J (the reference deposits them as
        species, tiling.F90:1697)."""
        for a in self.antennas:
            a.deposit(self)

    def _rebin(self, sp: Species):
        s_src = sp.struct(sp.cur)
        s_dst = sp.struct(1 - sp.cur)
        _lib.call("pxrhip_tiles_rebin_sort", C.byref(self.grid), C.byref(s_src), C.byref(s_dst))
        sp.cur = 1 - sp.cur
        if self.absorbing:  # particles may have been removed (this synchronises)
            sp.npart = int(sp.active.tile_off.numpy()[-1])

    # ------------------------------------------------------------------
    def check(self):
        """Raise if any kernel of the tiled path skipped a particle."""
        _lib.call("pxrhip_tiles_check")

    # ------------------------------------------------------------------
    # the reference's managers
    # ------------------------------------------------------------------
    def gather_push_deposit(self):
        """field_gathering_plus_particle_pusher + pxrdepose_currents_on_grid_jxjyjz
        for every species (J is zeroed first, current_deposition_manager_3d.F90:247)
        in one kernel per species, then the antennas."""
        for k in ("jx", "jy", "jz"):
            self.fields[k].fill(0.0)
        for sp in self.species:
            self.gather_push_deposit_species(sp)
        self.laser_push_deposit()

    def _tiled_args(self, sp: Species):
        return C.byref(self.grid), C.byref(sp.struct(sp.cur))

    def gather_push_deposit_species(self, sp: Species):
        """One launch of the fused kernel for one species (J accumulates):
        gather and push unless the species is frozen (lfreeze skips the push,
        particle_pusher_manager_3d.F90:523), deposition unless it does not
        deposit (ldodepos, current_deposition_manager_3d.F90:595)."""
        if sp.npart == 0:
            return
        if sp.lfreeze and not sp.ldodepos:
            return
        f = self.fields
        if sp.lfreeze:
            self.depose_currents_species(sp)
        elif not sp.ldodepos:
            self.gather_push_species(sp)
        else:
            _lib.call("pxrhip_tiles_gather_push_deposit", *self._tiled_args(sp),
                      f["ex"].ptr, f["ey"].ptr, f["ez"].ptr, f["bx"].ptr, f["by"].ptr,
                      f["bz"].ptr, f["jx"].ptr, f["jy"].ptr, f["jz"].ptr, self.dt, self.order,
                      int(self.l_lower_order_in_v), self.pusher, self.depo, self.chunk)

    def gather_push_species(self, sp: Species):
        """Gather, push and position push of one binned species (no deposition)."""
        f = self.fields
        _lib.call("pxrhip_tiles_gather_push", *self._tiled_args(sp), f["ex"].ptr, f["ey"].ptr,
                  f["ez"].ptr, f["bx"].ptr, f["by"].ptr, f["bz"].ptr, self.dt, self.order,
                  int(self.l_lower_order_in_v), self.pusher, self.chunk)

    def depose_currents_species(self, sp: Species):
        """The current of one binned species from its particles as they stand,
        added into J."""
        f = self.fields
        _lib.call("pxrhip_tiles_depose_currents", *self._tiled_args(sp), f["jx"].ptr,
                  f["jy"].ptr, f["jz"].ptr, self.dt, self.order, self.depo, self.chunk)

    def gather_push(self):
        """field_gathering_plus_particle_pusher (particle_pusher_manager_3d.F90:
        431-672): every species that is not frozen."""
        for sp in self.species:
            if not sp.lfreeze and sp.npart:
                self.gather_push_species(sp)

    def depose_currents(self):
        """pxrdepose_currents_on_grid_jxjyjz (current_deposition_manager_3d.F90:
        224-505): J = 0, then the current of every species that deposits and
        of the antennas."""
        for k in ("jx", "jy", "jz"):
            self.fields[k].fill(0.0)
        for sp in self.species:
            if sp.ldodepos and sp.npart:
                self.depose_currents_species(sp)
        self.deposit_antennas()

    def particle_bcs_and_sorting(self):
        """particle_bcs (domain boundaries + tile exchange) and
        particle_sorting_sub; a frozen species does not move and keeps its
        binning."""
        for sp in self.species:
            if sp.npart and not sp.lfreeze:
                self._rebin(sp)

    def current_bcs(self):
        n, g = (self.nx, self.ny, self.nz), self.ng
        for k in ("jx", "jy", "jz"):
            _pxr.summation_bcs(self.fields[k], g, g, g, *n)

    def efield_bcs(self):
        n, g = (self.nx, self.ny, self.nz), self.ng
        for k in ("ex", "ey", "ez"):
            _pxr.field_bc(self.fields[k], g, g, g, *n)

    def bfield_bcs(self):
        n, g = (self.nx, self.ny, self.nz), self.ng
        for k in ("bx", "by", "bz"):
            _pxr.field_bc(self.fields[k], g, g, g, *n)

    def push_bfield(self):
        """push_bfield (maxwell_solver_manager.F90:44): half step, 0.5*dt/dx."""
        f, g = self.fields, self.ng
        h = 0.5 * self.dt
        if self.norder == 2:
            lo, hi = (0, 0, 0), (self.nx, self.ny, self.nz)
            alo, ahi = (-g, -g, -g), (self.nx + g, self.ny + g, self.nz + g)
            _pxr.pxrpush_em3d_bvec(lo, hi, lo, hi, lo, hi,
                                   *sum(([f[k], alo, ahi] for k in ("ex", "ey", "ez", "bx", "by",
                                                                     "bz")), []),
                                   h / self.dx, h / self.dy, h / self.dz)
        else:
            c = self.coefs
            _pxr.pxrpush_em3d_bvec_norder(f["ex"], f["ey"], f["ez"], f["bx"], f["by"], f["bz"],
                                          h / self.dx * c, h / self.dy * c, h / self.dz * c,
                                          self.nx, self.ny, self.nz, self.norder, self.norder,
                                          self.norder, g, g, g, 0, 0, 0, False)

    def push_efield(self):
        """push_efield (maxwell_solver_manager.F90:393): c^2*dt/dx, mu0*c^2*dt."""
        f, g = self.fields, self.ng
        dt = self.dt
        if self.norder == 2:
            lo, hi = (0, 0, 0), (self.nx, self.ny, self.nz)
            alo, ahi = (-g, -g, -g), (self.nx + g, self.ny + g, self.nz + g)
            _pxr.pxrpush_em3d_evec(lo, hi, lo, hi, lo, hi,
                                   *sum(([f[k], alo, ahi] for k in ("ex", "ey", "ez", "bx", "by",
                                                                     "bz", "jx", "jy", "jz")), []),
                                   dt, dt / self.dx, dt / self.dy, dt / self.dz)
        else:
            c = self.coefs
            _pxr.pxrpush_em3d_evec_norder(f["ex"], f["ey"], f["ez"], f["bx"], f["by"], f["bz"],
                                          f["jx"], f["jy"], f["jz"], dt, dt / self.dx * c,
                                          dt / self.dy * c, dt / self.dz * c, self.nx, self.ny,
                                          self.nz, self.norder, self.norder, self.norder, g, g, g,
                                          0, 0, 0, False)

    def field_step(self):
        """submain.F90:160-175: B half, E, B half, each with its guard exchange."""
        self.push_bfield()
        self.bfield_bcs()
        self.push_efield()
        self.efield_bcs()
        self.push_bfield()
        self.bfield_bcs()

    def _one_step(self):
        if self.split_step:
            # submain.F90:99-140: push, particle boundaries and sort, then the
            # deposition from the particles as the boundaries left them
            self.gather_push()
            self.push_antennas()
            self.particle_bcs_and_sorting()
            self.depose_currents()
            self.current_bcs()
            self.field_step()
        else:
            self.gather_push_deposit()
            self.current_bcs()
            self.field_step()
            self.particle_bcs_and_sorting()

    def step(self, nst=1, check=True):
        for _ in range(int(nst)):
            self._one_step()
            self.it += 1
        if check:
            self.check()

    # ------------------------------------------------------------------
    # HIP graph of whole steps
    # ------------------------------------------------------------------
    def capture_steps(self, nst=2) -> "StepGraph":
        """Record ``nst`` steps (an even number: each step swaps the two
        buffers of every species, so after an even number the buffers are
        where they were) into a HIP graph without running them.  Run one
        step first so that the library's scratch buffers have their sizes;
        particle counts must not change while the graph is in use."""
        nst = int(nst)
        if nst < 2 or nst % 2:
            raise ValueError("capture an even number of steps")
        if self.antennas:
            raise RuntimeError("steps with a laser antenna cannot be captured: its push "
                               "depends on the time of each step")
        if self.absorbing:
            raise RuntimeError("steps with absorbing particle boundaries cannot be captured: "
                               "the particle counts change")
        if self.it == 0:
            raise RuntimeError("run one step before capturing (scratch buffers are sized then)")
        curs = [sp.cur for sp in self.species]
        _lib.call("pxrhip_graph_begin")
        try:
            for _ in range(nst):
                self._one_step()
        except Exception:
            h = C.c_void_p(0)
            try:
                _lib.call("pxrhip_graph_end", C.byref(h))
            finally:
                if h.value:
                    _lib.call("pxrhip_graph_destroy", h)
            raise
        h = C.c_void_p(0)
        _lib.call("pxrhip_graph_end", C.byref(h))
        for sp, c in zip(self.species, curs):
            if sp.cur != c:  # recording flipped the buffers; nothing ran
                raise RuntimeError("species buffers out of step after capture")
        return StepGraph(self, h.value, nst)

    def step_graph(self, graph: "StepGraph", nreplay=1, check=True):
        """Replay a captured graph ``nreplay`` times."""
        for _ in range(int(nreplay)):
            graph.launch()
        if check:
            self.check()

    # ------------------------------------------------------------------
    # diagnostics (diags.F90:847-1114, kin_energy.F90:33-143)
    # ------------------------------------------------------------------
    def kinetic_energy(self) -> float:
        """Sum over species of m c^2 sum (1/gaminv - 1) w (kin_energy.F90:33-143);
        for a mixed-precision species gamma - 1 is taken from the momenta
        in FP64 (pxrhip_kinetic_energy_f32)."""
        tot = 0.0
        for sp in self.species:
            out = C.c_double(0.0)
            b = sp.active
            if sp.precision == "mixed":
                _lib.call("pxrhip_kinetic_energy_f32", sp.npart, b.arrays["ux"].ptr,
                          b.arrays["uy"].ptr, b.arrays["uz"].ptr, b.arrays["w"].ptr,
                          C.byref(out))
            else:
                _lib.call("pxrhip_kinetic_energy", sp.npart, b.arrays["gaminv"].ptr,
                          b.arrays["w"].ptr, C.byref(out))
            tot += out.value * sp.m
        return tot

    def field_energy(self) -> float:
        """(1/2) sum (E^2 + B^2) dV over nodes 0..n-1 (diags.F90:1072-1114)."""
        vol = self.dx * self.dy * self.dz
        e = 0.0
        for k in ("ex", "ey", "ez", "bx", "by", "bz"):
            out = C.c_double(0.0)
            _lib.call("pxrhip_field_energy", self.fields[k].ptr, self.nx, self.ny, self.nz,
                      self.ng, self.ng, self.ng, C.byref(out))
            e += out.value
        return 0.5 * e * vol

    def total_particles(self) -> int:
        return sum(sp.npart for sp in self.species)

    def pushed_particles(self) -> int:
        """Particles that go through gather + push (frozen species excluded)."""
        return sum(sp.npart for sp in self.species if not sp.lfreeze)

    # ------------------------------------------------------------------
    # charge density and Gauss's law (charge_deposition_manager.F90:63-743,
    # field_boundaries.F90:1910-1930, diags.F90:63-197, 1222-1345)
    # ------------------------------------------------------------------
    def diag_grid(self, name: str) -> DeviceGrid:
        """The diagnostic grid array ``name`` (rho, dive, divb), zeroed when
        it is first used."""
        g = self.diag_fields.get(name)
        if g is None:
            g = DeviceGrid(self.nx, self.ny, self.nz, self.ng)
            g.fill(0.0)
            self.diag_fields[name] = g
        return g

    def pxrdepose_rho_on_grid(self):
        """rho = 0, then the charge density of every species that deposits
        (frozen ones included) and of the laser antennas, which the
        reference holds as species that deposit (tiling.F90:1684-1697), at
        the simulation's shape order: the species with the tiled kernel
        (LDS blocks per work item), the antennas with the leaf one."""
        rho = self.diag_grid("rho")
        rho.fill(0.0)
        g = self.ng
        for sp in self.species:
            if sp.ldodepos and sp.npart:
                s = sp.struct(sp.cur)
                _lib.call("pxrhip_tiles_depose_rho", C.byref(self.grid), C.byref(s), rho.ptr,
                          self.order)
        for ant in self.antennas:
            if ant.npart:  # pid column 1 holds the weights
                _lib.call("pxrhip_depose_rho_n", rho.ptr, ant.npart, ant.x.ptr, ant.y.ptr,
                          ant.z.ptr, ant.pid.ptr, ant.charge, self.xmin, self.ymin, self.zmin,
                          self.dx, self.dy, self.dz, self.nx, self.ny, self.nz, g, g, g,
                          self.order, self.order, self.order, 1, 0)

    def charge_bcs(self):
        """Periodic fold of the rho guards (summation_bcs on rho)."""
        g = self.ng
        _pxr.summation_bcs(self.diag_grid("rho"), g, g, g, self.nx, self.ny, self.nz)

    def calc_field_div(self) -> DeviceGrid:
        """div E at the nodes 0..n (backward differences)."""
        f, g = self.fields, self.ng
        d = self.diag_grid("dive")
        _pxr.calc_field_div(d, f["ex"], f["ey"], f["ez"], self.nx, self.ny, self.nz, g, g, g,
                            self.dx, self.dy, self.dz)
        return d

    def calc_field_divB(self) -> DeviceGrid:
        """div B at the nodes 1..n (forward differences)."""
        f, g = self.fields, self.ng
        d = self.diag_grid("divb")
        _pxr.calc_field_divB(d, f["bx"], f["by"], f["bz"], self.nx, self.ny, self.nz, g, g, g,
                             self.dx, self.dy, self.dz)
        return d

    def get_norm_divErho(self) -> float:
        """sqrt(sum over the nodes 1..n of (div E * eps0 - rho)^2) of the
        current dive and rho arrays."""
        g = self.ng
        return _pxr.get_norm_divErho(self.diag_grid("dive"), self.diag_grid("rho"), self.nx,
                                     self.ny, self.nz, g, g, g)

    def gauss_residual(self) -> float:
        """Gauss's law residual of the present state: rho deposited and
        folded, div E computed, then the norm of div E*eps0 - rho over the
        nodes 1..n.  With the Esirkepov current and the Yee solver the
        residual array keeps its initial value to round-off on the nodes
        0..n-1 of each axis; node n of rho holds the node-n deposit twice
        (the periodic fold of one rank, field_boundaries.F90:431-441), as in
        the reference, and enters the norm."""
        self.pxrdepose_rho_on_grid()
        self.charge_bcs()
        self.calc_field_div()
        return self.get_norm_divErho()


class StepGraph:
    """A HIP graph of ``nst`` whole steps of one Simulation."""

    def __init__(self, sim: "Simulation", handle: int, nst: int):
        self.sim, self.handle, self.nst = sim, handle, nst

    def launch(self):
        _lib.call("pxrhip_graph_launch", C.c_void_p(self.handle))
        self.sim.it += self.nst

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                _lib.call("pxrhip_graph_destroy", C.c_void_p(h))
            except Exception:
                pass
            self.handle = None


def _fd_weights(norder: int) -> np.ndarray:
    """FD_weights (submain.F90:927-984) for a staggered grid, as the
    reference computes them (including its column update inside the IF)."""
    n = norder + 1
    z = 0.5
    m = 1
    x = np.array([(i - n // 2 + 1) * 1.0 for i in range(n)])
    c = np.zeros((m + 1, n))
    c1 = 1.0
    c4 = x[0] - z
    c[0, 0] = 1.0
    for i in range(1, n):
        mn = min(i + 1, m + 1)
        c2 = 1.0
        c5 = c4
        c4 = x[i] - z
        for j in range(i):
            c3 = x[i] - x[j]
            c2 = c2 * c3
            if j == i - 1:
                for k in range(1, mn):
                    c[k, i] = c1 * (k * c[k - 1, i - 1] - c5 * c[k, i - 1]) / c2
                c[0, i] = -c1 * c5 * c[0, i - 1] / c2
                for k in range(1, mn):
                    c[k, j] = (c4 * c[k, j] - k * c[k - 1, j]) / c3
                c[0, j] = c4 * c[0, j] / c3
        c1 = c2
    return np.array([c[m, norder // 2 + i - 1] for i in range(1, norder // 2 + 1)])
