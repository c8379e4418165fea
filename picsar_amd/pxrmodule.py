"""The module state and set-up routines of ``picsar_python.picsarpy.picsar``.

PICSAR's Python drivers (examples/example_scripts_python/test.py) configure
the Fortran modules through variables of the Forthon object ``picsar``
(``pxr.nx_global_grid = 100``, ``pxr.nox = 1``, ``pxr.ntilex = 5`` ...), call
the set-up routines (default_init, set_particle_species_properties,
mpi_minimal_init, mpi_initialise, set_tile_split, init_tile_arrays,
py_add_particles_to_species, init_stencil_coefficients) and then step with
the zero-argument managers.  ``PicsarModule`` keeps the same variables and
routines for one rank and runs the managers on the GPU: init_tile_arrays
builds a ``picsar_amd.sim.Simulation`` from the variables, particles added
by py_add_particles_to_species are uploaded into tiled device species before
the next manager runs, and the field arrays read and assign as numpy arrays
in the reference's layout.  Leaf routines (geteb3d_energy_conserving,
depose_jxjyjz, ...) come from ``picsar_amd.picsarpy.picsar``.

Units are those of the reference's -DLIBRARY build (clight = eps0 = mu0 = 1).
Differences from the Fortran module, each raised as an error where it
matters: one rank (nprocx = nprocy = nprocz = 1); equal shape orders
nox = noy = noz in 1..3 and equal FDTD orders; one guard width for every
grid array (the largest of nxguards ... nzjguards); the particle's pid
column wpid = 1 (the weight) is kept, other columns are not; every manager
step sorts (the tiled kernels need cell-sorted tiles), whatever
sorting_activated says.
"""
from __future__ import annotations

from typing import Dict, List

import numpy as np

from .picsarpy import picsar as _leaf

# module variables and their values after default_init
# (control_file.F90:77-239)
_DEFAULTS = dict(
    c_dim=3, ntilex=1, ntiley=1, ntilez=1, norderx=2, nordery=2, norderz=2,
    l_nodalgrid=False, l_staggered=True, nox=1, noy=1, noz=1,
    nxguards=2, nyguards=2, nzguards=2, nxjguards=2, nyjguards=2, nzjguards=2,
    topology=0, mpicom_curr=1, currdepo=0, rhodepo=0, fieldgathe=0, partcom=0,
    particle_pusher=0, lvec_curr_depo=8, lvec_charge_depo=64, lvec_fieldgathe=256,
    mpi_buf_size=2000, sorting_activated=0, sorting_dx=1.0, sorting_dy=1.0, sorting_dz=1.0,
    sorting_shiftx=0.0, sorting_shifty=0.0, sorting_shiftz=0.0, l_lower_order_in_v=True,
    dtcoef=0.7, tmax=0.0, nsteps=0, nlab=1.0e23, g0=1.0, nspecies=0,
    pbound_x_min=0, pbound_x_max=0, pbound_y_min=0, pbound_y_max=0, pbound_z_min=0,
    pbound_z_max=0, nprocx=1, nprocy=1, nprocz=1,
    # not set by default_init; set here so that drivers can read them
    nx_global_grid=0, ny_global_grid=0, nz_global_grid=0,
    xmin=0.0, xmax=0.0, ymin=0.0, ymax=0.0, zmin=0.0, zmax=0.0, dt=0.0, it=0,
    # this library's own choice of particle storage: "fp64" or "mixed"
    precision="fp64",
)

_FIELDS = ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")


class PicsarModule:
    """One rank of PICSAR's module state, stepped on the GPU."""

    clight = 1.0
    eps0 = 1.0
    mu0 = 1.0
    imu0 = 1.0
    pi = np.pi

    def __init__(self):
        object.__setattr__(self, "_sim", None)
        object.__setattr__(self, "_staged", {})
        object.__setattr__(self, "_species_props", {})
        object.__setattr__(self, "_ready", set())
        for k, v in _DEFAULTS.items():
            object.__setattr__(self, k, v)
        self.b0 = 0.0
        self.nc = self.nlab * self.g0
        self.rank = 0
        self.nproc = 1
        self.xcoeffs = self.ycoeffs = self.zcoeffs = None

    # ------------------------------------------------------------------
    # module variables that live on the device: the field arrays
    # ------------------------------------------------------------------
    def __getattr__(self, name):
        sim = object.__getattribute__(self, "_sim")
        if name in _FIELDS and sim is not None:
            return sim.fields[name].numpy()
        if name in ("rho", "dive", "divb") and sim is not None:
            return sim.diag_grid(name).numpy()
        try:
            return getattr(_leaf, name)  # the leaf routines
        except AttributeError:
            raise AttributeError(f"picsar has no variable or routine '{name}'") from None

    def __setattr__(self, name, value):
        sim = self.__dict__.get("_sim")
        if name in _FIELDS and sim is not None:
            sim.fields[name].upload(np.asfortranarray(value, dtype=np.float64))
            return
        object.__setattr__(self, name, value)

    # ------------------------------------------------------------------
    # set-up
    # ------------------------------------------------------------------
    def default_init(self):
        """control_file.F90:77-239: every module variable to its default."""
        for k, v in _DEFAULTS.items():
            if k not in ("nx_global_grid", "ny_global_grid", "nz_global_grid", "xmin", "xmax",
                         "ymin", "ymax", "zmin", "zmax", "dt", "it", "precision"):
                object.__setattr__(self, k, v)
        self.b0 = float(np.sqrt(1.0 - 1.0 / self.g0 ** 2))
        self.nc = self.nlab * self.g0

    def set_particle_species_properties(self, nsp, sname, mss, chrg, nppc, xsmin, ysmin, zsmin,
                                        xsmax, ysmax, zsmax, vdxs, vdys, vdzs, vthxs, vthys,
                                        vthzs, sorting_period=0, sorting_start=0, dodepos=True):
        """tiling.F90:1335-1367 (species nsp is 1-based; dodepos is
        particle_species%ldodepos)."""
        self._species_props[int(nsp)] = dict(
            name=str(sname), mass=float(mss), charge=float(chrg), nppcell=int(nppc),
            box=(xsmin, ysmin, zsmin, xsmax, ysmax, zsmax), vdrift=(vdxs, vdys, vdzs),
            vth=(vthxs, vthys, vthzs), sorting_period=int(sorting_period),
            sorting_start=int(sorting_start), ldodepos=bool(dodepos), lfreeze=False)
        self.nspecies = max(int(self.nspecies), int(nsp))

    def set_species_freeze(self, nsp, lfreeze=True):
        """particle_species%lfreeze of species nsp (modules.F90:570)."""
        self._species_props[int(nsp)]["lfreeze"] = bool(lfreeze)

    def mpi_minimal_init(self, comm=-1):
        """One rank: rank 0 of 1."""
        if int(self.nprocx) * int(self.nprocy) * int(self.nprocz) != 1:
            raise NotImplementedError("this build runs one rank (nprocx = nprocy = nprocz = 1)")
        self.rank = 0
        self.nproc = 1

    def mpi_minimal_init_python(self, comm=-1):
        self.mpi_minimal_init(comm)

    def mpi_initialise(self):
        """The grid of the one rank (mpi_routines.F90:209, 1327-1698):
        nx_global = nx_global_grid - 1 cells, dx = (xmax - xmin)/nx_global,
        local extents equal to the global ones."""
        self.mpi_minimal_init()
        for a in "xyz":
            ng = int(getattr(self, f"n{a}_global_grid"))
            if ng < 2:
                raise ValueError(f"n{a}_global_grid must be >= 2")
            n = ng - 1
            lo, hi = float(getattr(self, f"{a}min")), float(getattr(self, f"{a}max"))
            setattr(self, f"n{a}_global", n)
            setattr(self, f"n{a}", n)
            setattr(self, f"n{a}_grid", n + 1)
            setattr(self, f"d{a}", (hi - lo) / float(n))
            setattr(self, f"length_{a}", hi - lo)
            for k in (f"{a}_min_local", f"{a}_grid_min_local", f"{a}_grid_min"):
                setattr(self, k, lo)
            for k in (f"{a}_max_local", f"{a}_grid_max_local", f"{a}_grid_max"):
                setattr(self, k, hi)

    def set_tile_split(self):
        """tiling.F90:66-80: checks the tile counts (each tile at least 4
        nodes wide, tiling.F90:110-126); the split itself is the library's
        (set_tile_split_for_species' arithmetic)."""
        for a in "xyz":
            n, nt = int(getattr(self, f"n{a}")), int(getattr(self, f"ntile{a}"))
            if nt < 1 or (n + 1) // nt < 4:
                raise ValueError(f"ntile{a} = {nt}: tiles must be at least 4 nodes wide")

    def init_tile_arrays(self):
        """Allocates the rank's field arrays and tiled species on the GPU."""
        from .sim import Simulation
        orders = {int(self.nox), int(self.noy), int(self.noz)}
        norders = {int(self.norderx), int(self.nordery), int(self.norderz)}
        if len(orders) != 1:
            raise NotImplementedError("nox, noy and noz must be equal")
        if len(norders) != 1:
            raise NotImplementedError("norderx, nordery and norderz must be equal")
        order, norder = orders.pop(), norders.pop()
        if int(self.particle_pusher) not in (0, 1):
            raise NotImplementedError("particle_pusher must be 0 (Boris) or 1 (Vay); the "
                                      "radiation-reaction pushers are not built")
        ng = max(int(self.nxguards), int(self.nyguards), int(self.nzguards),
                 int(self.nxjguards), int(self.nyjguards), int(self.nzjguards),
                 order, norder // 2)
        dt = float(self.dt) if self.dt else None
        # sorting_d* hold absolute sizes once the driver has scaled them by dx
        # (test.py:125-127); Simulation takes them in cells
        sim = Simulation(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz, self.xmin,
                         self.ymin, self.zmin, self.ntilex, self.ntiley, self.ntilez,
                         order=order, l_lower_order_in_v=bool(self.l_lower_order_in_v),
                         pusher="vay" if int(self.particle_pusher) == 1 else "boris",
                         depo="direct" if int(self.currdepo) in (3, 4, 5) else "esirkepov",
                         norder=norder, dt=dt, nguard=ng,
                         sorting_dx=float(self.sorting_dx) / self.dx,
                         sorting_dy=float(self.sorting_dy) / self.dy,
                         sorting_dz=float(self.sorting_dz) / self.dz,
                         precision=self.precision,
                         pbound=[int(self.pbound_x_min), int(self.pbound_x_max),
                                 int(self.pbound_y_min), int(self.pbound_y_max),
                                 int(self.pbound_z_min), int(self.pbound_z_max)],
                         partcom=int(self.partcom))
        object.__setattr__(self, "_sim", sim)
        self.nxguards = self.nyguards = self.nzguards = ng
        self.nxjguards = self.nyjguards = self.nzjguards = ng

    def py_add_particles_to_species(self, nsp, npart, *args):
        """tiling.F90:1382-1397: append npart particles to species nsp.
        Arguments after npart: npiid, x, y, z, ux, uy, uz, gaminv, pid
        (pid of shape (npart, npiid), column 1 the weight); the older form
        x, y, z, ux, uy, uz, gaminv, w is accepted too."""
        n = int(npart)
        extra = None
        if len(args) == 9:
            npiid, x, y, z, ux, uy, uz, gaminv, pid = args
            pid = np.asarray(pid, dtype=np.float64).reshape(n, int(npiid), order="F")
            w = pid[:, 0]
            if int(npiid) > 1:  # columns 2.. (ids, old positions ...) travel with the particles
                extra = np.ascontiguousarray(pid[:, 1:])
        elif len(args) == 8:
            x, y, z, ux, uy, uz, gaminv, w = args
        else:
            raise TypeError("py_add_particles_to_species(nsp, npart, npiid, x, y, z, ux, uy, "
                            "uz, gaminv, pid)")
        arrs = {k: np.asarray(v, dtype=np.float64).ravel()[:n].copy()
                for k, v in zip(("x", "y", "z", "ux", "uy", "uz", "gaminv", "w"),
                                (x, y, z, ux, uy, uz, gaminv, w))}
        for k, v in arrs.items():
            if v.size != n:
                raise ValueError(f"{k}: {v.size} values, expected {n}")
        if extra is not None:
            arrs["pid"] = extra
        staged = self._staged.setdefault(int(nsp), [])
        if staged and (("pid" in staged[0]) != ("pid" in arrs) or
                       ("pid" in arrs and staged[0]["pid"].shape[1] != arrs["pid"].shape[1])):
            raise ValueError("every call for one species must pass the same number of pid "
                             "columns")
        staged.append(arrs)

    def init_stencil_coefficients(self):
        """submain.F90:890-908: FD_weights of the FDTD orders."""
        from .sim import _fd_weights
        self.xcoeffs = _fd_weights(int(self.norderx))
        self.ycoeffs = _fd_weights(int(self.nordery))
        self.zcoeffs = _fd_weights(int(self.norderz))

    def mpi_close(self):
        """Frees the rank's device state."""
        object.__setattr__(self, "_sim", None)
        self._ready.clear()
        self._staged.clear()

    # ------------------------------------------------------------------
    # particles: staged on the host, uploaded before a manager runs
    # ------------------------------------------------------------------
    def _state(self):
        sim = self._sim
        if sim is None:
            raise RuntimeError("call init_tile_arrays first")
        if self.dt:
            sim.dt = float(self.dt)
        else:
            self.dt = sim.dt
        sim.it = int(self.it)
        for nsp in sorted(self._staged):
            new = self._staged[nsp]
            if not new:
                continue
            props = self._species_props.get(nsp)
            if props is None:
                raise ValueError(f"species {nsp}: call set_particle_species_properties first")
            parts = []
            idx = self._species_index(nsp)
            if idx is not None:  # species already on the device: append
                sp = sim.species[idx]
                a = sp.arrays()
                old = {k: v.astype(np.float64) for k, v in a.items() if not k.startswith("pid")}
                if sp.npid_extra:
                    old["pid"] = np.stack([a[f"pid{c + 2}"] for c in range(sp.npid_extra)],
                                          axis=1)
                parts.append(old)
            parts.extend(new)
            if any(set(p) != set(parts[0]) or
                   ("pid" in p and p["pid"].shape[1] != parts[0]["pid"].shape[1])
                   for p in parts):
                raise ValueError(f"species {nsp}: particles added with different numbers of "
                                 "pid columns")
            if idx is not None:
                sim.species.pop(idx)
                self._ready.discard(nsp)
            cat = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
            sp = sim.add_species(props["name"], props["charge"], props["mass"], cat["x"],
                                 cat["y"], cat["z"], cat["ux"], cat["uy"], cat["uz"], cat["w"],
                                 gaminv=cat["gaminv"], lfreeze=props["lfreeze"],
                                 ldodepos=props["ldodepos"], pid=cat.get("pid"))
            sp.pxr_index = nsp
            self._ready.add(nsp)
            self._staged[nsp] = []
        sim.species.sort(key=lambda s: getattr(s, "pxr_index", 0))
        return sim

    def _species_index(self, nsp):
        for i, sp in enumerate(self._sim.species):
            if getattr(sp, "pxr_index", None) == nsp:
                return i
        return None

    def species_arrays(self, nsp) -> Dict[str, np.ndarray]:
        """Host copies of species nsp's particle arrays (tile by tile)."""
        sim = self._state()
        i = self._species_index(int(nsp))
        if i is None:
            return {}
        return sim.species[i].arrays()

    def _managers(self):
        from .managers import Managers
        return Managers(self._state())

    def _done(self, sim):
        self.it = sim.it

    # ------------------------------------------------------------------
    # the zero-argument managers (submain.F90:99-199)
    # ------------------------------------------------------------------
    def field_gathering_plus_particle_pusher(self):
        self._managers().field_gathering_plus_particle_pusher()

    def push_laser_particles(self):
        self._managers().push_laser_particles()

    def particle_bcs(self):
        self._managers().particle_bcs()

    def particle_sorting_sub(self):
        """Done by particle_bcs: its re-binning writes every tile sorted."""
        self._state()

    def pxr_particle_sorting(self):
        self.particle_sorting_sub()

    def pxrdepose_currents_on_grid_jxjyjz(self):
        self._managers().pxrdepose_currents_on_grid_jxjyjz()

    def current_bcs(self):
        self._managers().current_bcs()

    def push_bfield(self):
        self._managers().push_bfield()

    def bfield_bcs(self):
        self._managers().bfield_bcs()

    def push_efield(self):
        self._managers().push_efield()

    def efield_bcs(self):
        self._managers().efield_bcs()

    def pxrdepose_rho_on_grid(self):
        self._managers().pxrdepose_rho_on_grid()

    def charge_bcs(self):
        self._managers().charge_bcs()

    def calc_diags(self):
        self._managers().calc_diags()

    def calc_field_div(self):
        self._managers().calc_field_div()

    def get_norm_divErho(self) -> float:
        return self._managers().get_norm_divErho()

    def step(self, nst=1):
        """submain.F90:51-199: nst whole steps with the managers; it advances."""
        mgr = self._managers()
        mgr.step(int(nst))
        self._done(mgr.sim)

    def get_kinetic_energy(self) -> float:
        """Total kinetic energy of the species (kin_energy.F90:33-143)."""
        return self._state().kinetic_energy()

    def get_field_energy(self) -> float:
        """Electromagnetic field energy (diags.F90:1072-1114)."""
        return self._state().field_energy()


picsar = PicsarModule()
