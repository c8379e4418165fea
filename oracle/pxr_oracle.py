"""ctypes bindings for the oracle libraries (test infrastructure only).

``Oracle`` wraps ``liboracle.so`` (the C restatement, value arguments).
``Reference`` wraps ``_ref/libpxrref.so`` (the reference's own Fortran
kernels, every argument passed by reference, symbol names lower case with a
trailing underscore).

Grid arrays are numpy float64 arrays in Fortran order with shape
(nx+1+2*ngx, ny+1+2*ngy, nz+1+2*ngz); particle arrays are contiguous float64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpxrref.so")
# the same kernels built with the reference's production flags (-O3 -fopenmp)
REF_O3_SO = os.path.join(HERE, "_ref", "libpxrref_o3.so")
# OpenMP tile loop over the production kernels (ref_tile_driver.c)
REF_DRIVER_SO = os.path.join(HERE, "_ref", "libpxrref_driver.so")

_i64 = C.c_int64
_f64 = C.c_double
_p = C.c_void_p


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_p)


def _chk_part(*arrs):
    for a in arrs:
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError("particle arrays must be contiguous float64")


def _chk_grid(*arrs):
    for a in arrs:
        if a.dtype != np.float64 or not a.flags.f_contiguous:
            raise TypeError("grid arrays must be Fortran-ordered float64")


def build_oracle(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(ORACLE_SO) or (
        os.path.getmtime(ORACLE_SO) < os.path.getmtime(os.path.join(HERE, "pxr_oracle.c"))
    ):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return ORACLE_SO


class Oracle:
    """numpy front end of the C restatement."""

    def __init__(self, path: str = ORACLE_SO):
        self.lib = C.CDLL(path)
        L = self.lib
        P, I, F = _p, _i64, _f64
        L.orc_geteb3d_energy_conserving.restype = I
        L.orc_geteb3d_energy_conserving.argtypes = [I] + [P] * 9 + [F] * 6 + [I] * 9 + [P] * 6 + [I, I]
        L.orc_boris_push_u_3d.restype = None
        L.orc_boris_push_u_3d.argtypes = [I] + [P] * 10 + [F] * 4
        L.orc_set_gamma.restype = None
        L.orc_set_gamma.argtypes = [I] + [P] * 4 + [F]
        L.orc_pushxyz.restype = None
        L.orc_pushxyz.argtypes = [I] + [P] * 7 + [F]
        L.orc_ebcancelpush3d.restype = None
        L.orc_ebcancelpush3d.argtypes = [I] + [P] * 10 + [F] * 3 + [I, F]
        L.orc_depose_direct.restype = I
        L.orc_depose_direct.argtypes = [P] * 3 + [I] + [P] * 8 + [F] * 8 + [I] * 6 + [I, I]
        L.orc_depose_esirkepov.restype = I
        L.orc_depose_esirkepov.argtypes = [P] * 3 + [I] + [P] * 8 + [F] * 8 + [I] * 6 + [I]
        L.orc_depose_rho.restype = I
        L.orc_depose_rho.argtypes = [P, I] + [P] * 4 + [F] * 7 + [I] * 6 + [I]
        L.orc_depose_rho_n.restype = I
        L.orc_depose_rho_n.argtypes = [P, I] + [P] * 4 + [F] * 7 + [I] * 6 + [I] * 5
        L.orc_depose_esirkepov_n.restype = I
        L.orc_depose_esirkepov_n.argtypes = [P] * 3 + [I] + [P] * 8 + [F] * 8 + [I] * 6 + [I] * 3
        L.orc_geteb3d_n.restype = I
        L.orc_geteb3d_n.argtypes = [I] + [P] * 9 + [F] * 6 + [I] * 9 + [P] * 6 + [I] * 3
        L.orc_push_em3d_evec.restype = None
        L.orc_push_em3d_evec.argtypes = [I] * 6 + [P] * 9 + [F] * 4 + [I] * 6
        L.orc_push_em3d_bvec.restype = None
        L.orc_push_em3d_bvec.argtypes = [I] * 6 + [P] * 6 + [F] * 3 + [I] * 6
        L.orc_push_em3d_evec_norder.restype = None
        L.orc_push_em3d_evec_norder.argtypes = [P] * 9 + [F] + [P] * 3 + [I] * 13
        L.orc_push_em3d_bvec_norder.restype = None
        L.orc_push_em3d_bvec_norder.argtypes = [P] * 6 + [P] * 3 + [I] * 13
        L.orc_fd_weights.restype = None
        L.orc_fd_weights.argtypes = [P, I, I]
        L.orc_particle_bin_sorting.restype = I
        L.orc_particle_bin_sorting.argtypes = [I] + [P] * 8 + [F] * 9 + [P]
        L.orc_field_bc_periodic.restype = None
        L.orc_field_bc_periodic.argtypes = [P] + [I] * 6
        L.orc_current_bc_periodic.restype = None
        L.orc_current_bc_periodic.argtypes = [P] + [I] * 6
        L.orc_calc_field_div.restype = None
        L.orc_calc_field_div.argtypes = [P] * 4 + [I] * 6 + [F] * 3
        L.orc_calc_field_divb.restype = None
        L.orc_calc_field_divb.argtypes = [P] * 4 + [I] * 6 + [F] * 3
        L.orc_node_norm2.restype = F
        L.orc_node_norm2.argtypes = [P, P] + [I] * 6
        L.orc_laser_profile.restype = None
        L.orc_laser_profile.argtypes = [I, I, P, P, P, F, F, P, P, F, F, F, F, F, F, I, F]
        L.orc_laserp_pusher.restype = None
        L.orc_laserp_pusher.argtypes = [I, I, I] + [P] * 8 + [F] * 4 + [P, P] + [F] * 7 + [I, F]

    # -- gather ------------------------------------------------------------
    def geteb3d_energy_conserving(self, xp, yp, zp, ex, ey, ez, bx, by, bz, xmin, ymin, zmin,
                                  dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, nox, noy, noz,
                                  exg, eyg, ezg, bxg, byg, bzg, l_lower_order_in_v=True,
                                  l_nodal=False):
        _chk_part(xp, yp, zp, ex, ey, ez, bx, by, bz)
        _chk_grid(exg, eyg, ezg, bxg, byg, bzg)
        return self.lib.orc_geteb3d_energy_conserving(
            len(xp), _ptr(xp), _ptr(yp), _ptr(zp), _ptr(ex), _ptr(ey), _ptr(ez), _ptr(bx),
            _ptr(by), _ptr(bz), xmin, ymin, zmin, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz,
            nox, noy, noz, _ptr(exg), _ptr(eyg), _ptr(ezg), _ptr(bxg), _ptr(byg), _ptr(bzg),
            int(bool(l_lower_order_in_v)), int(bool(l_nodal)))

    # -- pushers -----------------------------------------------------------
    def boris_push_u_3d(self, uxp, uyp, uzp, gaminv, ex, ey, ez, bx, by, bz, q, m, dt,
                        clight=1.0):
        _chk_part(uxp, uyp, uzp, gaminv, ex, ey, ez, bx, by, bz)
        self.lib.orc_boris_push_u_3d(len(uxp), *map(_ptr, (uxp, uyp, uzp, gaminv, ex, ey, ez,
                                                         bx, by, bz)), q, m, dt, clight)

    def set_gamma(self, uxp, uyp, uzp, gaminv, clight=1.0):
        _chk_part(uxp, uyp, uzp, gaminv)
        self.lib.orc_set_gamma(len(uxp), *map(_ptr, (uxp, uyp, uzp, gaminv)), clight)

    def pushxyz(self, xp, yp, zp, uxp, uyp, uzp, gaminv, dt):
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv)
        self.lib.orc_pushxyz(len(xp), *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv)), dt)

    def ebcancelpush3d(self, uxp, uyp, uzp, gi, ex, ey, ez, bx, by, bz, q, m, dt, which=0,
                       clight=1.0):
        _chk_part(uxp, uyp, uzp, gi, ex, ey, ez, bx, by, bz)
        self.lib.orc_ebcancelpush3d(len(uxp), *map(_ptr, (uxp, uyp, uzp, gi, ex, ey, ez, bx,
                                                        by, bz)), q, m, dt, which, clight)

    # -- deposition --------------------------------------------------------
    def depose_direct(self, jx, jy, jz, xp, yp, zp, uxp, uyp, uzp, gaminv, w, q, xmin, ymin,
                      zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, order, l_nodal=False):
        _chk_grid(jx, jy, jz)
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv, w)
        return self.lib.orc_depose_direct(
            _ptr(jx), _ptr(jy), _ptr(jz), len(xp),
            *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv, w)),
            q, xmin, ymin, zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, order,
            int(bool(l_nodal)))

    def depose_esirkepov(self, jx, jy, jz, xp, yp, zp, uxp, uyp, uzp, gaminv, w, q, xmin, ymin,
                         zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, order):
        _chk_grid(jx, jy, jz)
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv, w)
        return self.lib.orc_depose_esirkepov(
            _ptr(jx), _ptr(jy), _ptr(jz), len(xp),
            *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv, w)),
            q, xmin, ymin, zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, order)

    def depose_rho(self, rho, xp, yp, zp, w, q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                   ngx, ngy, ngz, order):
        _chk_grid(rho)
        _chk_part(xp, yp, zp, w)
        return self.lib.orc_depose_rho(_ptr(rho), len(xp), _ptr(xp), _ptr(yp), _ptr(zp),
                                       _ptr(w), q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                                       ngx, ngy, ngz, order)

    # -- generic orders (one shape order per axis) ---------------------------
    def depose_rho_n(self, rho, xp, yp, zp, w, q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                     ngx, ngy, ngz, nox, noy, noz, l_particles_weight=True, l4symtry=False):
        """pxr_depose_rho_n (charge_deposition_3d.F90:2081), orders 0..3 per axis."""
        _chk_grid(rho)
        _chk_part(xp, yp, zp)
        w = np.ascontiguousarray(w, dtype=np.float64)
        return self.lib.orc_depose_rho_n(_ptr(rho), len(xp), _ptr(xp), _ptr(yp), _ptr(zp),
                                         _ptr(w), q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                                         ngx, ngy, ngz, nox, noy, noz,
                                         int(bool(l_particles_weight)), int(bool(l4symtry)))

    def depose_esirkepov_n(self, jx, jy, jz, xp, yp, zp, uxp, uyp, uzp, gaminv, w, q, xmin,
                           ymin, zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, nox, noy, noz):
        """pxr_depose_jxjyjz_esirkepov_n (esirkepov_3d.F90:3480), orders 0..3 per axis."""
        _chk_grid(jx, jy, jz)
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv, w)
        return self.lib.orc_depose_esirkepov_n(
            _ptr(jx), _ptr(jy), _ptr(jz), len(xp),
            *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv, w)),
            q, xmin, ymin, zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, nox, noy, noz)

    def geteb3d_n(self, xp, yp, zp, ex, ey, ez, bx, by, bz, xmin, ymin, zmin, dx, dy, dz, nx,
                  ny, nz, ngx, ngy, ngz, nox, noy, noz, exg, eyg, ezg, bxg, byg, bzg,
                  l4symtry=False, l_lower_order_in_v=True, l_nodal=False):
        """pxr_gete3d_n_energy_conserving then pxr_getb3d_n_energy_conserving
        (field_gathering_on_3d.F90:1042, 697), orders 1..3 per axis."""
        _chk_part(xp, yp, zp, ex, ey, ez, bx, by, bz)
        _chk_grid(exg, eyg, ezg, bxg, byg, bzg)
        return self.lib.orc_geteb3d_n(
            len(xp), _ptr(xp), _ptr(yp), _ptr(zp), _ptr(ex), _ptr(ey), _ptr(ez), _ptr(bx),
            _ptr(by), _ptr(bz), xmin, ymin, zmin, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz,
            nox, noy, noz, _ptr(exg), _ptr(eyg), _ptr(ezg), _ptr(bxg), _ptr(byg), _ptr(bzg),
            int(bool(l4symtry)), int(bool(l_lower_order_in_v)), int(bool(l_nodal)))

    # -- FDTD ---------------------------------------------------------------
    def push_em3d_evec(self, lo, hi, ex, ey, ez, bx, by, bz, jx, jy, jz, mudt, dtsdx, dtsdy,
                       dtsdz, nx, ny, nz, ngx, ngy, ngz):
        _chk_grid(ex, ey, ez, bx, by, bz, jx, jy, jz)
        self.lib.orc_push_em3d_evec(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                                    *map(_ptr, (ex, ey, ez, bx, by, bz, jx, jy, jz)),
                                    mudt, dtsdx, dtsdy, dtsdz, nx, ny, nz, ngx, ngy, ngz)

    def push_em3d_bvec(self, lo, hi, ex, ey, ez, bx, by, bz, dtsdx, dtsdy, dtsdz, nx, ny, nz,
                       ngx, ngy, ngz):
        _chk_grid(ex, ey, ez, bx, by, bz)
        self.lib.orc_push_em3d_bvec(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                                    *map(_ptr, (ex, ey, ez, bx, by, bz)),
                                    dtsdx, dtsdy, dtsdz, nx, ny, nz, ngx, ngy, ngz)

    def push_em3d_evec_norder(self, ex, ey, ez, bx, by, bz, jx, jy, jz, mudt, cx, cy, cz,
                              nx, ny, nz, nox, noy, noz, ngx, ngy, ngz, nxs, nys, nzs,
                              l_nodal=False):
        _chk_grid(ex, ey, ez, bx, by, bz, jx, jy, jz)
        cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in (cx, cy, cz))
        self.lib.orc_push_em3d_evec_norder(*map(_ptr, (ex, ey, ez, bx, by, bz, jx, jy, jz)),
                                           mudt, _ptr(cx), _ptr(cy), _ptr(cz), nx, ny, nz,
                                           nox, noy, noz, ngx, ngy, ngz, nxs, nys, nzs,
                                           int(bool(l_nodal)))

    def push_em3d_bvec_norder(self, ex, ey, ez, bx, by, bz, cx, cy, cz, nx, ny, nz, nox, noy,
                              noz, ngx, ngy, ngz, nxs, nys, nzs, l_nodal=False):
        _chk_grid(ex, ey, ez, bx, by, bz)
        cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in (cx, cy, cz))
        self.lib.orc_push_em3d_bvec_norder(*map(_ptr, (ex, ey, ez, bx, by, bz)), _ptr(cx),
                                           _ptr(cy), _ptr(cz), nx, ny, nz, nox, noy, noz,
                                           ngx, ngy, ngz, nxs, nys, nzs, int(bool(l_nodal)))

    def fd_weights(self, norder, l_nodal=False):
        c = np.zeros(max(norder // 2, 1))
        self.lib.orc_fd_weights(_ptr(c), norder, int(bool(l_nodal)))
        return c[: norder // 2]

    # -- sorting and guards ------------------------------------------------
    def particle_bin_sorting(self, xp, yp, zp, ux, uy, uz, gam, pid, xmin2, ymin2, zmin2,
                             xmax2, ymax2, zmax2, dxf, dyf, dzf, return_keys=False):
        _chk_part(xp, yp, zp, ux, uy, uz, gam, pid)
        keys = np.zeros(len(xp), dtype=np.int64) if return_keys else None
        bad = self.lib.orc_particle_bin_sorting(
            len(xp), *map(_ptr, (xp, yp, zp, ux, uy, uz, gam, pid)), xmin2, ymin2, zmin2,
            xmax2, ymax2, zmax2, dxf, dyf, dzf, _ptr(keys))
        return (bad, keys) if return_keys else bad

    def field_bc_periodic(self, f, nx, ny, nz, gx, gy, gz):
        _chk_grid(f)
        self.lib.orc_field_bc_periodic(_ptr(f), nx, ny, nz, gx, gy, gz)

    def current_bc_periodic(self, a, nx, ny, nz, gx, gy, gz):
        _chk_grid(a)
        self.lib.orc_current_bc_periodic(_ptr(a), nx, ny, nz, gx, gy, gz)

    # -- divergence diagnostics ----------------------------------------------
    def calc_field_div(self, dv, ex, ey, ez, nx, ny, nz, gx, gy, gz, dx, dy, dz):
        _chk_grid(dv, ex, ey, ez)
        self.lib.orc_calc_field_div(*map(_ptr, (dv, ex, ey, ez)), nx, ny, nz, gx, gy, gz, dx, dy,
                                    dz)

    def calc_field_divb(self, dv, bx, by, bz, nx, ny, nz, gx, gy, gz, dx, dy, dz):
        _chk_grid(dv, bx, by, bz)
        self.lib.orc_calc_field_divb(*map(_ptr, (dv, bx, by, bz)), nx, ny, nz, gx, gy, gz, dx, dy,
                                     dz)

    def norm_diverho(self, dv, rho, nx, ny, nz, gx, gy, gz):
        """get_norm_divErho on one domain (square-rooted)."""
        _chk_grid(dv, rho)
        return float(np.sqrt(self.lib.orc_node_norm2(_ptr(dv), _ptr(rho), nx, ny, nz, gx, gy,
                                                     gz)))

    def loc_norm_2(self, a, nx, ny, nz, gx, gy, gz):
        _chk_grid(a)
        return self.lib.orc_node_norm2(_ptr(a), None, nx, ny, nz, gx, gy, gz)

    # -- laser antenna -------------------------------------------------------
    def laser_profile(self, window, xx, yy, emax1, emax2, pv1, pv2, k0, q_z, tau, t, t_peak,
                      temporal_order, polangle):
        """(n, 3) amplitudes: window 0 gaussian_profile, 1 hanning_profile."""
        _chk_part(xx, yy)
        amp = np.zeros((len(xx), 3))
        pv1 = np.ascontiguousarray(pv1, dtype=np.float64)
        pv2 = np.ascontiguousarray(pv2, dtype=np.float64)
        self.lib.orc_laser_profile(window, len(xx), _ptr(xx), _ptr(yy), _ptr(amp), emax1, emax2,
                                   _ptr(pv1), _ptr(pv2), k0, complex(q_z).real,
                                   complex(q_z).imag, tau, t, t_peak, temporal_order, polangle)
        return amp

    def laserp_pusher(self, window, pid, xp, yp, zp, uxp, uyp, uzp, gaminv, dt, emax, emax1,
                      emax2, pv1, pv2, k0, q_z, charge, tau, t, t_peak, temporal_order,
                      polangle):
        """laserp_pusher_gaussian (window 0) / laserp_pusher_hanning (1);
        pid is (np, npid) Fortran-ordered."""
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv)
        assert pid.flags.f_contiguous and pid.dtype == np.float64 and pid.shape[0] == len(xp)
        pv1 = np.ascontiguousarray(pv1, dtype=np.float64)
        pv2 = np.ascontiguousarray(pv2, dtype=np.float64)
        self.lib.orc_laserp_pusher(window, len(xp), pid.shape[1], _ptr(pid),
                                   *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv)), dt, emax,
                                   emax1, emax2, _ptr(pv1), _ptr(pv2), k0, complex(q_z).real,
                                   complex(q_z).imag, charge, tau, t, t_peak, temporal_order,
                                   polangle)


# ---------------------------------------------------------------------------
# The reference's own kernels (Fortran, by reference)
# ---------------------------------------------------------------------------
def _ri(v):
    return C.byref(_i64(int(v)))


def _rf(v):
    return C.byref(_f64(float(v)))


def _iv(*v):
    a = np.array(v, dtype=np.int64)
    return a


class Reference:
    """numpy front end of the reference's compiled FP64 leaf kernels."""

    GATHER_DEFAULT = {1: "geteb3d_energy_conserving_vecv4_1_1_1_",
                      2: "geteb3d_energy_conserving_vecv4_2_2_2_",
                      3: "geteb3d_energy_conserving_vecv3_3_3_3_"}

    def __init__(self, path: str = REF_SO):
        self.lib = C.CDLL(path)

    def _f(self, name):
        fn = getattr(self.lib, name)
        fn.restype = None
        return fn

    # geteb3d_energy_conserving_generic(..., field_gathe_algo=0) for equal
    # orders 1..3 calls the vecV4/vecV3 routines
    # (field_gathering_manager_3d.F90:598-627); algo=1 calls the scalar ones.
    def geteb3d_energy_conserving(self, xp, yp, zp, ex, ey, ez, bx, by, bz, xmin, ymin, zmin,
                                  dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, order, exg, eyg, ezg,
                                  bxg, byg, bzg, l_lower_order_in_v=True, l_nodal=False,
                                  algo=0, lvect=64):
        _chk_part(xp, yp, zp, ex, ey, ez, bx, by, bz)
        _chk_grid(exg, eyg, ezg, bxg, byg, bzg)
        ng = _iv(ngx, ngy, ngz)
        nv = _iv(nx + 1, ny + 1, nz + 1)
        lo = _ri(1 if l_lower_order_in_v else 0)
        nod = _ri(1 if l_nodal else 0)
        npart = _ri(len(xp))
        if algo == 0:
            fn = self._f(self.GATHER_DEFAULT[order])
            fn(npart, _ptr(xp), _ptr(yp), _ptr(zp), _ptr(ex), _ptr(ey), _ptr(ez), _ptr(bx),
               _ptr(by), _ptr(bz), _rf(xmin), _rf(ymin), _rf(zmin), _rf(dx), _rf(dy), _rf(dz),
               _ptr(exg), _ptr(ng), _ptr(nv), _ptr(eyg), _ptr(ng), _ptr(nv), _ptr(ezg),
               _ptr(ng), _ptr(nv), _ptr(bxg), _ptr(ng), _ptr(nv), _ptr(byg), _ptr(ng),
               _ptr(nv), _ptr(bzg), _ptr(ng), _ptr(nv), _ri(lvect), lo, nod)
        else:
            o = f"{order}_{order}_{order}"
            fe = self._f(f"gete3d_energy_conserving_scalar_{o}_")
            fb = self._f(f"getb3d_energy_conserving_scalar_{o}_")
            fe(npart, _ptr(xp), _ptr(yp), _ptr(zp), _ptr(ex), _ptr(ey), _ptr(ez), _rf(xmin),
               _rf(ymin), _rf(zmin), _rf(dx), _rf(dy), _rf(dz), _ptr(exg), _ptr(ng), _ptr(nv),
               _ptr(eyg), _ptr(ng), _ptr(nv), _ptr(ezg), _ptr(ng), _ptr(nv), lo, nod)
            fb(npart, _ptr(xp), _ptr(yp), _ptr(zp), _ptr(bx), _ptr(by), _ptr(bz), _rf(xmin),
               _rf(ymin), _rf(zmin), _rf(dx), _rf(dy), _rf(dz), _ptr(bxg), _ptr(ng), _ptr(nv),
               _ptr(byg), _ptr(ng), _ptr(nv), _ptr(bzg), _ptr(ng), _ptr(nv), lo, nod)

    def boris_push_u_3d(self, uxp, uyp, uzp, gaminv, ex, ey, ez, bx, by, bz, q, m, dt):
        _chk_part(uxp, uyp, uzp, gaminv, ex, ey, ez, bx, by, bz)
        self._f("pxr_boris_push_u_3d_")(_ri(len(uxp)), *map(_ptr, (uxp, uyp, uzp, gaminv, ex,
                                                                 ey, ez, bx, by, bz)),
                                        _rf(q), _rf(m), _rf(dt))

    def set_gamma(self, uxp, uyp, uzp, gaminv):
        self._f("pxr_set_gamma_")(_ri(len(uxp)), *map(_ptr, (uxp, uyp, uzp, gaminv)))

    def pushxyz(self, xp, yp, zp, uxp, uyp, uzp, gaminv, dt):
        self._f("pxr_pushxyz_")(_ri(len(xp)), *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv)),
                                _rf(dt))

    def ebcancelpush3d(self, uxp, uyp, uzp, gi, ex, ey, ez, bx, by, bz, q, m, dt, which=0):
        self._f("pxr_ebcancelpush3d_")(_ri(len(uxp)), *map(_ptr, (uxp, uyp, uzp, gi, ex, ey,
                                                                ez, bx, by, bz)),
                                       _rf(q), _rf(m), _rf(dt), _ri(which))

    def depose_jxjyjz(self, jx, jy, jz, xp, yp, zp, uxp, uyp, uzp, gaminv, w, q, xmin, ymin,
                      zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, order, algo=0,
                      l_nodal=False):
        """algo 0: Esirkepov (depose_jxjyjz_esirkepov_n_n_n); algo 3: direct scalar."""
        _chk_grid(jx, jy, jz)
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv, w)
        ng = _iv(ngx, ngy, ngz)
        nv = _iv(nx + 1, ny + 1, nz + 1)
        o = f"{order}_{order}_{order}"
        common = [_ptr(jx), _ptr(ng), _ptr(nv), _ptr(jy), _ptr(ng), _ptr(nv), _ptr(jz),
                  _ptr(ng), _ptr(nv), _ri(len(xp))] + list(map(_ptr, (xp, yp, zp, uxp, uyp,
                                                                     uzp, gaminv, w))) + \
            [_rf(q), _rf(xmin), _rf(ymin), _rf(zmin), _rf(dt), _rf(dx), _rf(dy), _rf(dz)]
        if algo == 0:
            self._f(f"depose_jxjyjz_esirkepov_{o}_")(*common)
        elif algo == 3:
            self._f(f"depose_jxjyjz_scalar_{o}_")(*common, _ri(1 if l_nodal else 0))
        else:
            raise ValueError("algo must be 0 or 3")

    def depose_rho(self, rho, xp, yp, zp, w, q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz, ngx,
                   ngy, ngz, order, scalar=False):
        """pxr_depose_rho_n (charge_deposition_3d.F90:2081), or with scalar=True
        depose_rho_scalar_n_n_n (:77/:174/:324)."""
        _chk_grid(rho)
        _chk_part(xp, yp, zp, w)
        head = [_ptr(rho), _ri(len(xp)), _ptr(xp), _ptr(yp), _ptr(zp), _ptr(w), _rf(q),
                _rf(xmin), _rf(ymin), _rf(zmin), _rf(dx), _rf(dy), _rf(dz), _ri(nx), _ri(ny),
                _ri(nz), _ri(ngx), _ri(ngy), _ri(ngz)]
        if scalar:
            o = f"{order}_{order}_{order}"
            self._f(f"depose_rho_scalar_{o}_")(*head, _ri(64))
        else:
            self._f("pxr_depose_rho_n_")(*head, _ri(order), _ri(order), _ri(order), _ri(1),
                                         _ri(0))

    def depose_rho_n(self, rho, xp, yp, zp, w, q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                     ngx, ngy, ngz, nox, noy, noz, l_particles_weight=True, l4symtry=False):
        """pxr_depose_rho_n with one order per axis."""
        _chk_grid(rho)
        _chk_part(xp, yp, zp)
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._f("pxr_depose_rho_n_")(
            _ptr(rho), _ri(len(xp)), _ptr(xp), _ptr(yp), _ptr(zp), _ptr(w), _rf(q), _rf(xmin),
            _rf(ymin), _rf(zmin), _rf(dx), _rf(dy), _rf(dz), _ri(nx), _ri(ny), _ri(nz),
            _ri(ngx), _ri(ngy), _ri(ngz), _ri(nox), _ri(noy), _ri(noz),
            _ri(1 if l_particles_weight else 0), _ri(1 if l4symtry else 0))

    def depose_esirkepov_n(self, jx, jy, jz, xp, yp, zp, uxp, uyp, uzp, gaminv, w, q, xmin,
                           ymin, zmin, dt, dx, dy, dz, nx, ny, nz, ngx, ngy, ngz, nox, noy, noz):
        """pxr_depose_jxjyjz_esirkepov_n as depose_jxjyjz_generic calls it
        (current_deposition_manager_3d.F90:186: particle weights, no symmetry)."""
        _chk_grid(jx, jy, jz)
        _chk_part(xp, yp, zp, uxp, uyp, uzp, gaminv, w)
        ng = _iv(ngx, ngy, ngz)
        nv = _iv(nx + 1, ny + 1, nz + 1)
        self._f("pxr_depose_jxjyjz_esirkepov_n_")(
            _ptr(jx), _ptr(ng), _ptr(nv), _ptr(jy), _ptr(ng), _ptr(nv), _ptr(jz), _ptr(ng),
            _ptr(nv), _ri(len(xp)), *map(_ptr, (xp, yp, zp, uxp, uyp, uzp, gaminv, w)),
            _rf(q), _rf(xmin), _rf(ymin), _rf(zmin), _rf(dt), _rf(dx), _rf(dy), _rf(dz),
            _ri(nox), _ri(noy), _ri(noz), _ri(1), _ri(0))

    def geteb3d_n(self, xp, yp, zp, ex, ey, ez, bx, by, bz, xmin, ymin, zmin, dx, dy, dz, nx,
                  ny, nz, ngx, ngy, ngz, nox, noy, noz, exg, eyg, ezg, bxg, byg, bzg,
                  l4symtry=False, l_lower_order_in_v=True, l_nodal=False):
        """pxr_gete3d_n_energy_conserving then pxr_getb3d_n_energy_conserving,
        as geteb3d_energy_conserving_generic calls them for field_gathe_algo 2
        or orders other than 1_1_1, 2_2_2, 3_3_3."""
        _chk_part(xp, yp, zp, ex, ey, ez, bx, by, bz)
        _chk_grid(exg, eyg, ezg, bxg, byg, bzg)
        ng = _iv(ngx, ngy, ngz)
        nv = _iv(nx + 1, ny + 1, nz + 1)
        flags = [_ri(1 if l4symtry else 0), _ri(1 if l_lower_order_in_v else 0),
                 _ri(1 if l_nodal else 0)]
        head = [_ri(len(xp)), _ptr(xp), _ptr(yp), _ptr(zp)]
        tail = [_rf(xmin), _rf(ymin), _rf(zmin), _rf(dx), _rf(dy), _rf(dz), _ri(nox), _ri(noy),
                _ri(noz)]
        self._f("pxr_gete3d_n_energy_conserving_")(
            *head, _ptr(ex), _ptr(ey), _ptr(ez), *tail, _ptr(exg), _ptr(ng), _ptr(nv),
            _ptr(eyg), _ptr(ng), _ptr(nv), _ptr(ezg), _ptr(ng), _ptr(nv), *flags)
        self._f("pxr_getb3d_n_energy_conserving_")(
            *head, _ptr(bx), _ptr(by), _ptr(bz), *tail, _ptr(bxg), _ptr(ng), _ptr(nv),
            _ptr(byg), _ptr(ng), _ptr(nv), _ptr(bzg), _ptr(ng), _ptr(nv), *flags)

    def push_em3d_evec(self, lo, hi, ex, ey, ez, bx, by, bz, jx, jy, jz, mudt, dtsdx, dtsdy,
                       dtsdz, nx, ny, nz, ngx, ngy, ngz):
        _chk_grid(ex, ey, ez, bx, by, bz, jx, jy, jz)
        lo_, hi_ = _iv(*lo), _iv(*hi)
        alo, ahi = _iv(-ngx, -ngy, -ngz), _iv(nx + ngx, ny + ngy, nz + ngz)
        args = [_ptr(lo_), _ptr(hi_)] * 3
        for a in (ex, ey, ez, bx, by, bz, jx, jy, jz):
            args += [_ptr(a), _ptr(alo), _ptr(ahi)]
        self._f("pxrpush_em3d_evec_")(*args, _rf(mudt), _rf(dtsdx), _rf(dtsdy), _rf(dtsdz))

    def push_em3d_bvec(self, lo, hi, ex, ey, ez, bx, by, bz, dtsdx, dtsdy, dtsdz, nx, ny, nz,
                       ngx, ngy, ngz):
        _chk_grid(ex, ey, ez, bx, by, bz)
        lo_, hi_ = _iv(*lo), _iv(*hi)
        alo, ahi = _iv(-ngx, -ngy, -ngz), _iv(nx + ngx, ny + ngy, nz + ngz)
        args = [_ptr(lo_), _ptr(hi_)] * 3
        for a in (ex, ey, ez, bx, by, bz):
            args += [_ptr(a), _ptr(alo), _ptr(ahi)]
        self._f("pxrpush_em3d_bvec_")(*args, _rf(dtsdx), _rf(dtsdy), _rf(dtsdz))

    def push_em3d_evec_norder(self, ex, ey, ez, bx, by, bz, jx, jy, jz, mudt, cx, cy, cz, nx,
                              ny, nz, nox, noy, noz, ngx, ngy, ngz, nxs, nys, nzs,
                              l_nodal=False):
        _chk_grid(ex, ey, ez, bx, by, bz, jx, jy, jz)
        cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in (cx, cy, cz))
        self._f("pxrpush_em3d_evec_norder_")(
            *map(_ptr, (ex, ey, ez, bx, by, bz, jx, jy, jz)), _rf(mudt), _ptr(cx), _ptr(cy),
            _ptr(cz), *map(_ri, (nx, ny, nz, nox, noy, noz, ngx, ngy, ngz, nxs, nys, nzs)),
            _ri(1 if l_nodal else 0))

    def push_em3d_bvec_norder(self, ex, ey, ez, bx, by, bz, cx, cy, cz, nx, ny, nz, nox, noy,
                              noz, ngx, ngy, ngz, nxs, nys, nzs, l_nodal=False):
        _chk_grid(ex, ey, ez, bx, by, bz)
        cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in (cx, cy, cz))
        self._f("pxrpush_em3d_bvec_norder_")(
            *map(_ptr, (ex, ey, ez, bx, by, bz)), _ptr(cx), _ptr(cy), _ptr(cz),
            *map(_ri, (nx, ny, nz, nox, noy, noz, ngx, ngy, ngz, nxs, nys, nzs)),
            _ri(1 if l_nodal else 0))

    def fd_weights(self, norder, l_nodal=False):
        c = np.zeros(max(norder // 2, 1))
        self._f("fd_weights_")(_ptr(c), _ri(norder), _ri(1 if l_nodal else 0))
        return c[: norder // 2]

    def _div(self, name, dv, fx, fy, fz, nx, ny, nz, gx, gy, gz, dx, dy, dz):
        _chk_grid(dv, fx, fy, fz)
        self._f(name)(*map(_ptr, (dv, fx, fy, fz)), _ri(nx), _ri(ny), _ri(nz), _ri(gx), _ri(gy),
                      _ri(gz), _rf(dx), _rf(dy), _rf(dz))

    # diags.F90:129 and :170, compiled inside a module (flang's module
    # procedure names)
    def calc_field_div(self, dv, ex, ey, ez, nx, ny, nz, gx, gy, gz, dx, dy, dz):
        self._div("_QMpxr_ref_diagsPcalc_field_div", dv, ex, ey, ez, nx, ny, nz, gx, gy, gz, dx,
                  dy, dz)

    def calc_field_divb(self, dv, bx, by, bz, nx, ny, nz, gx, gy, gz, dx, dy, dz):
        self._div("_QMpxr_ref_diagsPcalc_field_divb", dv, bx, by, bz, nx, ny, nz, gx, gy, gz, dx,
                  dy, dz)

    def laser_profile(self, window, xx, yy, emax1, emax2, pv1, pv2, k0, q_z, tau, t, t_peak,
                      temporal_order, polangle, emax=0.0):
        """gaussian_profile_ / hanning_profile_ point by point: (n, 3)."""
        amp = np.zeros((len(xx), 3))
        a = [_f64(0.0) for _ in range(3)]
        pv1 = np.ascontiguousarray(pv1, dtype=np.float64)
        pv2 = np.ascontiguousarray(pv2, dtype=np.float64)
        qz = np.array([complex(q_z).real, complex(q_z).imag])
        for i in range(len(xx)):
            common = (_rf(xx[i]), _rf(yy[i]), C.byref(a[0]), C.byref(a[1]), C.byref(a[2]),
                      _rf(emax), _rf(emax1), _rf(emax2), _ptr(pv1), _ptr(pv2), _rf(k0),
                      _ptr(qz))
            if window == 0:
                self._f("gaussian_profile_")(*common, _rf(tau), _rf(t), _rf(t_peak),
                                             _ri(temporal_order), _rf(polangle))
            else:
                self._f("hanning_profile_")(*common, _rf(t), _rf(t_peak), _ri(temporal_order),
                                            _rf(polangle))
            amp[i] = [a[0].value, a[1].value, a[2].value]
        return amp

    def particle_bin_sorting(self, xp, yp, zp, ux, uy, uz, gam, pid, xmin2, ymin2, zmin2,
                             xmax2, ymax2, zmax2, dxf, dyf, dzf):
        _chk_part(xp, yp, zp, ux, uy, uz, gam, pid)
        self._f("pxr_particle_bin_sorting_")(
            _ri(len(xp)), *map(_ptr, (xp, yp, zp, ux, uy, uz, gam, pid)), _ri(1),
            *map(_rf, (xmin2, ymin2, zmin2, xmax2, ymax2, zmax2, dxf, dyf, dzf)))


def load_oracle() -> Oracle:
    build_oracle()
    return Oracle()


def load_reference(production: bool = False):
    """The compiled reference kernels, or None where they were not built:
    the golden-vector build (-O2 -ffp-contract=off), or with production
    the build with the reference's own flags (-O3 -fopenmp) that the CPU
    baseline times."""
    path = REF_O3_SO if production else REF_SO
    if not os.path.exists(path):
        return None
    return Reference(path)


class RefTileDriver:
    """The CPU baseline's tile loop (oracle/ref_tile_driver.c): the
    reference's production-built kernels over tiles on OpenMP threads."""

    def __init__(self, path: str = REF_DRIVER_SO):
        self.lib = C.CDLL(path)
        f = self.lib.pxrref_time_tiles
        f.restype = C.c_int
        f.argtypes = [_i64, C.c_double, _i64, _i64, _i64, C.POINTER(C.c_void_p), C.c_void_p,
                      C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                      C.POINTER(C.c_double)]

    def time_tiles(self, nthreads, seconds, nc, ng, fields, pool, q, m, dt):
        """fields: six Fortran-ordered (nc+1+2ng)^3 blocks; pool: float64
        array (nthreads, 2, 8, ntile); q, m: the two species.  Returns
        (particle-steps, wall seconds)."""
        pool = np.ascontiguousarray(pool, dtype=np.float64)
        assert pool.ndim == 4 and pool.shape[:3] == (nthreads, 2, 8)
        for f in fields:
            assert f.flags.f_contiguous and f.dtype == np.float64 and f.shape == (nc + 1 + 2 * ng,) * 3
        fp = (C.c_void_p * 6)(*[f.ctypes.data for f in fields])
        qa = np.ascontiguousarray(q, dtype=np.float64)
        ma = np.ascontiguousarray(m, dtype=np.float64)
        steps, wall = C.c_double(0), C.c_double(0)
        rc = self.lib.pxrref_time_tiles(nthreads, seconds, nc, ng, pool.shape[3], fp,
                                        pool.ctypes.data, qa.ctypes.data, ma.ctypes.data, dt,
                                        C.byref(steps), C.byref(wall))
        if rc != 0:
            raise RuntimeError("pxrref_time_tiles rejected its arguments")
        return steps.value, wall.value


def load_ref_driver():
    """The CPU baseline's OpenMP tile loop, or None where it was not built."""
    if not os.path.exists(REF_DRIVER_SO):
        return None
    return RefTileDriver(REF_DRIVER_SO)
