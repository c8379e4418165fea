"""Drop-in mirror of PICSAR's Python module for the hot path.

PICSAR exposes its Fortran subroutines through Forthon as
``picsar_python.picsarpy.picsar`` (examples/example_scripts_python/test.py:4,11).
This module exposes the same names with the same argument order, backed by
the HIP kernels of libpxrhip.so:

    from picsar_amd.picsarpy import picsar as pxr
    pxr.geteb3d_energy_conserving(np, xp, yp, zp, ex, ey, ez, bx, by, bz,
                                  xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                                  nxguard, nyguard, nzguard, nox, noy, noz,
                                  exg, eyg, ezg, bxg, byg, bzg,
                                  ll4symtry, l_lower_order_in_v, lvect,
                                  field_gathe_algo)

Arrays may be numpy arrays (copied to the GPU for the call and written
back, like Forthon's in-place update), device arrays of the library
(``picsar_amd.devmem.DeviceArray`` / ``DeviceGrid``, used in place) or torch
ROCm tensors (used in place).  Grid arrays are Fortran-ordered with bounds
(-ng:n+ng) on each axis: a numpy array with order='F' and shape
(nx+1+2ng, ny+1+2ng, nz+1+2ng), a DeviceGrid of that shape, or a tensor with
that shape and strides (1, NX, NX*NY) (or any contiguous tensor of that many
elements).

Module variables the reference kernels read (fields%l_nodalgrid) are
attributes of ``picsar``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PxrHipError  # noqa: F401
from .devmem import DeviceArray, DeviceGrid

_i64arr3 = C.c_int64 * 3


class _Staged:
    """A numpy argument validated but not yet copied to the device."""

    def __init__(self, a, order, write):
        self.a, self.order, self.write = a, order, write
        self.d = None


class _Call:
    """Collects the array arguments of one call.  Every argument is checked
    before anything is copied.  numpy arrays are staged in device memory of
    the library and the modified ones written back after the call; device
    arrays of the library (``picsar_amd.devmem``) are used in place.  torch
    tensors are used in place too; since torch runs its own HIP runtime, the
    call is then bracketed by a device synchronisation on both sides."""

    def __init__(self):
        self.staged = []
        self.torch = False

    @staticmethod
    def _is_torch(a):
        mod = type(a).__module__
        return mod.startswith("torch")

    def _tensor_ptr(self, t, need, name, fortran_shape=None):
        import torch
        if t.dtype != torch.float64:
            raise TypeError(f"{name}: expected float64, got {t.dtype}")
        if not t.is_cuda:
            raise TypeError(f"{name}: tensor must live on the GPU")
        if t.numel() < need:
            raise ValueError(f"{name}: {t.numel()} elements, need {need}")
        if fortran_shape is not None and t.dim() == 3:
            ok_f = tuple(t.shape) == tuple(fortran_shape) and t.stride() == (
                1, fortran_shape[0], fortran_shape[0] * fortran_shape[1])
            ok_c = t.is_contiguous() and tuple(t.shape) == tuple(reversed(fortran_shape))
            if not (ok_f or ok_c):
                raise ValueError(f"{name}: grid tensor must be Fortran-ordered "
                                 f"{tuple(fortran_shape)}")
        elif not t.is_contiguous():
            raise ValueError(f"{name}: tensor must be contiguous")
        self.torch = True
        return t.data_ptr()

    def arr(self, a, need, name, write=False, fortran_shape=None):
        if isinstance(a, DeviceArray):
            if a.dtype != np.float64:
                raise TypeError(f"{name}: expected float64")
            if a.n < need:
                raise ValueError(f"{name}: {a.n} elements, need {need}")
            if fortran_shape is not None and isinstance(a, DeviceGrid) and \
                    tuple(a.shape) != tuple(fortran_shape):
                raise ValueError(f"{name}: grid shape {a.shape} != {tuple(fortran_shape)}")
            return a.ptr
        if isinstance(a, np.ndarray):
            if a.dtype != np.float64:
                raise TypeError(f"{name}: expected float64, got {a.dtype}")
            if a.size < need:
                raise ValueError(f"{name}: {a.size} elements, need {need}")
            order = "F" if (fortran_shape is not None and a.ndim == 3) else "C"
            if order == "F" and not a.flags.f_contiguous:
                raise ValueError(f"{name}: grid array must be Fortran-ordered")
            h = _Staged(a, order, write)
            self.staged.append(h)
            return h
        if self._is_torch(a):
            return self._tensor_ptr(a, need, name, fortran_shape)
        raise TypeError(f"{name}: expected a numpy array, a picsar_amd device array or a torch "
                        "tensor")

    def ptrs(self, handles):
        """Copy the staged numpy arrays to the device; device pointers."""
        out = []
        for h in handles:
            if isinstance(h, _Staged):
                if h.d is None:
                    h.d = DeviceArray.from_numpy(h.a.ravel(order="F" if h.order == "F" else "K"))
                out.append(h.d.ptr)
            else:
                out.append(h)
        return out

    def before(self):
        if self.torch:
            import torch
            torch.cuda.synchronize()

    def finish(self):
        if self.torch:
            _lib.call("pxrhip_synchronize")
        for h in self.staged:
            if h.write and h.d is not None:
                v = h.d.numpy()
                if h.order == "F":
                    h.a[...] = v.reshape(h.a.shape, order="F")
                else:
                    h.a.reshape(-1)[...] = v
        self.staged.clear()


def _gshape(nx, ny, nz, gx, gy, gz):
    return (int(nx) + 1 + 2 * int(gx), int(ny) + 1 + 2 * int(gy), int(nz) + 1 + 2 * int(gz))


class _Picsar:
    """The ``picsar`` namespace: subroutines and module variables."""

    def __init__(self):
        self.l_nodalgrid = False  # fields%l_nodalgrid (modules.F90)

    # ------------------------------------------------------------------
    def set_checks(self, on: bool):
        """Leaf calls synchronise and raise on out-of-range particles (default on)."""
        _lib.call("pxrhip_set_checks", 1 if on else 0)

    # ------------------------------------------------------------------
    # field_gathering_manager_3d.F90:295
    def geteb3d_energy_conserving(self, np_, xp, yp, zp, ex, ey, ez, bx, by, bz, xmin, ymin,
                                  zmin, dx, dy, dz, nx, ny, nz, nxguard, nyguard, nzguard, nox,
                                  noy, noz, exg, eyg, ezg, bxg, byg, bzg, ll4symtry=False,
                                  l_lower_order_in_v=True, lvect=8, field_gathe_algo=0):
        n = int(np_)
        if n < 0:
            raise ValueError("np < 0")
        shp = _gshape(nx, ny, nz, nxguard, nyguard, nzguard)
        ng = shp[0] * shp[1] * shp[2]
        c = _Call()
        P = [c.arr(a, n, nm) for a, nm in ((xp, "xp"), (yp, "yp"), (zp, "zp"))]
        O = [c.arr(a, n, nm, write=True) for a, nm in
             ((ex, "ex"), (ey, "ey"), (ez, "ez"), (bx, "bx"), (by, "by"), (bz, "bz"))]
        G = [c.arr(a, ng, nm, fortran_shape=shp) for a, nm in
             ((exg, "exg"), (eyg, "eyg"), (ezg, "ezg"), (bxg, "bxg"), (byg, "byg"),
              (bzg, "bzg"))]
        P, O, G = c.ptrs(P), c.ptrs(O), c.ptrs(G)
        c.before()
        _lib.call("pxrhip_geteb3d_energy_conserving", n, *P, *O, float(xmin), float(ymin),
                  float(zmin), float(dx), float(dy), float(dz), int(nx), int(ny), int(nz),
                  int(nxguard), int(nyguard), int(nzguard), int(nox), int(noy), int(noz),
                  *G, int(bool(ll4symtry)), int(bool(l_lower_order_in_v)), int(lvect),
                  int(field_gathe_algo), int(bool(self.l_nodalgrid)))
        c.finish()

    # boris_3d.F90:51
    def pxr_boris_push_u_3d(self, np_, uxp, uyp, uzp, gaminv, ex, ey, ez, bx, by, bz, q, m, dt):
        n = int(np_)
        c = _Call()
        U = [c.arr(a, n, nm, write=True) for a, nm in
             ((uxp, "uxp"), (uyp, "uyp"), (uzp, "uzp"), (gaminv, "gaminv"))]
        E = [c.arr(a, n, nm) for a, nm in
             ((ex, "ex"), (ey, "ey"), (ez, "ez"), (bx, "bx"), (by, "by"), (bz, "bz"))]
        U, E = c.ptrs(U), c.ptrs(E)
        c.before()
        _lib.call("pxrhip_boris_push_u_3d", n, *U, *E, float(q), float(m), float(dt))
        c.finish()

    # vay_3d.F90:56
    def pxr_ebcancelpush3d(self, np_, uxp, uyp, uzp, gi, ex, ey, ez, bx, by, bz, q, m, dt,
                           which=0):
        n = int(np_)
        c = _Call()
        U = [c.arr(a, n, nm, write=True) for a, nm in
             ((uxp, "uxp"), (uyp, "uyp"), (uzp, "uzp"), (gi, "gi"))]
        E = [c.arr(a, n, nm) for a, nm in
             ((ex, "ex"), (ey, "ey"), (ez, "ez"), (bx, "bx"), (by, "by"), (bz, "bz"))]
        U, E = c.ptrs(U), c.ptrs(E)
        c.before()
        _lib.call("pxrhip_ebcancelpush3d", n, *U, *E, float(q), float(m), float(dt), int(which))
        c.finish()

    # boris_3d.F90:849
    def pxr_pushxyz(self, np_, xp, yp, zp, uxp, uyp, uzp, gaminv, dt):
        n = int(np_)
        c = _Call()
        X = [c.arr(a, n, nm, write=True) for a, nm in ((xp, "xp"), (yp, "yp"), (zp, "zp"))]
        U = [c.arr(a, n, nm) for a, nm in
             ((uxp, "uxp"), (uyp, "uyp"), (uzp, "uzp"), (gaminv, "gaminv"))]
        X, U = c.ptrs(X), c.ptrs(U)
        c.before()
        _lib.call("pxrhip_pushxyz", n, *X, *U, float(dt))
        c.finish()

    # laser_pusher_manager_3d.F90:195 and :281
    def _laser(self, entry, np_, npidd, pid, xp, yp, zp, uxp, uyp, uzp, gaminv, dtt, lvect,
               emax, emax1, emax2, polvector1, polvector2, k0_laser, q_z, charge, tail):
        n, npid = int(np_), int(npidd)
        if isinstance(pid, np.ndarray) and pid.ndim == 2 and not pid.flags.f_contiguous:
            raise ValueError("pid: expected a Fortran-ordered (np, npid) array")
        c = _Call()
        P = c.arr(pid, n * npid, "pid")
        X = [c.arr(a, n, nm, write=True) for a, nm in
             ((xp, "xp"), (yp, "yp"), (zp, "zp"), (uxp, "uxp"), (uyp, "uyp"), (uzp, "uzp"),
              (gaminv, "gaminv"))]
        (P,), X = c.ptrs([P]), c.ptrs(X)
        pv1 = np.ascontiguousarray(polvector1, dtype=np.float64).reshape(3)
        pv2 = np.ascontiguousarray(polvector2, dtype=np.float64).reshape(3)
        qz = np.array([complex(q_z).real, complex(q_z).imag])
        c.before()
        _lib.call(entry, n, npid, P, *X, float(dtt), int(lvect), float(emax), float(emax1),
                  float(emax2), pv1.ctypes.data, pv2.ctypes.data, float(k0_laser),
                  qz.ctypes.data, float(charge), *tail)
        c.finish()

    def laserp_pusher_gaussian(self, np_, npidd, pid, xp, yp, zp, uxp, uyp, uzp, gaminv, dtt,
                               lvect, emax, emax1, emax2, polvector1, polvector2, k0_laser, q_z,
                               charge, laser_tau, real_time, t_peak, temporal_order, polangle):
        self._laser("pxrhip_laserp_pusher_gaussian", np_, npidd, pid, xp, yp, zp, uxp, uyp,
                    uzp, gaminv, dtt, lvect, emax, emax1, emax2, polvector1, polvector2,
                    k0_laser, q_z, charge,
                    (float(laser_tau), float(real_time), float(t_peak), int(temporal_order),
                     float(polangle)))

    def laserp_pusher_hanning(self, np_, npidd, pid, xp, yp, zp, uxp, uyp, uzp, gaminv, dtt,
                              lvect, emax, emax1, emax2, polvector1, polvector2, k0_laser, q_z,
                              charge, real_time, t_peak, temporal_order, polangle):
        self._laser("pxrhip_laserp_pusher_hanning", np_, npidd, pid, xp, yp, zp, uxp, uyp,
                    uzp, gaminv, dtt, lvect, emax, emax1, emax2, polvector1, polvector2,
                    k0_laser, q_z, charge,
                    (float(real_time), float(t_peak), int(temporal_order), float(polangle)))

    # boris_3d.F90:1059
    def pxr_set_gamma(self, np_, uxp, uyp, uzp, gaminv):
        n = int(np_)
        c = _Call()
        U = [c.arr(a, n, nm) for a, nm in ((uxp, "uxp"), (uyp, "uyp"), (uzp, "uzp"))]
        g = c.arr(gaminv, n, "gaminv", write=True)
        U, (g,) = c.ptrs(U), c.ptrs([g])
        c.before()
        _lib.call("pxrhip_set_gamma", n, *U, g)
        c.finish()

    # current_deposition_manager_3d.F90:47
    def depose_jxjyjz(self, jx, jy, jz, np_, xp, yp, zp, uxp, uyp, uzp, gaminv, w, q, xmin, ymin,
                      zmin, dt, dx, dy, dz, nx, ny, nz, nxguard, nyguard, nzguard, nox, noy, noz,
                      current_depo_algo=0):
        n = int(np_)
        shp = _gshape(nx, ny, nz, nxguard, nyguard, nzguard)
        ng = shp[0] * shp[1] * shp[2]
        c = _Call()
        J = [c.arr(a, ng, nm, write=True, fortran_shape=shp) for a, nm in
             ((jx, "jx"), (jy, "jy"), (jz, "jz"))]
        P = [c.arr(a, n, nm) for a, nm in
             ((xp, "xp"), (yp, "yp"), (zp, "zp"), (uxp, "uxp"), (uyp, "uyp"), (uzp, "uzp"),
              (gaminv, "gaminv"), (w, "w"))]
        J, P = c.ptrs(J), c.ptrs(P)
        c.before()
        _lib.call("pxrhip_depose_jxjyjz", *J, n, *P, float(q), float(xmin), float(ymin),
                  float(zmin), float(dt), float(dx), float(dy), float(dz), int(nx), int(ny),
                  int(nz), int(nxguard), int(nyguard), int(nzguard), int(nox), int(noy), int(noz),
                  int(current_depo_algo), int(bool(self.l_nodalgrid)))
        c.finish()

    # charge_deposition_3d.F90:2081
    def pxr_depose_rho_n(self, rho, np_, xp, yp, zp, w, q, xmin, ymin, zmin, dx, dy, dz, nx, ny,
                         nz, nxguard, nyguard, nzguard, nox, noy, noz, l_particles_weight=True,
                         l4symtry=False):
        n = int(np_)
        shp = _gshape(nx, ny, nz, nxguard, nyguard, nzguard)
        c = _Call()
        R = [c.arr(rho, shp[0] * shp[1] * shp[2], "rho", write=True, fortran_shape=shp)]
        P = [c.arr(a, n, nm) for a, nm in ((xp, "xp"), (yp, "yp"), (zp, "zp"))]
        P.append(c.arr(w, n if l_particles_weight else 1, "w"))
        R, P = c.ptrs(R), c.ptrs(P)
        c.before()
        _lib.call("pxrhip_depose_rho_n", R[0], n, *P, float(q), float(xmin), float(ymin),
                  float(zmin), float(dx), float(dy), float(dz), int(nx), int(ny), int(nz),
                  int(nxguard), int(nyguard), int(nzguard), int(nox), int(noy), int(noz),
                  int(bool(l_particles_weight)), int(bool(l4symtry)))
        c.finish()

    # yee.F90:637
    def pxrpush_em3d_evec(self, xlo, xhi, ylo, yhi, zlo, zhi, ex, exlo, exhi, ey, eylo, eyhi, ez,
                          ezlo, ezhi, bx, bxlo, bxhi, by, bylo, byhi, bz, bzlo, bzhi, jx, jxlo,
                          jxhi, jy, jylo, jyhi, jz, jzlo, jzhi, mudt, dtsdx, dtsdy, dtsdz):
        c = _Call()
        B = [_i64arr3(*map(int, v)) for v in (xlo, xhi, ylo, yhi, zlo, zhi)]
        args = []
        for a, lo, hi, nm, wr in ((ex, exlo, exhi, "ex", True), (ey, eylo, eyhi, "ey", True),
                                  (ez, ezlo, ezhi, "ez", True), (bx, bxlo, bxhi, "bx", False),
                                  (by, bylo, byhi, "by", False), (bz, bzlo, bzhi, "bz", False),
                                  (jx, jxlo, jxhi, "jx", False), (jy, jylo, jyhi, "jy", False),
                                  (jz, jzlo, jzhi, "jz", False)):
            shp = tuple(int(h) - int(l) + 1 for l, h in zip(lo, hi))
            lo_, hi_ = _i64arr3(*map(int, lo)), _i64arr3(*map(int, hi))
            B += [lo_, hi_]
            args += [c.arr(a, shp[0] * shp[1] * shp[2], nm, write=wr, fortran_shape=shp), lo_, hi_]
        args = c.ptrs(args)
        c.before()
        _lib.call("pxrhip_push_em3d_evec", *B[:6], *args, float(mudt), float(dtsdx),
                  float(dtsdy), float(dtsdz))
        c.finish()

    # yee.F90:1258
    def pxrpush_em3d_bvec(self, xlo, xhi, ylo, yhi, zlo, zhi, ex, exlo, exhi, ey, eylo, eyhi, ez,
                          ezlo, ezhi, bx, bxlo, bxhi, by, bylo, byhi, bz, bzlo, bzhi, dtsdx, dtsdy,
                          dtsdz):
        c = _Call()
        B = [_i64arr3(*map(int, v)) for v in (xlo, xhi, ylo, yhi, zlo, zhi)]
        args = []
        for a, lo, hi, nm, wr in ((ex, exlo, exhi, "ex", False), (ey, eylo, eyhi, "ey", False),
                                  (ez, ezlo, ezhi, "ez", False), (bx, bxlo, bxhi, "bx", True),
                                  (by, bylo, byhi, "by", True), (bz, bzlo, bzhi, "bz", True)):
            shp = tuple(int(h) - int(l) + 1 for l, h in zip(lo, hi))
            lo_, hi_ = _i64arr3(*map(int, lo)), _i64arr3(*map(int, hi))
            B += [lo_, hi_]
            args += [c.arr(a, shp[0] * shp[1] * shp[2], nm, write=wr, fortran_shape=shp), lo_, hi_]
        args = c.ptrs(args)
        c.before()
        _lib.call("pxrhip_push_em3d_bvec", *B[:6], *args, float(dtsdx), float(dtsdy),
                  float(dtsdz))
        c.finish()

    @staticmethod
    def _coefs(v, norder):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).ravel())
        if a.size < int(norder) // 2:
            raise ValueError("stencil coefficient array shorter than norder/2")
        return a

    # yee.F90:46
    def pxrpush_em3d_evec_norder(self, ex, ey, ez, bx, by, bz, jx, jy, jz, mudt, dtsdx, dtsdy,
                                 dtsdz, nx, ny, nz, norderx, nordery, norderz, nxguard, nyguard,
                                 nzguard, nxs, nys, nzs, l_nodalgrid=None):
        shp = _gshape(nx, ny, nz, nxguard, nyguard, nzguard)
        ng = shp[0] * shp[1] * shp[2]
        c = _Call()
        F = [c.arr(a, ng, nm, write=(i < 3), fortran_shape=shp) for i, (a, nm) in enumerate(
            ((ex, "ex"), (ey, "ey"), (ez, "ez"), (bx, "bx"), (by, "by"), (bz, "bz"), (jx, "jx"),
             (jy, "jy"), (jz, "jz")))]
        cx, cy, cz = (self._coefs(v, o) for v, o in ((dtsdx, norderx), (dtsdy, nordery),
                                                      (dtsdz, norderz)))
        nod = self.l_nodalgrid if l_nodalgrid is None else l_nodalgrid
        F = c.ptrs(F)
        c.before()
        _lib.call("pxrhip_push_em3d_evec_norder", *F, float(mudt), cx.ctypes.data, cy.ctypes.data,
                  cz.ctypes.data, int(nx), int(ny), int(nz), int(norderx), int(nordery),
                  int(norderz), int(nxguard), int(nyguard), int(nzguard), int(nxs), int(nys),
                  int(nzs), int(bool(nod)))
        c.finish()

    # yee.F90:752
    def pxrpush_em3d_bvec_norder(self, ex, ey, ez, bx, by, bz, dtsdx, dtsdy, dtsdz, nx, ny, nz,
                                 norderx, nordery, norderz, nxguard, nyguard, nzguard, nxs, nys,
                                 nzs, l_nodalgrid=None):
        shp = _gshape(nx, ny, nz, nxguard, nyguard, nzguard)
        ng = shp[0] * shp[1] * shp[2]
        c = _Call()
        F = [c.arr(a, ng, nm, write=(i >= 3), fortran_shape=shp) for i, (a, nm) in enumerate(
            ((ex, "ex"), (ey, "ey"), (ez, "ez"), (bx, "bx"), (by, "by"), (bz, "bz")))]
        cx, cy, cz = (self._coefs(v, o) for v, o in ((dtsdx, norderx), (dtsdy, nordery),
                                                      (dtsdz, norderz)))
        nod = self.l_nodalgrid if l_nodalgrid is None else l_nodalgrid
        F = c.ptrs(F)
        c.before()
        _lib.call("pxrhip_push_em3d_bvec_norder", *F, cx.ctypes.data, cy.ctypes.data,
                  cz.ctypes.data, int(nx), int(ny), int(nz), int(norderx), int(nordery),
                  int(norderz), int(nxguard), int(nyguard), int(nzguard), int(nxs), int(nys),
                  int(nzs), int(bool(nod)))
        c.finish()

    # sorting.F90:247
    def pxr_particle_bin_sorting(self, np2, xp, yp, zp, ux, uy, uz, gam, pid, wpid, xmin2, ymin2,
                                 zmin2, xmax2, ymax2, zmax2, dxf, dyf, dzf):
        n = int(np2)
        c = _Call()
        A = [c.arr(a, n, nm, write=True) for a, nm in
             ((xp, "xp"), (yp, "yp"), (zp, "zp"), (ux, "ux"), (uy, "uy"), (uz, "uz"),
              (gam, "gam"), (pid, "pid"))]
        A = c.ptrs(A)
        c.before()
        _lib.call("pxrhip_particle_bin_sorting", n, *A, int(wpid), float(xmin2), float(ymin2),
                  float(zmin2), float(xmax2), float(ymax2), float(zmax2), float(dxf), float(dyf),
                  float(dzf))
        c.finish()

    # field_boundaries.F90:81 (field_bc on one periodic rank)
    def field_bc(self, field, nxg, nyg, nzg, nx, ny, nz):
        shp = _gshape(nx, ny, nz, nxg, nyg, nzg)
        c = _Call()
        p = c.arr(field, shp[0] * shp[1] * shp[2], "field", write=True, fortran_shape=shp)
        (p,) = c.ptrs([p])
        c.before()
        _lib.call("pxrhip_field_bc_periodic", p, int(nx), int(ny), int(nz), int(nxg), int(nyg),
                  int(nzg))
        c.finish()

    # field_boundaries.F90:403 (summation_bcs on one periodic rank)
    def summation_bcs(self, array, nxg, nyg, nzg, nx, ny, nz):
        shp = _gshape(nx, ny, nz, nxg, nyg, nzg)
        c = _Call()
        p = c.arr(array, shp[0] * shp[1] * shp[2], "array", write=True, fortran_shape=shp)
        (p,) = c.ptrs([p])
        c.before()
        _lib.call("pxrhip_current_bc_periodic", p, int(nx), int(ny), int(nz), int(nxg), int(nyg),
                  int(nzg))
        c.finish()

    def _div(self, entry, out, fx, fy, fz, nx, ny, nz, nxguard, nyguard, nzguard, dx, dy, dz,
             names):
        shp = _gshape(nx, ny, nz, nxguard, nyguard, nzguard)
        tot = shp[0] * shp[1] * shp[2]
        c = _Call()
        A = [c.arr(out, tot, names[0], write=True, fortran_shape=shp)]
        A += [c.arr(a, tot, nm, fortran_shape=shp) for a, nm in zip((fx, fy, fz), names[1:])]
        A = c.ptrs(A)
        c.before()
        _lib.call(entry, *A, int(nx), int(ny), int(nz), int(nxguard), int(nyguard), int(nzguard),
                  float(dx), float(dy), float(dz))
        c.finish()

    # diags.F90:129
    def calc_field_div(self, divee, eex, eey, eez, nx, ny, nz, nxguard, nyguard, nzguard, dx, dy,
                       dz):
        self._div("pxrhip_calc_field_div", divee, eex, eey, eez, nx, ny, nz, nxguard, nyguard,
                  nzguard, dx, dy, dz, ("divee", "eex", "eey", "eez"))

    # diags.F90:170
    def calc_field_divB(self, divbb, bbx, bby, bbz, nx, ny, nz, nxguard, nyguard, nzguard, dx, dy,
                        dz):
        self._div("pxrhip_calc_field_divb", divbb, bbx, bby, bbz, nx, ny, nz, nxguard, nyguard,
                  nzguard, dx, dy, dz, ("divbb", "bbx", "bby", "bbz"))

    def _norm(self, entry, arrays, names, nx2, ny2, nz2, nxguard, nyguard, nzguard):
        shp = _gshape(nx2, ny2, nz2, nxguard, nyguard, nzguard)
        tot = shp[0] * shp[1] * shp[2]
        c = _Call()
        A = c.ptrs([c.arr(a, tot, nm, fortran_shape=shp) for a, nm in zip(arrays, names)])
        out = C.c_double(0.0)
        c.before()
        _lib.call(entry, *A, int(nx2), int(ny2), int(nz2), int(nxguard), int(nyguard),
                  int(nzguard), C.byref(out))
        c.finish()
        return out.value

    # diags.F90:1305 (the norm, intent(out) in the reference, is returned)
    def get_norm_divErho(self, divee2, rho2, nx2, ny2, nz2, nxguard, nyguard, nzguard):
        return self._norm("pxrhip_get_norm_diverho", (divee2, rho2), ("divee2", "rho2"), nx2, ny2,
                          nz2, nxguard, nyguard, nzguard)

    # diags.F90:1265 (returned, not square-rooted)
    def get_loc_norm_2(self, array, nx2, ny2, nz2, nxguard, nyguard, nzguard):
        return self._norm("pxrhip_get_loc_norm_2", (array,), ("array",), nx2, ny2, nz2, nxguard,
                          nyguard, nzguard)


picsar = _Picsar()
