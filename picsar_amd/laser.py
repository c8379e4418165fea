"""Laser antenna: a plane of virtual particles whose momenta are set from a
laser profile every step and which deposit current, launching the laser
(laser_pusher_manager_3d.F90; antenna set-up tiling.F90:1654-1919).

The particles are built as ``load_laser_species`` builds them
(tiling.F90:1737-1919): one particle per cell of the plane through ``spot``
normal to ``vector``, at cell centres of the two transverse axes, with the
weight ``eps0 * Emax / 0.01 * min(sx) * min(sy)`` and the antenna-frame
coordinates ``(pos - spot) . polvector1`` and ``(pos - spot) . polvector2`` in
pid columns 2 and 3.  The field amplitudes ``emax1``/``emax2`` are given
directly, as the reference's Python entry ``init_laser_species_python``
(tiling.F90:1654) takes them; ``emax_from_a0`` converts a normalised
amplitude a0 in the unit system of the simulation (the reference's
``load_laser_species`` uses the SI electron mass and charge even in its
normalised-units build).

Each step (submain.F90:116, before the particle boundaries and the
deposition) the antenna particles are pushed by ``laserp_pusher_gaussian`` or
``laserp_pusher_hanning`` at ``real_time = it*dt - dt/2`` and their current is
deposited with the simulation's shape order and deposition algorithm.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from . import _lib
from .devmem import DeviceArray

EPS0 = 1.0  # normalised units


def emax_from_a0(a0: float, lambda_laser: float, m: float = 1.0, q: float = 1.0,
                 clight: float = 1.0) -> float:
    """E = a0 * c^2 * m * k0 / q (tiling.F90:1785)."""
    return a0 * clight ** 2 * m * (2.0 * math.pi / lambda_laser) / q


def antenna_plane(xmin, d, n, spot, vec, pv1, pv2, emax):
    """Antenna particles of load_laser_species (tiling.F90:1806-1893): one
    per cell of the plane normal to vec through spot, at the cell centres of
    the two transverse axes, kept where the normal coordinate lies in the box;
    weight eps0 * emax / 0.01 * min(sx) * min(sy); antenna-frame coordinates
    (pos - spot) . pv1 and (pos - spot) . pv2.  Returns pos (n, 3), w, xx, yy."""
    vec, pv1, pv2 = (np.asarray(v, dtype=np.float64) for v in (vec, pv1, pv2))
    spot = np.asarray(spot, dtype=np.float64)
    mins = np.asarray(xmin, dtype=np.float64)
    dst = np.asarray(d, dtype=np.float64)
    maxs = mins + np.asarray(n, dtype=np.float64) * dst
    intercept = vec[0] * spot[0] + vec[1] * spot[1] + vec[2] * spot[2]
    if vec[2] != 0.0:
        i1, i2, inonz = 0, 1, 2
    elif vec[1] != 0.0:
        i1, i2, inonz = 0, 2, 1
    else:
        i1, i2, inonz = 1, 2, 0
    jmax = int(np.rint((maxs[i1] - mins[i1]) / dst[i1]))
    lmax = int(np.rint((maxs[i2] - mins[i2]) / dst[i2]))
    sx = np.ones(3)
    sy = np.ones(3)
    for k in range(3):
        if pv1[k] != 0.0:
            sx[k] = abs(dst[k] / pv1[k])
        if pv2[k] != 0.0:
            sy[k] = abs(dst[k] / pv2[k])
    weight = EPS0 * emax / 0.01 * sx.min() * sy.min()
    a2 = (mins[i2] + np.arange(lmax) * dst[i2]) + dst[i2] / 2.0
    a1 = (mins[i1] + np.arange(jmax) * dst[i1]) + dst[i1] / 2.0
    P2, P1 = np.meshgrid(a2, a1, indexing="ij")  # l outer, j inner
    pos = np.zeros((P1.size, 3))
    pos[:, i1] = P1.ravel()
    pos[:, i2] = P2.ravel()
    pos[:, inonz] = (intercept - vec[i1] * pos[:, i1] - vec[i2] * pos[:, i2]) / vec[inonz]
    keep = (pos[:, inonz] >= mins[inonz]) & (pos[:, inonz] < maxs[inonz])
    pos = pos[keep]
    dp = pos - spot
    xx = (dp[:, 0] * pv1[0] + dp[:, 1] * pv1[1]) + dp[:, 2] * pv1[2]
    yy = (dp[:, 0] * pv2[0] + dp[:, 1] * pv2[1]) + dp[:, 2] * pv2[2]
    return pos, np.full(len(pos), weight), xx, yy


class LaserAntenna:
    def __init__(self, sim, spot: Sequence[float], vector: Sequence[float],
                 polvector1: Sequence[float], lambda_laser: float, laser_w0: float,
                 laser_ctau: float, t_peak: float, emax1: float, emax2: float = 0.0,
                 temporal_order: int = 2, time_window: int = 0, polangle: float = 0.0,
                 charge: float = 1.0):
        vec = np.asarray(vector, dtype=np.float64)
        pv1 = np.asarray(polvector1, dtype=np.float64)
        if not (vec ** 2).sum() > 0:
            raise ValueError("laser wave vector is null")
        vec = vec / math.sqrt((vec ** 2).sum())
        pv1 = pv1 / math.sqrt((pv1 ** 2).sum())
        if (pv1 * vec).sum() != 0.0:
            raise ValueError("laser vector and polvector1 are not orthogonal")
        pv2 = np.array([vec[1] * pv1[2] - vec[2] * pv1[1],
                        vec[2] * pv1[0] - vec[0] * pv1[2],
                        vec[0] * pv1[1] - vec[1] * pv1[0]])
        self.vector, self.pv1, self.pv2 = vec, pv1, pv2
        self.spot = np.asarray(spot, dtype=np.float64)
        self.lambda_laser = float(lambda_laser)
        self.k0 = 2.0 * math.pi / self.lambda_laser
        self.laser_tau = float(laser_ctau) / 1.0
        self.t_peak = float(t_peak)
        self.emax1, self.emax2 = float(emax1), float(emax2)
        self.emax = math.sqrt(((self.emax1 * pv1 + self.emax2 * pv2) ** 2).sum())
        if self.emax == 0.0:
            raise ValueError("laser amplitude is zero")
        self.q_z = 1j * float(laser_w0) ** 2 * math.pi / self.lambda_laser
        self.temporal_order = int(temporal_order)
        self.time_window = int(time_window)
        self.polangle = float(polangle)
        self.charge = float(charge)
        pos, w, xx, yy = antenna_plane(
            (sim.xmin, sim.ymin, sim.zmin), (sim.dx, sim.dy, sim.dz), (sim.nx, sim.ny, sim.nz),
            self.spot, vec, pv1, pv2, self.emax)
        self.npart = len(w)
        n = self.npart
        self.x, self.y, self.z = (DeviceArray.from_numpy(np.ascontiguousarray(pos[:, d]))
                                  for d in range(3))
        self.ux, self.uy, self.uz = (DeviceArray.from_numpy(np.zeros(n)) for _ in range(3))
        self.gaminv = DeviceArray.from_numpy(np.ones(n))
        # pid (np x 3, column-major): weight, antenna-frame x, antenna-frame y
        self.pid = DeviceArray.from_numpy(np.concatenate([w, xx, yy]))

    def push(self, sim):
        """push_laser_particles (laser_pusher_manager_3d.F90:15-91) at
        real_time = it*dt - dt/2."""
        if self.npart == 0:
            return
        t = sim.it * sim.dt - sim.dt / 2.0
        pv1 = np.ascontiguousarray(self.pv1)
        pv2 = np.ascontiguousarray(self.pv2)
        qz = np.array([self.q_z.real, self.q_z.imag])
        common = (self.npart, 3, self.pid.ptr, self.x.ptr, self.y.ptr, self.z.ptr, self.ux.ptr,
                  self.uy.ptr, self.uz.ptr, self.gaminv.ptr, sim.dt, 8, self.emax, self.emax1,
                  self.emax2, pv1.ctypes.data, pv2.ctypes.data, self.k0, qz.ctypes.data,
                  self.charge)
        if self.time_window == 0:
            _lib.call("pxrhip_laserp_pusher_gaussian", *common, self.laser_tau, t, self.t_peak,
                      self.temporal_order, self.polangle)
        else:
            _lib.call("pxrhip_laserp_pusher_hanning", *common, t, self.t_peak,
                      self.temporal_order, self.polangle)

    def deposit(self, sim):
        """The antenna's share of pxrdepose_currents_on_grid_jxjyjz: its
        current, with the simulation's shape order and algorithm, added to J."""
        if self.npart == 0:
            return
        f = sim.fields
        g = sim.ng
        _lib.call("pxrhip_depose_jxjyjz", f["jx"].ptr, f["jy"].ptr, f["jz"].ptr, self.npart,
                  self.x.ptr, self.y.ptr, self.z.ptr, self.ux.ptr, self.uy.ptr, self.uz.ptr,
                  self.gaminv.ptr, self.pid.ptr, self.charge, sim.xmin, sim.ymin, sim.zmin,
                  sim.dt, sim.dx, sim.dy, sim.dz, sim.nx, sim.ny, sim.nz, g, g, g, sim.order,
                  sim.order, sim.order, sim.depo, 0)

    def arrays(self):
        n = self.npart
        out = {k: getattr(self, k).numpy(n) for k in ("x", "y", "z", "ux", "uy", "uz", "gaminv")}
        p = self.pid.numpy(3 * n)
        out["w"], out["xx"], out["yy"] = p[:n], p[n:2 * n], p[2 * n:]
        return out
