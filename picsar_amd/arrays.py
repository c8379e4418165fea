"""Device array helpers for the reference's layouts.

A PICSAR grid array is Fortran-ordered with bounds (-ng:n+ng) per axis.  On
the device it is a float64 tensor of shape (NX, NY, NZ) with strides
(1, NX, NX*NY): element (i, j, k) of the reference sits at index
[i+ngx, j+ngy, k+ngz].
"""
from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


def grid_shape(nx, ny, nz, ngx, ngy=None, ngz=None):
    ngy = ngx if ngy is None else ngy
    ngz = ngx if ngz is None else ngz
    return (int(nx) + 1 + 2 * int(ngx), int(ny) + 1 + 2 * int(ngy), int(nz) + 1 + 2 * int(ngz))


def grid(nx, ny, nz, ngx, ngy=None, ngz=None, fill=0.0, device="cuda"):
    """A zero (or constant) Fortran-ordered float64 grid tensor."""
    torch = _torch()
    NX, NY, NZ = grid_shape(nx, ny, nz, ngx, ngy, ngz)
    t = torch.full((NZ, NY, NX), float(fill), dtype=torch.float64, device=device)
    return t.permute(2, 1, 0)


def from_numpy_grid(a: np.ndarray, device="cuda"):
    """Fortran-ordered numpy grid -> device tensor with the same layout."""
    torch = _torch()
    if a.ndim != 3:
        raise ValueError("expected a 3-D grid array")
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).to(device)
    return t.permute(2, 1, 0)


def to_numpy_grid(t) -> np.ndarray:
    """Device grid tensor -> Fortran-ordered numpy array."""
    return np.asfortranarray(t.permute(2, 1, 0).cpu().numpy().T)
