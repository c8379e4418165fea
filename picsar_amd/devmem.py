"""Device arrays owned by libpxrhip.so (its own HIP runtime and stream).

``DeviceArray`` is a flat float64 or int64 buffer; ``DeviceGrid`` is a
PICSAR grid array: Fortran order, bounds (-ng:n+ng) per axis, so element
(i, j, k) of the reference is at [i+ngx, j+ngy, k+ngz] of
``grid.numpy()``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_DT = {np.dtype(np.float64): 8, np.dtype(np.int64): 8, np.dtype(np.float32): 4}


class DeviceArray:
    """A device buffer of ``n`` float64, int64 or float32 values."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DT:
            raise TypeError("DeviceArray holds float64, int64 or float32")
        self.itemsize = _DT[self.dtype]
        self.nbytes = self.n * self.itemsize
        p = C.c_void_p(0)
        _lib.call("pxrhip_malloc", C.byref(p), max(self.nbytes, 8))
        self._ptr = p.value

    @property
    def ptr(self) -> int:
        return self._ptr

    def __len__(self):
        return self.n

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p:
            try:
                _lib.call("pxrhip_free", C.c_void_p(p))
            except Exception:
                pass
            self._ptr = None

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype or np.asarray(a).dtype).ravel())
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a, offset: int = 0):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype).ravel())
        if offset < 0 or offset + a.size > self.n:
            raise ValueError("upload past the end of the device array")
        _lib.call("pxrhip_memcpy_htod", C.c_void_p(self._ptr + self.itemsize * offset),
                  a.ctypes.data_as(C.c_void_p), a.nbytes)

    def numpy(self, count: int = None) -> np.ndarray:
        n = self.n if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        _lib.call("pxrhip_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self._ptr),
                  n * self.itemsize)
        return out

    def fill(self, value):
        if self.dtype == np.float64:
            _lib.call("pxrhip_fill_f64", C.c_void_p(self._ptr), self.n, float(value))
        elif self.dtype == np.int64:
            _lib.call("pxrhip_fill_i64", C.c_void_p(self._ptr), self.n, int(value))
        else:
            self.upload(np.full(self.n, value, dtype=self.dtype))

    def copy_from(self, other: "DeviceArray", count: int = None):
        if other.dtype != self.dtype:
            raise TypeError("copy between device arrays of different dtypes")
        n = min(self.n, other.n) if count is None else int(count)
        _lib.call("pxrhip_memcpy_dtod", C.c_void_p(self._ptr), C.c_void_p(other.ptr),
                  n * self.itemsize)


class DeviceGrid(DeviceArray):
    """A Fortran-ordered float64 grid array with guards."""

    def __init__(self, nx, ny, nz, ngx, ngy=None, ngz=None):
        ngy = ngx if ngy is None else ngy
        ngz = ngx if ngz is None else ngz
        self.shape = (int(nx) + 1 + 2 * int(ngx), int(ny) + 1 + 2 * int(ngy),
                      int(nz) + 1 + 2 * int(ngz))
        super().__init__(self.shape[0] * self.shape[1] * self.shape[2], np.float64)
        self.fill(0.0)

    def upload(self, a, offset: int = 0):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 3:
            if a.shape != self.shape:
                raise ValueError(f"grid shape {a.shape} != {self.shape}")
            a = a.ravel(order="F")
        super().upload(a, offset)

    def numpy(self, count: int = None) -> np.ndarray:
        return super().numpy().reshape(self.shape, order="F")


def synchronize():
    _lib.call("pxrhip_synchronize")
