"""ctypes binding of libpxrhip.so (the C ABI declared in include/pxrhip.h).

The library is loaded from this package directory and nowhere else; there is
no CPU fallback.  Every call goes through ``call()``, which raises
``PxrHipError`` with the library's message on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# PXR_LIB overrides the location of the shared library (a build placed
# elsewhere); the default is the in-tree build next to this file
LIB_PATH = os.environ.get("PXR_LIB") or os.path.join(HERE, "libpxrhip.so")

_i64 = C.c_int64
_f64 = C.c_double
_p = C.c_void_p
_i32 = C.c_int

# name -> argument types (restype is int for every entry point except the two
# accessors set below)
_P = _p
_I = _i64
_F = _f64
SIGNATURES = {
    "pxrhip_init": [_i32],
    "pxrhip_finalize": [],
    "pxrhip_set_stream": [_p],
    "pxrhip_synchronize": [],
    "pxrhip_set_checks": [_i32],
    "pxrhip_version": [],
    "pxrhip_bad_count": [C.POINTER(C.c_longlong), _i32],
    "pxrhip_malloc": [C.POINTER(C.c_void_p), _I],
    "pxrhip_free": [_P],
    "pxrhip_memcpy_htod": [_P, _P, _I],
    "pxrhip_memcpy_dtoh": [_P, _P, _I],
    "pxrhip_memcpy_dtod": [_P, _P, _I],
    "pxrhip_fill_f64": [_P, _I, _F],
    "pxrhip_fill_i64": [_P, _I, _I],
    "pxrhip_convert_f32_f64": [_P, _P, _I],
    "pxrhip_timer_record": [_i32],
    "pxrhip_timer_elapsed": [_i32, _i32, C.POINTER(C.c_double)],
    "pxrhip_kinetic_energy": [_I, _P, _P, C.POINTER(C.c_double)],
    "pxrhip_graph_begin": [],
    "pxrhip_graph_end": [C.POINTER(C.c_void_p)],
    "pxrhip_graph_launch": [_P],
    "pxrhip_graph_destroy": [_P],
    "pxrhip_kinetic_energy_f32": [_I, _P, _P, _P, _P, C.POINTER(C.c_double)],
    "pxrhip_field_energy": [_P] + [_I] * 6 + [C.POINTER(C.c_double)],
    "pxrhip_calc_field_div": [_P] * 4 + [_I] * 6 + [_F] * 3,
    "pxrhip_calc_field_divb": [_P] * 4 + [_I] * 6 + [_F] * 3,
    "pxrhip_get_norm_diverho": [_P, _P] + [_I] * 6 + [C.POINTER(C.c_double)],
    "pxrhip_get_loc_norm_2": [_P] + [_I] * 6 + [C.POINTER(C.c_double)],
    "pxrhip_geteb3d_energy_conserving": [_I] + [_P] * 9 + [_F] * 6 + [_I] * 9 + [_P] * 6 + [_I] * 5,
    "pxrhip_boris_push_u_3d": [_I] + [_P] * 10 + [_F] * 3,
    "pxrhip_ebcancelpush3d": [_I] + [_P] * 10 + [_F] * 3 + [_I],
    "pxrhip_pushxyz": [_I] + [_P] * 7 + [_F],
    "pxrhip_set_gamma": [_I] + [_P] * 4,
    "pxrhip_depose_jxjyjz": [_P] * 3 + [_I] + [_P] * 8 + [_F] * 8 + [_I] * 6 + [_I] * 3 + [_I, _I],
    "pxrhip_depose_rho_n": [_P, _I] + [_P] * 4 + [_F] * 7 + [_I] * 6 + [_I] * 3 + [_I, _I],
    "pxrhip_push_em3d_evec": [_P] * 6 + [_P] * 27 + [_F] * 4,
    "pxrhip_push_em3d_bvec": [_P] * 6 + [_P] * 18 + [_F] * 3,
    "pxrhip_push_em3d_evec_norder": [_P] * 9 + [_F] + [_P] * 3 + [_I] * 13,
    "pxrhip_push_em3d_bvec_norder": [_P] * 6 + [_P] * 3 + [_I] * 13,
    "pxrhip_field_bc_periodic": [_P] + [_I] * 6,
    "pxrhip_current_bc_periodic": [_P] + [_I] * 6,
    "pxrhip_particle_bin_sorting": [_I] + [_P] * 8 + [_I] + [_F] * 9,
    "pxrhip_particle_bin_sorting_noalloc": [_I] + [_P] * 8 + [_I] + [_F] * 9,
    "pxrhip_laserp_pusher_gaussian": [_I, _I] + [_P] * 8 + [_F, _I] + [_F] * 3 + [_P, _P, _F, _P]
                                     + [_F] * 4 + [_I, _F],
    "pxrhip_laserp_pusher_hanning": [_I, _I] + [_P] * 8 + [_F, _I] + [_F] * 3 + [_P, _P, _F, _P]
                                    + [_F] * 3 + [_I, _F],
    "pxrhip_tiles_nbins": [_P, _P],
    "pxrhip_tiles_check": [],
    "pxrhip_tiles_stats": [C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "pxrhip_set_tuning": [C.c_char_p, _I],
    "pxrhip_get_tuning": [C.c_char_p, C.POINTER(C.c_int64)],
    "pxrhip_tiles_gather_push_deposit": [_P, _P] + [_P] * 6 + [_P] * 3 + [_F] + [_I] * 5,
    "pxrhip_tiles_gather_push": [_P, _P] + [_P] * 6 + [_F] + [_I] * 4,
    "pxrhip_tiles_depose_currents": [_P, _P] + [_P] * 3 + [_F] + [_I] * 3,
    "pxrhip_tiles_rebin_sort": [_P, _P, _P],
    "pxrhip_tiles_depose_rho": [_P, _P, _P, _I],
}


class PxrHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pxrhip {code}] {msg}")
        self.code = code


ERR_ARG = -1
ERR_HIP = -2
ERR_BOUNDS = -3
ERR_UNSUPPORTED = -4
ERR_CAPACITY = -5

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it with `python -m picsar_amd.build` "
                "(there is no CPU fallback)")
        lib = C.CDLL(path)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:  # reported by call() if it is ever used
                continue
            fn.argtypes = args
            fn.restype = _i32
        lib.pxrhip_last_error.argtypes = []
        lib.pxrhip_last_error.restype = C.c_char_p
        lib.pxrhip_get_stream.argtypes = []
        lib.pxrhip_get_stream.restype = _p
        _lib = lib
        return lib


def last_error() -> str:
    return load().pxrhip_last_error().decode(errors="replace")


def call(name: str, *args) -> int:
    lib = load()
    fn = getattr(lib, name, None)
    if fn is None:
        raise PxrHipError(ERR_UNSUPPORTED, f"{name} is not exported by {LIB_PATH}")
    rc = fn(*args)
    if rc != 0:
        raise PxrHipError(rc, last_error())
    return rc


def exported_symbols(path: str = LIB_PATH):
    """Names of the pxrhip_* functions the shared object exports (no GPU needed)."""
    lib = C.CDLL(path)
    out = []
    for name in list(SIGNATURES) + ["pxrhip_last_error", "pxrhip_get_stream"]:
        if hasattr(lib, name):
            out.append(name)
    return out
