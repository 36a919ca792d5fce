"""ctypes binding of libawp_hip.so (the C ABI declared in include/awp_hip.h).

The library is loaded from the package directory.  There is no fallback:
if the shared library is missing, or no GPU is visible when a launcher is
called, the call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from .refstructs import Grid3

LIB_PATH = Path(__file__).resolve().parent / "libawp_hip.so"

_fp = C.c_void_p  # device pointers travel as plain addresses
_i = C.c_int
_f = C.c_float
_s = C.c_void_p  # hipStream_t


# name -> argtypes, in header order
SIGNATURES: dict[str, list] = {
    "awp_last_error": [],
    "awp_last_error_string": [],
    "awp_clear_error": [],
    "awp_set_tuning": [C.c_char_p, _i],
    "SetDeviceConstValue": [C.POINTER(_f), _f, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _i,
                            _f, _f, C.POINTER(_f), C.POINTER(_f)],
    "dvelcx_H_opt": [_fp] * 13 + [_i, _i, _s, _i, _i, _i, _i],
    "dstrqc_H": [_fp] * 23 + [_i, _i, _s] + [_fp] * 5 + [_i] * 9,
    "dstrqc_H_new": [_fp] * 23 + [_i, _i, _s] + [_fp] * 5 + [_i] * 9,
    "fstr_H": [_fp] * 3 + [_s, _i, _i, _i, _i],
    "addsrc_H": [_i, _i, _i, _fp, _i, _s] + [_fp] * 12 + [_i],
    "SetDeviceFilterParameters": [_i, _fp, _fp],
    "addkinsrc_H": [_i, _i, _fp, _i, _s] + [_fp] * 15 + [_i],
    "frcvel_H": [_i, _i, _i, _fp, _i, _i, _s] + [_fp] * 9 + [_i, _i, _i],
    "velbuffer_H": [_fp] * 8 + [_i] * 13 + [_s, _i, _fp, _i],
    "awp_receivers_gather": [_fp] * 4 + [_i, _fp, _i, _s, _i],
    "drprecpc_calc_H_opt": [_fp] * 13 + [_i] * 5 + [_s, _i],
    "awp_drprecpc_calc_tv": [_fp] * 13 + [_i] * 5 + [_s, _i, _f],
    "drprecpc_app_H": [_fp] * 9 + [_i] * 5 + [_s, _i],
    "awp_dvelcx_dm": [_fp] * 13 + [_i, _i, _s, _i, _i, _i, _i],
    "dvelc2_H": [_fp] * 13 + [_i, _i, _s, _i],
    "dstrqc2_H": [_fp] * 22 + [_i, _i, _s] + [_fp] * 5 + [_i] * 5,
    "awp_dstrqc2_tex": [_fp] * 22 + [_i, _i, _s] + [_fp] * 3 + [_i] * 5,
    "intp3d_H": [_fp] * 18 + [_i, _i, _i, _s, _i],
    "swap_H": [_fp] * 18 + [_i, _i] + [_fp] * 4 + [_i, _s, _i],
    "awp_dvelcx_s16": [_fp] * 13 + [_i, _i, _s, _i, _i, _i, _i, _i, _fp],
    "awp_dstrqc_s16": [_fp] * 23 + [_i, _i, _s] + [_fp] * 5 + [_i] * 9 + [_i, _fp],
    "awp_dstrqc_tex_s16": [_fp] * 23 + [_i, _i, _s] + [_fp] * 3 + [_i] * 9 + [_i, _fp],
    "awp_fstr_s16": [_fp] * 3 + [_s, _i, _i, _i, _i, _i, _fp],
    "awp_addsrc_s16": [_i, _i, _i, _fp, _i, _s] + [_fp] * 12 + [_i, _i, _fp],
    "awp_receivers_gather_s16": [_fp] * 4 + [_i, _fp, _i, _s, _i, _i, _fp],
    "awp_s16_encode": [_fp, _fp, C.c_longlong, _i, _f, _s],
    "awp_s16_decode": [_fp, _fp, C.c_longlong, _i, _f, _s],
    "awp_s16_rescale": [_fp, C.c_longlong, _i, _i, _s],
    "awp_s16_absmax": [_fp, C.c_longlong, _i, _fp, _s],
    "awp_adjoint_step": [_fp] * 13 + [_s],
    "awp_adjoint_settle": [_fp, _fp, _s],
    "awp_adjoint_inject": [_fp] * 4 + [_i, _fp, _s],
    "awp_adjoint_flush": [_fp, _fp, C.c_longlong, _s],
    "awp_misfit": [_fp, _fp, C.c_longlong, _fp, _fp, _fp, _s],
    "awp_misfit_noalloc": [_fp, _fp, C.c_longlong, _fp, _fp, _s],
    "awp_adjoint_map": [_fp] * 11 + [_s],
    "awp_dvelcx_oop": [_fp] * 16 + [_i, _i, _s, _i, _i, _i, _i],
    "awp_dstrqc_oop": [_fp] * 24 + [_i, _i, _s] + [_fp] * 5 + [_i] * 9,
    "awp_dstrqc_tex": [_fp] * 23 + [_i, _i, _s] + [_fp] * 3 + [_i] * 9,
    "awp_step": [_fp] * 11 + [_i, _i, _i, _s],
    "awp_stream_probe": [_fp, _i, _fp, _i, C.c_longlong, _i, _i, _i, _s],
    # topo_t* travels as a plain address (ctypes.byref of awp_amd.refstructs.Topo)
    "topo_set_constants": [_fp],
    "topo_velocity_interior_H": [_fp],
    "topo_velocity_front_H": [_fp],
    "topo_velocity_back_H": [_fp],
    "topo_stress_interior_H": [_fp],
    "topo_stress_left_H": [_fp],
    "topo_stress_right_H": [_fp],
    "awp_topo_stress_interior_rates_H": [_fp, _fp],
    "awp_topo_stress_interior_tex_H": [_fp, _fp, _fp, _i],
    "awp_topo_velocity_s16": [_fp, _i, _fp],
    "awp_topo_stress_s16": [_fp, _i, _fp, _fp],
    "awp_topo_geom_custom": [_fp, _i, _i, _i, _f, _s],
    "awp_topo_metrics_f": [_fp, _i, _i, _f, _s],
    "awp_topo_geom_no_stretching": [_fp, _i, _f, _s],
    "awp_topo_metrics_g": [_fp, _i, _f, _s],
    # cu_interp_t* travels as a plain address (awp_amd.refstructs.CuInterp)
    "awp_interp_set_stream": [_fp, _s],
    "awp_interp_num_cells": [_fp],
    "cuinterp_init": [_fp] * 4 + [Grid3] + [_fp] * 4 + [_i, _i],
    "cuinterp_lagrange_h": [_fp] * 4 + [Grid3] + [_fp] * 3,
    "cuinterp_htod": [_fp],
    "cuinterp_dtoh": [_fp],
    "cuinterp_malloc": [_fp],
    "cuinterp_finalize": [_fp],
    "cuinterp_interp_H": [_fp] * 3,
    "cusource_add_cartesian_H": [_fp] * 3 + [_f, _f],
    "cusource_add_curvilinear_H": [_fp] * 3 + [_f, _f, _fp, _i, _fp],
    "cusource_add_force_H": [_fp] * 4 + [_f, _f, _f, _fp, _i, _i, _i, _fp],
}
RESTYPES = {
    "awp_last_error": _i,
    "awp_last_error_string": C.c_char_p,
    "awp_set_tuning": _i,
    "awp_interp_set_stream": _i,
    "awp_interp_num_cells": _i,
    "cuinterp_init": _i,
    "cuinterp_lagrange_h": _i,
    "cuinterp_htod": _i,
    "cuinterp_dtoh": _i,
    "cuinterp_malloc": _i,
    "awp_stream_probe": _i,
}

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m awp_amd.build` "
            "(the HIP path has no CPU fallback)")
    # torch ships its own HIP runtime (libamdhip64.so.7 in torch/lib).  Loading
    # torch first makes the library bind to that copy by soname, so kernels and
    # torch's allocations share one runtime; loaded the other way round the
    # process holds two runtimes and the second reports no device.
    import torch  # noqa: F401
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = RESTYPES.get(name, None)
    _lib = lib
    return lib


class AwpError(RuntimeError):
    pass


def check() -> None:
    """Raise if the library recorded an error since the last check."""
    lib = load()
    code = lib.awp_last_error()
    if code:
        msg = lib.awp_last_error_string().decode()
        lib.awp_clear_error()
        raise AwpError(f"awp_hip error {code}: {msg}")


def set_tuning(key: str, value: int) -> None:
    if load().awp_set_tuning(key.encode(), int(value)) != 0:
        raise ValueError(f"unknown tuning key or bad value: {key}={value}")
