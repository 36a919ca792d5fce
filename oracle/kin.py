"""Python driver of the CPU oracle for kinematic finite-fault sources and
prescribed-velocity sources (oracle/awp_oracle_kin.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle.oracle import load as _load_lib

P, I, F, D, L = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_longlong
_ready = False


def _lib():
    global _ready
    lib = _load_lib()
    if not _ready:
        lib.awp_oracle_compmt.argtypes = [F, F, F, P]
        lib.awp_oracle_brune.argtypes = [F, F]
        lib.awp_oracle_brune.restype = F
        lib.awp_oracle_liu.argtypes = [F, F]
        lib.awp_oracle_liu.restype = F
        lib.awp_oracle_lfilter.argtypes = [I, P, P, D, P]
        lib.awp_oracle_lfilter.restype = D
        lib.awp_oracle_addkinsrc.argtypes = ([I, I, P, I] + [P] * 13 + [P, I, P, P] +
                                             [F, F, I, L])
        lib.awp_oracle_frcvel.argtypes = [I, I, I, P, I, I] + [P] * 6 + [I, I, F, I, L]
        _ready = True
    return lib


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data


def compmt(strike, dip, rake) -> np.ndarray:
    """xx, yy, zz, xz, yz, xy of compmt (angles in radians, single precision)."""
    m = np.zeros(6, np.float32)
    _lib().awp_oracle_compmt(float(np.float32(strike)), float(np.float32(dip)),
                             float(np.float32(rake)), _p(m))
    return m


def brune(freq, t) -> float:
    return _lib().awp_oracle_brune(float(np.float32(freq)), float(np.float32(t)))


def liu(tau, t) -> float:
    return _lib().awp_oracle_liu(float(np.float32(tau)), float(np.float32(t)))


def lfilter(b, a, x) -> np.ndarray:
    """The reference's lfilter applied sample by sample from a zero state."""
    b = np.ascontiguousarray(b, np.float64)
    a = np.ascontiguousarray(a, np.float64)
    order = len(b) - 1
    d = np.zeros(order + 1, np.float64)
    lib = _lib()
    return np.array([lib.awp_oracle_lfilter(order, _p(b), _p(a), float(v), _p(d)) for v in x])


def addkinsrc(step, psrc, mu, cols, fields, mom, filt_d, b, a, DT, DH, dims) -> None:
    """One addkinsrc launch on host arrays (fields: xx, yy, zz, xy, yz, xz)."""
    order = 0 if b is None else len(b) - 1
    b = np.zeros(1) if b is None else np.ascontiguousarray(b, np.float64)
    a = np.ones(1) if a is None else np.ascontiguousarray(a, np.float64)
    _lib().awp_oracle_addkinsrc(step, psrc.shape[1], _p(psrc), psrc.shape[0], _p(mu),
                                *(_p(c) for c in cols), *(_p(f) for f in fields), _p(mom),
                                _p(filt_d) if filt_d is not None else None, order, _p(b), _p(a),
                                DT, DH, dims.yline, dims.slice)


def frcvel(step, read_step, psrc, tskp, cols, fields, xmin, xmax, DT, dims) -> None:
    """One frcvel launch on host arrays (cols: axx, ayy, azz; fields: u1, v1, w1)."""
    _lib().awp_oracle_frcvel(step, read_step, psrc.shape[1], _p(psrc), psrc.shape[0], tskp,
                             *(_p(c) for c in cols), *(_p(f) for f in fields), xmin, xmax, DT,
                             dims.yline, dims.slice)
