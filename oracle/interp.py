"""Python driver of the CPU oracle for Lagrange-interpolated receivers and
sources (oracle/awp_oracle_interp.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
Each wrapper calls the C restatement of one reference function; the file and
line each one follows are listed in awp_oracle_interp.cpp.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle.oracle import load as _load_lib

P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
_ready = False


def _lib():
    global _ready
    lib = _load_lib()
    if not _ready:
        lib.awp_oracle_lagrange_node_poly.argtypes = [F, P, I]
        lib.awp_oracle_lagrange_node_poly.restype = D
        lib.awp_oracle_lagrange_basis.argtypes = [P, I, P, F]
        lib.awp_oracle_lagrange_basis.restype = I
        lib.awp_oracle_argnearest.argtypes = [P, P, I, F]
        lib.awp_oracle_argnearest.restype = I
        lib.awp_oracle_argnearest_range.argtypes = [P, P, I, I, I, I, F]
        lib.awp_oracle_argnearest_range.restype = I
        lib.awp_oracle_lagrange1_coef.argtypes = [P, P, P, P, I, F, I]
        lib.awp_oracle_lagrange1_coef.restype = I
        lib.awp_oracle_lagrange1.argtypes = [P, P, P, I, P, I, I]
        lib.awp_oracle_lagrange1.restype = I
        lib.awp_oracle_lagrange_h.argtypes = [P] * 9 + [I, I, I] + [P] * 3 + [I, I]
        lib.awp_oracle_lagrange_h.restype = I
        lib.awp_oracle_interp.argtypes = [P] * 5 + [I] + [P] * 4 + [I] * 6
        lib.awp_oracle_source_add.argtypes = ([I] + [P] * 5 + [I] + [P] * 4 + [I] * 6 +
                                              [F, F, F] + [P] * 3 + [I, I])
        _ready = True
    return lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def node_poly(x0: float, x) -> float:
    x = _f32(x)
    return _lib().awp_oracle_lagrange_node_poly(float(np.float32(x0)), x.ctypes.data, len(x))


def lagrange_basis(x, x0: float):
    """(err, l): lagrange_basis on nodes x at x0."""
    x = _f32(x)
    l = np.zeros(len(x), np.float32)
    err = _lib().awp_oracle_lagrange_basis(l.ctypes.data, len(x), x.ctypes.data,
                                           float(np.float32(x0)))
    return err, l


def argnearest(points, query: float):
    p = _f32(points)
    nearest = C.c_int(-1)
    err = _lib().awp_oracle_argnearest(C.addressof(nearest), p.ctypes.data, len(p),
                                       float(np.float32(query)))
    return err, nearest.value


def argnearest_range(lower: int, upper: int, nearest: int, n: int, query: float = 0.0):
    first, last = C.c_int(0), C.c_int(0)
    err = _lib().awp_oracle_argnearest_range(C.addressof(first), C.addressof(last), lower, upper,
                                             nearest, n, float(query))
    return err, first.value, last.value


def lagrange1_coef(x, query: float, deg: int):
    """(err, xloc, l, first) of interp_lagrange1_coef."""
    x = _f32(x)
    xloc = np.zeros(deg + 1, np.float32)
    l = np.zeros(deg + 1, np.float32)
    first = C.c_int(-1)
    err = _lib().awp_oracle_lagrange1_coef(xloc.ctypes.data, l.ctypes.data, C.addressof(first),
                                           x.ctypes.data, len(x), float(np.float32(query)), deg)
    return err, xloc, l, first.value


def lagrange1(x, values, queries, deg: int) -> np.ndarray:
    x, values, queries = _f32(x), _f32(values), _f32(queries)
    out = np.zeros(len(queries), np.float32)
    _lib().awp_oracle_lagrange1(out.ctypes.data, x.ctypes.data, values.ctypes.data, len(x),
                                queries.ctypes.data, len(queries), deg)
    return out


def lagrange_h(x, y, z, qx, qy, qz, deg: int):
    """(err, lx, ly, lz, ix, iy, iz) of cuinterp_lagrange_h."""
    x, y, z, qx, qy, qz = map(_f32, (x, y, z, qx, qy, qz))
    nq, nb = len(qx), deg + 1
    lx, ly, lz = (np.zeros((nq, nb), np.float32) for _ in range(3))
    ix, iy, iz = (np.zeros(nq, np.int32) for _ in range(3))
    err = _lib().awp_oracle_lagrange_h(lx.ctypes.data, ly.ctypes.data, lz.ctypes.data,
                                       ix.ctypes.data, iy.ctypes.data, iz.ctypes.data,
                                       x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                       len(x), len(y), len(z), qx.ctypes.data, qy.ctypes.data,
                                       qz.ctypes.data, nq, deg)
    return err, lx, ly, lz, ix, iy, iz


def _grid(grid):
    return (int(grid.off_x), int(grid.off_y), int(grid.off_z), int(grid.line), int(grid.slice))


def interp(out: np.ndarray, field: np.ndarray, coef, ridx, grid) -> None:
    """cuinterp_dinterp on the host: out[ridx[q]] = interpolant at query q."""
    lx, ly, lz, ix, iy, iz = coef
    lx, ly, lz = map(_f32, (lx, ly, lz))
    ix, iy, iz, ridx = map(_i32, (ix, iy, iz, ridx))
    assert out.dtype == np.float32 and out.flags.c_contiguous
    assert field.dtype == np.float32 and field.flags.c_contiguous
    _lib().awp_oracle_interp(out.ctypes.data, field.ctypes.data, lx.ctypes.data, ly.ctypes.data,
                             lz.ctypes.data, lx.shape[1], ix.ctypes.data, iy.ctypes.data,
                             iz.ctypes.data, ridx.ctypes.data, len(ix), *_grid(grid))


KINDS = {"cartesian": 0, "curvilinear": 1, "force": 2}


def source_add(kind: str, out: np.ndarray, values: np.ndarray, coef, lidx, grid, h: float,
               dt: float, quad_weight: float = 0.0, f=None, dg=None, d1=None, ny: int = 0,
               nz: int = 0) -> None:
    """cusource_add_{cartesian,curvilinear,force} on the host, queries in
    ascending order."""
    lx, ly, lz, ix, iy, iz = coef
    lx, ly, lz = map(_f32, (lx, ly, lz))
    ix, iy, iz, lidx = map(_i32, (ix, iy, iz, lidx))
    values = _f32(values)
    assert out.dtype == np.float32 and out.flags.c_contiguous
    keep = [None if a is None else _f32(a) for a in (f, dg, d1)]
    fp, gp, dp = (None if a is None else a.ctypes.data for a in keep)
    _lib().awp_oracle_source_add(KINDS[kind], out.ctypes.data, values.ctypes.data,
                                 lx.ctypes.data, ly.ctypes.data, lz.ctypes.data, lx.shape[1],
                                 ix.ctypes.data, iy.ctypes.data, iz.ctypes.data, lidx.ctypes.data,
                                 len(ix), *_grid(grid), float(h), float(dt), float(quad_weight),
                                 fp, gp, dp, int(ny), int(nz))
