"""Lagrange-interpolated receivers and sources: the host side of the
reference's cu_interp_t path (src/interpolation/interpolation.cu,
src/topography/sources/source.cu).

`Interpolation` calls cuinterp_init in libawp_hip.so, which computes the
Lagrange coefficients of every query point on the host and uploads them;
`interp` samples a field at the queries (receivers, cuinterp_interp_H) and
`add_cartesian` / `add_curvilinear` / `add_force` inject source values
(cusource_add_*_H).  Field arrays are device tensors of the padded layout of
awp_amd.grid; receiver outputs and source inputs are 1-D device tensors
indexed by the query's `indices` entry.

A grid is the reference's grid3_t (awp_amd.refstructs.Grid3), read as
grid_index does (include/grid/grid_3d.h:113-115): point (i, j, k) of the
coordinate arrays is field element
offset1.z + k + (offset1.y + j)*line + (offset1.x + i)*slice.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .grid import ALIGN, NGSL, Dims
from .refstructs import CuInterp, Grid3, boundary_size, grid3, numel

# include/awp/error.h:6-35
ERRORS = {
    0: "SUCCESS",
    400: "ERR_GRID_OUT_OF_BOUNDS",
    600: "ERR_LAGRANGE_LOWER_EXTRAPOLATION",
    601: "ERR_LAGRANGE_UPPER_EXTRAPOLATION",
    700: "ERR_OUT_OF_BOUNDS_LOWER",
    701: "ERR_OUT_OF_BOUNDS_UPPER",
    800: "ERR_NON_POSITIVE",
    900: "ERR_INTERP_MALLOC",
    902: "ERR_INCONSISTENT_SIZE",
}


_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)


class GridDesc:
    """A grid3_t (`c`) with the quantities the Lagrange path derives from it:
    size_x/y/z, the coordinate-array lengths (grid_boundary_size), off_x/y/z
    (offset1), line, slice, and numel, the values in one field (mem.x mem.y
    mem.z)."""

    def __init__(self, c: Grid3):
        self.c = c
        self.size_x, self.size_y, self.size_z = boundary_size(c)
        self.off_x, self.off_y, self.off_z = c.offset1.x, c.offset1.y, c.offset1.z
        self.line, self.slice = c.line, c.slice
        self.numel = numel(c)

    @classmethod
    def plain(cls, off_x: int, off_y: int, off_z: int, line: int, slice_: int, size_x: int,
              size_y: int, size_z: int, numel_: int) -> "GridDesc":
        """A cell-centred grid (no point dropped) of the given coordinate sizes on
        a field of at least numel_ values with strides line and slice."""
        mem = (-(-int(numel_) // int(slice_)), int(slice_) // int(line), int(line))
        return cls(grid3((size_x, size_y, size_z), (1, 1, 1), (off_x, off_y, off_z), mem, line,
                         slice_))


def metric_grid(dims: Dims, size_z: int | None = None, h: float = 0.0) -> GridDesc:
    """The reference's metric grid of one rank (grid_init_metric_grid,
    src/grid/grid_3d.c:26-102: padding ngsl in x and y) on the padded field
    layout: coordinate index 0 in x and y is field index 2 (ngsl cells
    outside the interior), index 0 in z is field level `align`.  A node grid
    (shift 0) with an open top keeps nzt points in z (size_z = nzt, the
    default); with a closed top (size_z = nzt - 1) grid_boundary_size drops
    the top point."""
    nz = dims.nzt
    if size_z is None or size_z == nz:
        b2 = (0, 0, 0)
    elif size_z == nz - 1:
        b2 = (0, 0, 1)
    else:
        raise ValueError(f"size_z must be nzt ({nz}) or nzt - 1")
    mem = (dims.nxt + 4 + 2 * NGSL, dims.nyt + 4 + 2 * NGSL, nz + 2 * ALIGN)
    return GridDesc(grid3((dims.nxt + 2 * NGSL, dims.nyt + 2 * NGSL, nz), (0, 0, 0),
                          (2, 2, ALIGN), mem, mem[2], mem[1] * mem[2], boundary2=b2,
                          gridspacing=h))


def grid_coordinates(size: int, h: float, shift: int = 0, padding: int = 0) -> np.ndarray:
    """Coordinates of one direction of a single-rank grid, the interior
    formula of grid_fill1 (src/grid/grid_3d.c:258-263):
    h (i - shift/2 - padding), evaluated in double from the single h."""
    i = np.arange(size, dtype=np.float64)
    return (np.float64(np.float32(h)) * (i - 0.5 * shift - padding)).astype(np.float32)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class Interpolation:
    """One cu_interp_t: num_query points, Lagrange degree `degree`."""

    def __init__(self, x, y, z, grid: GridDesc, qx, qy, qz, indices=None, degree: int = 3):
        lib = _lib.load()
        self._x, self._y, self._z = _f32(x), _f32(y), _f32(z)
        if (len(self._x), len(self._y), len(self._z)) != (grid.size_x, grid.size_y, grid.size_z):
            raise ValueError("coordinate arrays do not match the grid's sizes")
        qx, qy, qz = _f32(qx), _f32(qy), _f32(qz)
        nq = len(qx)
        if len(qy) != nq or len(qz) != nq:
            raise ValueError("qx, qy, qz differ in length")
        idx = np.arange(nq, dtype=np.int32) if indices is None else \
            np.ascontiguousarray(np.asarray(indices, dtype=np.int32))
        if len(idx) != nq:
            raise ValueError("indices must have one entry per query")
        if nq and idx.min() < 0:
            raise ValueError("indices must be non-negative")
        self._slots = int(idx.max()) + 1 if nq else 0
        self.grid = grid
        self.I = CuInterp()
        self._live = False
        fp = lambda a: a.ctypes.data_as(_FP)  # noqa: E731
        err = lib.cuinterp_init(C.byref(self.I), fp(self._x), fp(self._y), fp(self._z), grid.c,
                                fp(qx), fp(qy), fp(qz), idx.ctypes.data_as(_IP), nq, int(degree))
        self._live = self.I.lx is not None and bool(self.I.lx)
        if err != 0:
            self.close()
            _lib.check()
            raise ValueError(f"cuinterp_init failed: {ERRORS.get(err, err)}")
        _lib.check()

    # host copies of the coefficients (num_query x num_basis) and indices
    def coefficients(self):
        nq, nb = self.I.num_query, self.I.num_basis
        arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()  # noqa: E731
        return (arr(self.I.lx, nq * nb).reshape(nq, nb), arr(self.I.ly, nq * nb).reshape(nq, nb),
                arr(self.I.lz, nq * nb).reshape(nq, nb), arr(self.I.ix, nq), arr(self.I.iy, nq),
                arr(self.I.iz, nq))

    @property
    def num_query(self) -> int:
        return self.I.num_query

    @property
    def num_cells(self) -> int:
        """Grid points of the source map."""
        return int(_lib.load().awp_interp_num_cells(C.byref(self.I)))

    def _stream(self) -> None:
        import torch
        st = torch.cuda.current_stream().cuda_stream
        if st != getattr(self, "_st", None):
            if _lib.load().awp_interp_set_stream(C.byref(self.I), st) != 0:
                raise RuntimeError("awp_interp_set_stream: no source map (cuinterp_htod)")
            self._st = st

    def _check_field(self, t, slots) -> None:
        if t.numel() < self.grid.numel:
            raise ValueError(f"field has {t.numel()} elements, the grid needs {self.grid.numel}")
        if slots.numel() < self._slots:
            raise ValueError(f"{slots.numel()} values, the indices address {self._slots}")

    def interp(self, out, field) -> None:
        """out[indices[q]] = the degree-p Lagrange interpolant of field at query q."""
        self._check_field(field, out)
        self._stream()
        _lib.load().cuinterp_interp_H(C.byref(self.I), out.data_ptr(), field.data_ptr())
        _lib.check()

    def add_cartesian(self, field, values, h: float, dt: float) -> None:
        """field -= dt/h^3 * weights * values[indices[q]] (a moment-rate source)."""
        self._check_field(field, values)
        self._stream()
        _lib.load().cusource_add_cartesian_H(C.byref(self.I), field.data_ptr(), values.data_ptr(),
                                             float(h), float(dt))
        _lib.check()

    @staticmethod
    def _check_len(name, t, n) -> None:
        if t.numel() < n:
            raise ValueError(f"{name} has {t.numel()} elements, its layout needs {n}")

    def add_curvilinear(self, field, values, h: float, dt: float, f, ny: int, dg) -> None:
        """As add_cartesian, divided by the topography Jacobian f(x, y) dg(z)."""
        self._check_field(field, values)
        self._check_len("f", f, (self.grid.numel // self.grid.slice) * (ny + 4 + 2 * NGSL + 2 * ALIGN))
        self._check_len("dg", dg, self.grid.line)
        self._stream()
        _lib.load().cusource_add_curvilinear_H(C.byref(self.I), field.data_ptr(),
                                               values.data_ptr(), float(h), float(dt),
                                               f.data_ptr(), int(ny), dg.data_ptr())
        _lib.check()

    def add_force(self, field, values, d1, h: float, dt: float, quad_weight: float, f,
                  nx: int, ny: int, nz: int, dg) -> None:
        """A body force on a velocity field, scaled by -quad_weight / (f dg rho)."""
        self._check_field(field, values)
        self._check_len("f", f, (nx + 4 + 2 * NGSL) * (ny + 4 + 2 * NGSL + 2 * ALIGN))
        self._check_len("dg", dg, nz + 2 * ALIGN)
        self._check_len("d1", d1, (nx + 4 + 2 * NGSL) * (ny + 4 + 2 * NGSL) * (nz + 2 * ALIGN))
        self._stream()
        _lib.load().cusource_add_force_H(C.byref(self.I), field.data_ptr(), values.data_ptr(),
                                         d1.data_ptr(), float(h), float(dt), float(quad_weight),
                                         f.data_ptr(), int(nx), int(ny), int(nz), dg.data_ptr())
        _lib.check()

    def close(self) -> None:
        if getattr(self, "_live", False):
            _lib.load().cuinterp_finalize(C.byref(self.I))
            self._live = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lagrange_h(x, y, z, grid: GridDesc, qx, qy, qz, degree: int):
    """The library's host coefficient computation alone (cuinterp_lagrange_h),
    no device memory: returns (err, lx, ly, lz, ix, iy, iz)."""
    lib = _lib.load()
    x, y, z, qx, qy, qz = map(_f32, (x, y, z, qx, qy, qz))
    nq, nb = len(qx), degree + 1
    lx, ly, lz = (np.zeros(nq * nb, np.float32) for _ in range(3))
    ix, iy, iz = (np.zeros(nq, np.int32) for _ in range(3))
    H = CuInterp()
    H.num_basis, H.num_query = nb, nq
    H.lx, H.ly, H.lz = (a.ctypes.data_as(_FP) for a in (lx, ly, lz))
    H.ix, H.iy, H.iz = (a.ctypes.data_as(_IP) for a in (ix, iy, iz))
    fp = lambda a: a.ctypes.data_as(_FP)  # noqa: E731
    err = lib.cuinterp_lagrange_h(C.byref(H), fp(x), fp(y), fp(z), grid.c, fp(qx), fp(qy),
                                  fp(qz))
    return (err, lx.reshape(nq, nb), ly.reshape(nq, nb), lz.reshape(nq, nb), ix, iy, iz)
