"""ctypes mirrors of the reference's structs that the topography and
Lagrange launchers take (restated in include/awp_hip.h): int3_t, grid3_t
(fcn_grid_t), f_grid_t, g_grid_t, topo_t and cu_interp_t, member for member
in the reference's order (include/awp/definitions.h:55-58,
include/grid/grid_3d.h:90-111, include/topography/metrics/metrics.h:58-127,
include/topography/topography.h:136-273,
include/interpolation/interpolation.cuh:21-32).  _prec and prec are float,
CUDA's dim3 is three 32-bit unsigned integers, streams are opaque pointers.
"""
from __future__ import annotations

import ctypes as C

_p = C.c_void_p  # device and host pointers travel as plain addresses


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Dim3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class Grid3(C.Structure):
    """grid3_t = fcn_grid_t."""
    _fields_ = ([(n, Int3) for n in ("size", "inner_size", "outer_size", "mem", "coordinate",
                                     "shift", "offset1", "offset2", "boundary1", "boundary2",
                                     "alignment")]
                + [("gridspacing", C.c_float), ("padding", Int3), ("num_bytes", C.c_int),
                   ("line", C.c_int), ("slice", C.c_int), ("exclude_top_row", C.c_int)])


F_ARRAYS = ("f", "f_1", "f_2", "f_c", "f1_1", "f1_2", "f1_c", "f2_1", "f2_2", "f2_c")
G_ARRAYS = ("g", "g_c", "g3", "g3_c")


class FGrid(C.Structure):
    """f_grid_t: host arrays f .. f2_c, then the device arrays d_f .. d_f2_c."""
    _fields_ = ([("size", C.c_int * 3), ("mem", C.c_int * 3), ("offset", C.c_int * 3),
                 ("bounds_x", C.c_int * 2), ("bounds_y", C.c_int * 2),
                 ("bounds_stress_x", C.c_int * 2), ("bounds_stress_y", C.c_int * 2),
                 ("hi", C.c_float), ("line", C.c_int), ("slice", C.c_int)]
                + [(n, _p) for n in F_ARRAYS] + [("d_" + n, _p) for n in F_ARRAYS])


class GGrid(C.Structure):
    """g_grid_t: host arrays g g_c g3 g3_c, then their device arrays."""
    _fields_ = ([("size", C.c_int), ("mem", C.c_int), ("line", C.c_int), ("offset", C.c_int),
                 ("bounds_z", C.c_int * 2), ("hi", C.c_float)]
                + [(n, _p) for n in G_ARRAYS] + [("d_" + n, _p) for n in G_ARRAYS])


TOPO_MEDIUM = ("rho", "lami", "mui", "lam_mu")
TOPO_WAVE = ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6")
TOPO_Q = ("qpi", "qsi", "ww", "wwo", "vx1", "vx2", "coeff")


class Topo(C.Structure):
    """topo_t."""
    _fields_ = ([("use", C.c_int), ("dbg", C.c_int), ("verbose", C.c_int), ("rank", C.c_int),
                 ("topography_file", C.c_char_p),
                 ("x_rank_l", C.c_int), ("x_rank_r", C.c_int), ("y_rank_b", C.c_int),
                 ("y_rank_f", C.c_int), ("coord", C.c_int * 2), ("px", C.c_int), ("py", C.c_int),
                 ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                 ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
                 ("off_x", C.c_int * 4), ("off_y", C.c_int * 4), ("off_z", C.c_int * 4),
                 ("velocity_bounds_left", C.c_int * 2), ("velocity_bounds_right", C.c_int * 2),
                 ("velocity_bounds_front", C.c_int * 2), ("velocity_bounds_back", C.c_int * 2),
                 ("velocity_bounds_xdir", C.c_int * 2), ("velocity_offset_x", C.c_int * 6),
                 ("velocity_offset_y", C.c_int * 6), ("boundary_top_size", C.c_int),
                 ("boundary_bottom_size", C.c_int), ("velocity_grid_interior", Dim3),
                 ("velocity_grid_front", Dim3), ("velocity_grid_back", Dim3),
                 ("stress_bounds_left", C.c_int * 2), ("stress_bounds_right", C.c_int * 2),
                 ("stress_bounds_ydir", C.c_int * 2), ("stress_offset_x", C.c_int * 6),
                 ("stress_offset_y", C.c_int * 4), ("stress_grid_interior", Dim3),
                 ("stress_grid_left", Dim3), ("stress_grid_right", Dim3)]
                + [(n, C.c_int * 3) for n in ("sxx", "syy", "szz", "sxy", "sxz", "syz", "su1",
                                              "sv1", "sw1")]
                + [("gridsize", C.c_int), ("slice", C.c_int), ("slice_gl", C.c_int),
                   ("line", C.c_int), ("dth", C.c_float), ("timestep", C.c_float),
                   ("gridspacing", C.c_float)]
                + [(n, _p) for n in TOPO_MEDIUM + TOPO_WAVE + TOPO_Q]
                + [(n, _p) for n in ("u1", "v1", "w1", "f_u1", "f_v1", "f_w1", "b_u1", "b_v1",
                                     "b_w1")]
                + [("metrics_f", FGrid), ("metrics_g", GGrid), ("topography_grid", Grid3),
                   ("stress_grid", Grid3), ("velocity_grid", Grid3)]
                + [(n, _p) for n in ("x1", "y1", "z1", "dcrjx", "dcrjy", "dcrjz", "stream_1",
                                     "stream_2", "stream_i")])


class CuInterp(C.Structure):
    """cu_interp_t."""
    _fields_ = ([("num_basis", C.c_int), ("num_query", C.c_int), ("size_l", C.c_size_t),
                 ("size_i", C.c_size_t)]
                + [(n, C.POINTER(C.c_float)) for n in ("lx", "ly", "lz")]
                + [(n, C.POINTER(C.c_int)) for n in ("ix", "iy", "iz", "ridx")]
                + [(n, _p) for n in ("d_lx", "d_ly", "d_lz", "d_ix", "d_iy", "d_iz", "d_ridx")]
                + [("grid", Grid3)])


def grid3(size, shift, offset1, mem, line: int, slice_: int, boundary2=(0, 0, 0),
          gridspacing: float = 0.0) -> Grid3:
    """A grid3_t with the members the Lagrange path reads; the others zero."""
    g = Grid3()
    g.size = Int3(*size)
    g.inner_size = Int3(*size)
    g.outer_size = Int3(*size)
    g.shift = Int3(*shift)
    g.offset1 = Int3(*offset1)
    g.mem = Int3(*mem)
    g.boundary2 = Int3(*boundary2)
    g.line, g.slice = int(line), int(slice_)
    g.gridspacing = float(gridspacing)
    g.num_bytes = C.c_int(4 * mem[0] * mem[1] * mem[2]).value  # an int, as the reference's
    return g


def boundary_size(g: Grid3) -> tuple[int, int, int]:
    """grid_boundary_size (src/grid/grid_3d.c:143-160)."""
    def one(size, shift, b2):
        return size - 1 if shift == 0 and b2 == 1 else size
    return (one(g.size.x, g.shift.x, g.boundary2.x), one(g.size.y, g.shift.y, g.boundary2.y),
            one(g.size.z, g.shift.z, g.boundary2.z))


def numel(g: Grid3) -> int:
    return g.mem.x * g.mem.y * g.mem.z
