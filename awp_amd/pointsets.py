"""Input-file-driven point sets on one rank's grid stack: receivers, moment-
tensor sources, boundary forces and strain-Green-tensor (SGT) output
(reference src/topography/receivers/{receivers,receiver,sgt}.c,
src/topography/sources/{sources,source,forces}.c, src/topography/grids.c,
src/grid/grid_3d.c, src/grid/shift.c, src/mpi/distribute.c).

Each point set reads a v2 input file (awp_amd.inputs), assigns every point to
a grid of the stack by depth, maps its user coordinates to the grid's
internal coordinates (the half-cell material shift, the discontinuous-mesh
offsets, and on a topography grid the depth mapping through the metric f),
and builds one cu_interp_t per grid on that grid's staggered positions
(awp_amd.interpolation, the library's cuinterp_* entry points).  Receivers
sample through cuinterp_interp_H into a step buffer and write receiver files;
sources read their time series from files and inject them with
cusource_add_*_H.

One rank holds the whole mesh: the MPI broadcast of the input, the
communicator split of the users, and the partition of points between ranks
(dist_indices keeps every point inside the rank's x/y extent) are done for
that single rank.  Host arithmetic follows the reference's single-precision
(_prec) steps so the coordinates, and hence the interpolation coefficients,
are the reference's.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

from .grid import ALIGN, NGSL
from .inputs import (INPUT_SURFACE_COORD, INPUT_VOLUME_COORD, Input, StepBuffer, read_input,
                     receiver_filename, write_receiver_block)
from .refstructs import Grid3, Int3

f32 = np.float32

# enum grid_types (include/topography/grids.h:30)
X, Y, Z, XX, YY, ZZ, XY, XZ, YZ, NODE = range(10)
# shift of each grid type (src/grid/shift.c:76-150): 1 = shifted by half a cell
SHIFT = {X: (1, 1, 1), Y: (0, 0, 1), Z: (0, 1, 0), XX: (0, 1, 1), YY: (0, 1, 1),
         ZZ: (0, 1, 1), XY: (1, 0, 1), XZ: (1, 1, 0), YZ: (0, 0, 0), NODE: (1, 0, 0)}
# x/y padding of each grid type (grids.c:22-38): velocity and node grids 0,
# stress grids ngsl/2
PADDING = {t: (NGSL // 2 if t in (XX, YY, ZZ, XY, XZ, YZ) else 0) for t in SHIFT}
CLOSED = 1                     # enum grid_boundary_types
OVERLAP = 7.0                  # source.c:19
SOURCE_DM_OFFSET_X = 0         # include/topography/sources/source.h:19-25
SOURCE_DM_OFFSET_Y = -1
SOURCE_OFFSET_X = -0.5


def grid_init(size, shift, coord, bnd1, bnd2, padding: int, h: float) -> Grid3:
    """grid_init (src/grid/grid_3d.c:26-75)."""
    nx, ny, nz = size
    g = Grid3()
    g.size = g.inner_size = g.outer_size = Int3(nx, ny, nz)
    g.mem = Int3(nx + 4 + 2 * NGSL, ny + 4 + 2 * NGSL, nz + 2 * ALIGN)
    g.num_bytes = C.c_int(4 * g.mem.x * g.mem.y * g.mem.z).value  # an int, as there
    g.line = g.mem.z
    g.slice = g.mem.z * g.mem.y
    g.boundary1 = Int3(*bnd1)
    g.boundary2 = Int3(*bnd2)
    g.gridspacing = float(f32(h))
    g.coordinate = Int3(*coord)
    g.shift = Int3(*shift)
    g.padding = Int3(padding, padding, 0)
    g.size = Int3(nx + 2 * padding, ny + 2 * padding, nz)
    g.alignment = Int3(2 + NGSL - padding, 2 + NGSL - padding, ALIGN)
    g.offset1 = Int3(2 + NGSL - padding, 2 + NGSL - padding, ALIGN)
    g.offset2 = Int3(2 + NGSL + nx + padding, 2 + NGSL + ny + padding, ALIGN + nz)
    g.exclude_top_row = 1 if shift[2] == 0 else 0
    if g.alignment.x < 0 or g.offset2.x > g.mem.x:
        raise ValueError("padding too large for the grid")
    return g


def metric_grid_like(g: Grid3, shift=None) -> Grid3:
    """grid_init_metric_grid(g.inner_size, shift, ...) (grid_3d.c:96-102):
    padding ngsl."""
    s = (g.shift.x, g.shift.y, g.shift.z) if shift is None else shift
    return grid_init((g.inner_size.x, g.inner_size.y, g.inner_size.z), s,
                     (g.coordinate.x, g.coordinate.y, g.coordinate.z),
                     (g.boundary1.x, g.boundary1.y, g.boundary1.z),
                     (g.boundary2.x, g.boundary2.y, g.boundary2.z), NGSL, g.gridspacing)


def grids_init(nx: int, ny: int, nz: int, topography: int, h: float,
               coord=(0, 0, 0)) -> dict:
    """grids_init (src/topography/grids.c:10-43): the grid of every type of
    one block; boundary2.z is closed on a topography block."""
    c = (coord[0], coord[1], 0)
    return {t: grid_init((nx, ny, nz), SHIFT[t], c, (0, 0, 0), (0, 0, topography), PADDING[t], h)
            for t in SHIFT}


@dataclass
class Grid1:
    id: int
    size: int
    shift: int
    boundary1: int
    boundary2: int
    padding: int
    end: int
    h: np.float32


def grid1(g: Grid3, axis: int) -> Grid1:
    """grid_grid1_x/_y/_z (grid_3d.c:222-257)."""
    get = lambda v: (v.x, v.y, v.z)[axis]  # noqa: E731
    end = get(g.size) - 1 - get(g.boundary2) if axis == 2 else 0
    return Grid1(get(g.coordinate), get(g.size), get(g.shift), get(g.boundary1),
                 get(g.boundary2), get(g.padding), end, f32(g.gridspacing))


def grid_fill1(g: Grid1) -> np.ndarray:
    """grid_fill1 (grid_3d.c:258-275): h (i + id (n - 2 p) - shift/2 - p),
    the factor in double, stored as _prec; the boundary points of closed
    boundaries in single precision."""
    i = np.arange(g.size, dtype=np.float64)
    out = (np.float64(g.h) * (i + g.id * (g.size - 2 * g.padding) - 0.5 * g.shift
                              - g.padding)).astype(f32)
    if g.shift and g.boundary1:
        out[0] = f32(g.h * f32(g.id)) * f32(g.size)
    if g.shift and g.boundary2:
        out[g.size - 1] = g.h * f32((g.id + 1) * (g.size - 2 * g.padding) - 2 - g.padding)
    if not g.shift and g.boundary2:
        out[g.size - 1] = f32(0.0)
    return out


def grid_in_bounds_ext1(x: np.ndarray, q: np.ndarray, g: Grid1) -> np.ndarray:
    """grid_in_bounds_ext1 (grid_3d.c:301-311) for an array of queries: True
    where x[0] - h/2 <= q < x[-1] + h/2 (single precision)."""
    half = f32(g.h / f32(2))
    lo = (q - (x[0] - half)).astype(f32) < 0
    hi = (q - (x[g.size - 1] + half)).astype(f32) >= 0
    return ~lo & ~hi


def find_grid_number(inp: Input, stack: list, indices: np.ndarray) -> np.ndarray:
    """source_find_grid_number (source.c:71-126): the block each point lies
    in by its z (<= 0 below the surface), blocks stacked downward with
    OVERLAP cells of overlap; surface coordinates and z > 0 go to block 0."""
    n = len(indices)
    out = np.full(n, -1, dtype=np.int64)
    lower = upper = overlap = f32(0)
    z = inp.z[indices].astype(f32)
    typ = inp.type[indices]
    for i, grids in enumerate(stack):
        gz = grid1(grids[Z], 2)
        z1 = grid_fill1(gz)
        upper = lower
        lower = f32(lower - z1[gz.end])
        sel = (z > lower) & (z <= f32(upper - overlap))
        out[sel] = i
        out[z > 0] = 0
        out[typ == INPUT_SURFACE_COORD] = 0
        overlap = f32(gz.h * OVERLAP) if i + 1 != len(stack) else f32(0)
        lower = f32(lower + overlap)
    bad = np.nonzero(out == -1)[0]
    if len(bad):
        raise ValueError(f"failed to assign source/receiver id={int(bad[0])} to a grid")
    return out


def metrics_interpolate_f_point(f_host: np.ndarray, mgrid: Grid3, qx, qy, deg: int,
                                nx: int, ny: int) -> np.ndarray:
    """metrics_interpolate_f_point (src/topography/metrics/metrics.c:171-204):
    the degree-deg Lagrange interpolant of the metric f at (qx, qy), with the
    library's coefficients on the metric grid's node coordinates; f_host is
    the (nx+20) x (ny+84) f array."""
    from .interpolation import GridDesc, lagrange_h
    x1 = grid_fill1(grid1(mgrid, 0))
    y1 = grid_fill1(grid1(mgrid, 1))
    z1 = np.zeros(max(deg + 1, 1), dtype=f32)
    n = len(qx)
    desc = GridDesc(mgrid)
    # a 1-point z axis: only the x/y coefficients are used
    zgrid = Grid3.from_buffer_copy(mgrid)
    zgrid.size = Int3(mgrid.size.x, mgrid.size.y, len(z1))
    zgrid.shift = Int3(mgrid.shift.x, mgrid.shift.y, 1)
    err, lx, ly, _, ix, iy, _ = lagrange_h(x1, y1, z1, GridDesc(zgrid), qx, qy,
                                           np.zeros(n, dtype=f32), deg)
    del desc
    flat = f_host.reshape(-1)
    cols = ny + 4 + 2 * NGSL + 2 * ALIGN
    out = np.zeros(n, dtype=f32)
    for q in range(n):
        acc = f32(0)
        for i in range(deg + 1):
            for j in range(deg + 1):
                pos = (NGSL + ALIGN + 2 - NGSL + iy[q] + j) + (NGSL + 2 - NGSL + ix[q] + i) * cols
                acc = f32(acc + f32(f32(lx[q, i] * ly[q, j]) * flat[pos]))
        out[q] = acc
    return out


class PointSet:
    """source_init_common (source.c:129-470) for one component (grid type)
    on one rank: the points of the input file mapped onto the grid stack,
    one Interpolation per grid that holds points.

    stack: per grid, the dict of grids_init; f_host: the metric f of a
    topography grid 0 (host copy), or None."""

    def __init__(self, inp: Input, suffix: str, grid_type: int, stack: list,
                 f_host: np.ndarray | None = None, upload: bool = True):
        from .interpolation import GridDesc
        self.inp, self.grid_type, self.ngrids = inp, grid_type, len(stack)
        self.filename = f"{inp.file}_{suffix}"
        n = inp.length
        allidx = np.arange(n)
        gn_all = find_grid_number(inp, stack, allidx)
        top = stack[0][grid_type]
        htop = f32(top.gridspacing)
        dmx = np.zeros(self.ngrids, f32)
        dmy = np.zeros(self.ngrids, f32)
        dmz = np.zeros(self.ngrids, f32)
        for j in range(1, self.ngrids):
            pre, cur = stack[j - 1][grid_type], stack[j][grid_type]
            hp, hc = f32(pre.gridspacing), f32(cur.gridspacing)
            dmx[j] = f32(np.float64(f32(dmx[j - 1] + f32(SOURCE_DM_OFFSET_X) * hp))
                         - (cur.shift.x * 0.5 * np.float64(hc) - pre.shift.x * 0.5 * np.float64(hp)))
            dmy[j] = f32(np.float64(f32(dmy[j - 1] + f32(SOURCE_DM_OFFSET_Y) * hp))
                         - (cur.shift.y * 0.5 * np.float64(hc) - pre.shift.y * 0.5 * np.float64(hp)))
            dmz[j] = f32(top.shift.z * 0.5 * np.float64(htop) - cur.shift.z * 0.5 * np.float64(hc))
        x = (inp.x.astype(np.float64) + SOURCE_OFFSET_X * np.float64(htop)).astype(f32)
        y = inp.y.astype(f32).copy()
        z = inp.z.astype(f32).copy()
        x = (x + dmx[gn_all]).astype(f32)
        y = (y + dmy[gn_all]).astype(f32)
        z = (z + dmz[gn_all]).astype(f32)
        # dist_indices (src/mpi/distribute.c:9-52) against the top grid
        gx, gy = grid1(top, 0), grid1(top, 1)
        inb = grid_in_bounds_ext1(grid_fill1(gx), x, gx) & \
            grid_in_bounds_ext1(grid_fill1(gy), y, gy)
        self.indices = np.nonzero(inb)[0].astype(np.int64)
        self.length = len(self.indices)
        gn = find_grid_number(inp, stack, self.indices) if self.length else np.zeros(0, int)
        self.lengths = [int((gn == j).sum()) for j in range(self.ngrids)]
        self.global_indices, self.interp, self.grids = [], [], []
        self.xq, self.yq, self.zq, self.coords = [], [], [], []
        overlap, lower = f32(0), f32(0)
        for j in range(self.ngrids):
            grid = stack[j][grid_type]
            local = np.nonzero(gn == j)[0]
            src = self.indices[local]
            self.global_indices.append(local.astype(np.int32))
            xj, yj, zj = x[src].copy(), y[src].copy(), z[src].copy()
            typ = inp.type[src]
            h = f32(grid.gridspacing)
            topo = f_host is not None and j == 0
            block_height = f32(h * f32(grid.size.z - (2 if topo else 1)))
            lower = f32(f32(lower - block_height) + overlap)
            if len(local) and topo:
                mg = metric_grid_like(grid, SHIFT[NODE])
                z1 = grid_fill1(grid1(mg, 2))
                fi = metrics_interpolate_f_point(f_host, mg, xj, yj, inp.degree,
                                                 grid.inner_size.x, grid.inner_size.y)
                vol = typ == INPUT_VOLUME_COORD
                zj = np.where(vol, (f32(block_height) + zj).astype(f32) / fi,
                              np.where(typ == INPUT_SURFACE_COORD, z1[len(z1) - 2], zj)
                              ).astype(f32)
            elif len(local):
                if np.any((typ == INPUT_SURFACE_COORD) & (j != 0)):
                    raise ValueError("surface coordinates lie in the top block only")
                zj = np.where(typ == INPUT_VOLUME_COORD, (zj - lower).astype(f32),
                              np.where(typ == INPUT_SURFACE_COORD, block_height, zj)).astype(f32)
            overlap = f32(h * OVERLAP)
            self.xq.append(xj)
            self.yq.append(yj)
            self.zq.append(zj)
            # the grid covering interior and halo on which the coefficients live
            full = metric_grid_like(grid)
            desc = GridDesc(full)
            xs, ys, zs = (grid_fill1(grid1(full, a)) for a in range(3))
            self.grids.append(desc)
            self.coords.append((xs[:desc.size_x], ys[:desc.size_y], zs[:desc.size_z]))
        self.buffer = None
        if upload:
            self.upload()

    def host_coefficients(self, j: int):
        """cuinterp_lagrange_h of grid j's points on the host (no device):
        (err, lx, ly, lz, ix, iy, iz)."""
        from .interpolation import lagrange_h
        x, y, z = self.coords[j]
        return lagrange_h(x, y, z, self.grids[j], self.xq[j], self.yq[j], self.zq[j],
                          self.inp.degree)

    def upload(self) -> None:
        """cuinterp_init of every grid that holds points, and the step buffer."""
        from .interpolation import Interpolation
        self.interp = []
        for j in range(self.ngrids):
            if not self.lengths[j]:
                self.interp.append(None)
                continue
            x, y, z = self.coords[j]
            self.interp.append(Interpolation(x, y, z, self.grids[j], self.xq[j], self.yq[j],
                                             self.zq[j], indices=self.global_indices[j],
                                             degree=self.inp.degree))
        inp = self.inp
        self.buffer = StepBuffer(self.length, inp.gpu_buffer_size, inp.cpu_buffer_size,
                                 inp.stride)

    def close(self) -> None:
        for I in self.interp:
            if I is not None:
                I.close()


class Receiver(PointSet):
    """recv_t (receiver.c): receiver_write samples grid grid_num's points
    into this step's device record, copies device-fulls to the host after
    the last grid, and writes each host-full, transposed to receiver-major
    order, into `outputname` (num_writes host-fulls per file, each receiver's
    blocks together, mpi_io_idx_write)."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.current_write = 0

    def write(self, step: int, outputname: str, field, grid_num: int) -> bool:
        import torch
        b = self.buffer
        if self.length == 0:
            return False
        if self.lengths[grid_num] and b.is_device_ready(step):
            self.interp[grid_num].interp(b.device_slot(step), field)
        if grid_num + 1 == self.ngrids and b.is_device_full(step):
            torch.cuda.current_stream().synchronize()
            b.copy_to_host(step)
        if grid_num + 1 == self.ngrids and b.is_host_full(step):
            rows = b.num_host * b.num_device
            data = b.h_buffer.reshape(rows, b.num_elements).T
            full = np.zeros((self.inp.length, rows), dtype=f32)
            full[self.indices] = data
            if self.current_write == 0 and os.path.exists(outputname):
                os.remove(outputname)
            write_receiver_block(outputname, full, self.current_write, self.inp.num_writes)
            self.current_write = (self.current_write + 1) % self.inp.num_writes
            return True
        return False


class Source(PointSet):
    """source_t (source.c): the component's time series are read from
    `<file>_<suffix>`, each source's `steps` samples contiguous in input
    order (mpi_io_idx_read with the indexed view of source_init_indexed),
    one host-full per read, transposed to time-major and copied to the
    device a device-full at a time."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        inp = self.inp
        self.num_reads = inp.steps // (inp.cpu_buffer_size * inp.gpu_buffer_size)
        self.num_elements = inp.steps // self.num_reads if self.num_reads else 0
        self.read_no = 0

    def read(self, step: int) -> None:
        b = self.buffer
        if self.length == 0:
            return
        if b.is_host_empty(step):
            ne, nr = self.num_elements, self.num_reads
            raw = np.fromfile(self.filename, dtype=f32)
            blk = np.empty((self.length, ne), dtype=f32)
            for i, g in enumerate(self.indices):
                o = (int(g) * nr + self.read_no) * ne
                if o + ne > raw.size:
                    raise ValueError(f"{self.filename}: too short for source {g}")
                blk[i] = raw[o:o + ne]
            self.read_no += 1
            b.h_buffer[:] = blk.T.reshape(-1)
        if b.is_device_empty(step):
            b.copy_to_device(step)

    def ready(self, step: int, grid_num: int) -> bool:
        return self.length > 0 and self.buffer.is_device_ready(step) and \
            self.lengths[grid_num] > 0

    def add_cartesian(self, out, step: int, h: float, dt: float, grid_num: int) -> None:
        if self.ready(step, grid_num):
            self.interp[grid_num].add_cartesian(out, self.buffer.device_slot(step), h, dt)

    def add_curvilinear(self, out, step: int, h: float, dt: float, f, ny: int, dg,
                        grid_num: int) -> None:
        if self.ready(step, grid_num):
            self.interp[grid_num].add_curvilinear(out, self.buffer.device_slot(step), h, dt, f,
                                                  ny, dg)

    def add_force(self, out, d1, step: int, h: float, dt: float, quad_weight: float, f,
                  nx: int, ny: int, nz: int, dg, grid_num: int) -> None:
        if self.ready(step, grid_num):
            self.interp[grid_num].add_force(out, self.buffer.device_slot(step), d1, h, dt,
                                            quad_weight, f, nx, ny, nz, dg)


def _outputname(base: str, inp: Input, step: int, num_steps: int) -> str:
    return receiver_filename(base, inp, step, num_steps)


class Receivers:
    """receivers.c: velocity receivers u, v, w on the X, Y, Z grids, files
    `<file>_x_<step>` etc."""

    def __init__(self, filename: str, stack: list, f_host=None, directory: str = "."):
        self.inp = read_input(filename)
        self.dir = directory
        self.r = [Receiver(self.inp, s, t, stack, f_host) for s, t in (("x", X), ("y", Y),
                                                                        ("z", Z))]
        self.files: list[str] = []

    def write(self, u, v, w, step: int, num_steps: int, grid_num: int) -> None:
        for rec, fld in zip(self.r, (u, v, w)):
            name = os.path.join(self.dir, _outputname(rec.filename, self.inp, step, num_steps))
            if rec.write(step, name, fld, grid_num) and name not in self.files:
                self.files.append(name)

    def close(self) -> None:
        for r in self.r:
            r.close()


class Sgt:
    """sgt.c: strain Green tensor output.  Six receivers on the stress grids
    (files `<file>_Gxx_<step>` ..) and an unbuffered material receiver on
    the node grid (files `<file>_d1`, `_lami`, `_mui`)."""

    COMPONENTS = (("Gxx", XX), ("Gyy", YY), ("Gzz", ZZ), ("Gxy", XY), ("Gxz", XZ), ("Gyz", YZ))

    def __init__(self, filename: str, stack: list, f_host=None, directory: str = "."):
        import copy
        self.inp = read_input(filename)
        self.dir = directory
        self.g = [Receiver(self.inp, s, t, stack, f_host) for s, t in self.COMPONENTS]
        mi = copy.deepcopy(self.inp)
        mi.gpu_buffer_size = mi.cpu_buffer_size = mi.steps = mi.num_writes = 1
        self.mat = Receiver(mi, "", NODE, stack, f_host)
        self.files: list[str] = []

    def write_material_properties(self, d1, lami, mui, grid_num: int) -> None:
        """sgt_write_material_properties (sgt.c:77-92).  As in the reference,
        the file `<base>d1` receives the lami samples (its d1 argument is
        unused)."""
        del d1
        for suffix, fld in (("d1", lami), ("lami", lami), ("mui", mui)):
            name = os.path.join(self.dir, f"{self.mat.filename}{suffix}")
            if self.mat.write(0, name, fld, grid_num) and name not in self.files:
                self.files.append(name)

    def write(self, xx, yy, zz, xy, xz, yz, step: int, num_steps: int, grid_num: int) -> None:
        """sgt_write (sgt.c:94-131).  pmcl3d passes grid 0's xy for every grid
        (pmcl3d.c:1605, 2248); the caller keeps that choice."""
        for rec, fld in zip(self.g, (xx, yy, zz, xy, xz, yz)):
            name = os.path.join(self.dir, _outputname(rec.filename, self.inp, step, num_steps))
            if rec.write(step, name, fld, grid_num) and name not in self.files:
                self.files.append(name)

    def close(self) -> None:
        for r in self.g + [self.mat]:
            r.close()


class Sources:
    """sources.c: moment-tensor sources on the six stress grids, time series
    from `<file>_xx` .. `<file>_yz`."""

    COMPONENTS = (("xx", XX), ("yy", YY), ("zz", ZZ), ("xy", XY), ("xz", XZ), ("yz", YZ))

    def __init__(self, filename: str, stack: list, f_host=None):
        self.inp = read_input(filename)
        self.m = [Source(self.inp, s, t, stack, f_host) for s, t in self.COMPONENTS]

    def read(self, step: int) -> None:
        for s in self.m:
            s.read(step)

    def add_cartesian(self, fields6, step: int, h: float, dt: float, grid_num: int) -> None:
        for s, fld in zip(self.m, fields6):
            s.add_cartesian(fld, step, h, dt, grid_num)

    def add_curvilinear(self, fields6, step: int, h: float, dt: float, metrics: dict,
                        ny: int, grid_num: int) -> None:
        """sources_add_curvilinear (sources.c:73-95): xx yy zz divided by
        f_c g3_c, xy by f g3_c, xz by f_1 g3, yz by f_2 g3."""
        jac = (("f_c", "g3_c"), ("f_c", "g3_c"), ("f_c", "g3_c"), ("f", "g3_c"),
               ("f_1", "g3"), ("f_2", "g3"))
        for s, fld, (fn, gname) in zip(self.m, fields6, jac):
            s.add_curvilinear(fld, step, h, dt, metrics[fn], ny, metrics[gname], grid_num)

    def close(self) -> None:
        for s in self.m:
            s.close()


# forces_add (forces.c:55-79): the SBP boundary quadrature weight, metric f
# and g arrays of each velocity component (u, v, w)
FORCE_Q = 3.55599789310935      # nodal grid
FORCE_QH = 2.9022824945274315   # cell-centred grid
FORCE_WEIGHTS = ((FORCE_QH, "f_1", "g3_c"), (FORCE_QH, "f_2", "g3_c"), (FORCE_Q, "f_c", "g3"))


class Forces:
    """forces.c: boundary forces on the velocity grids X, Y, Z, time series
    from `<file>_fx` .. `<file>_fz`; every force must be a surface
    coordinate (forces_boundary_check)."""

    def __init__(self, filename: str, stack: list, f_host=None):
        self.inp = read_input(filename)
        self.f = [Source(self.inp, s, t, stack, f_host) for s, t in (("fx", X), ("fy", Y),
                                                                      ("fz", Z))]
        # forces_boundary_check (forces.c:35-46): its loop leaves the type of
        # the last grid-0 force that differs from the running value, so the
        # check passes when that last force is a surface coordinate
        fx = self.f[0]
        t = 1
        for k in fx.global_indices[0]:
            typ = int(self.inp.type[fx.indices[k]])
            if typ != t:
                t = typ
        if t != 1:
            raise ValueError("forces need surface coordinates (type 1): "
                             "ERR_INCOMPATIBLE_SOURCE_TYPE")

    def read(self, step: int) -> None:
        for s in self.f:
            s.read(step)

    def add(self, u, v, w, d1, step: int, h: float, dt: float, metrics: dict, nx: int,
            ny: int, nz: int, grid_num: int) -> None:
        """forces_add (forces.c:55-79): u with f_1 g3_c and weight qh, v with
        f_2 g3_c and qh, w with f_c g3 and q."""
        for s, fld, (qw, fn, gname) in zip(self.f, (u, v, w), FORCE_WEIGHTS):
            s.add_force(fld, d1, step, h, dt, float(f32(qw)), metrics[fn], nx, ny, nz,
                        metrics[gname], grid_num)

    def close(self) -> None:
        for s in self.f:
            s.close()


def leading_zeros(num_steps: int) -> int:
    return int(math.ceil(math.log10(float(num_steps)))) + 1


class PointIO:
    """The input-file point sets of one run in pmcl3d's call order
    (pmcl3d.c:1560-1608, 1987-2006, 2243-2249): sources and forces read
    their first host-full before the loop and inject every step before the
    point sources; receivers and SGT sample every grid after the free-surface
    pass (and at step 0 before the loop).  With topography, grid 0's sources
    take the curvilinear Jacobians and forces act on grid 0; otherwise every
    grid's sources are Cartesian.

    sources, receivers, forces, sgt: input-file names or None; stack: the
    grids_init dict of every grid; f_host: grid 0's metric f (topography)."""

    def __init__(self, stack: list, sources: str | None = None, receivers: str | None = None,
                 forces: str | None = None, sgt: str | None = None, f_host=None,
                 directory: str = "."):
        self.stack, self.f_host = stack, f_host
        self.sources = Sources(sources, stack, f_host) if sources else None
        self.receivers = Receivers(receivers, stack, f_host, directory) if receivers else None
        self.forces = Forces(forces, stack, f_host) if forces else None
        self.sgt = Sgt(sgt, stack, f_host, directory) if sgt else None
        self.nt = 0

    def start(self, solver, nt: int) -> None:
        self.nt = int(nt)
        if self.sgt is not None:
            for p in range(solver.ngrids):
                m = solver.levels[p].m
                self.sgt.write_material_properties(m["d1"], m["lam"], m["mu"], p)
        self._output(solver, 0)
        if self.sources is not None:
            self.sources.read(0)
        if self.forces is not None:
            self.forces.read(0)

    def inject(self, solver) -> None:
        step = solver.cur_step
        topo = getattr(solver, "topo", None)
        if self.sources is not None:
            self.sources.read(step)
        if topo is not None:
            L = solver.levels[0]
            mets = topo.metric_tensors()
            d = L.dims
            if self.forces is not None:
                self.forces.read(step)
                self.forces.add(L.f["u1"], L.f["v1"], L.f["w1"], L.m["d1"], step, L.medium.DH,
                                L.medium.DT, mets, d.nxt, d.nyt, d.nzt, 0)
            if self.sources is not None:
                self.sources.add_curvilinear(self._six(L), step, L.medium.DH, L.medium.DT, mets,
                                             d.nyt, 0)
            first = 1
        else:
            first = 0
        if self.sources is not None:
            for p in range(first, solver.ngrids):
                L = solver.levels[p]
                self.sources.add_cartesian(self._six(L), step, L.medium.DH, L.medium.DT, p)

    def output(self, solver) -> None:
        self._output(solver, solver.cur_step)

    def _output(self, solver, step: int) -> None:
        for p in range(solver.ngrids):
            f = solver.levels[p].f
            if self.receivers is not None:
                self.receivers.write(f["u1"], f["v1"], f["w1"], step, self.nt, p)
            if self.sgt is not None:
                # pmcl3d passes grid 0's xy for every grid (pmcl3d.c:1605, 2248)
                self.sgt.write(f["xx"], f["yy"], f["zz"], solver.levels[0].f["xy"], f["xz"],
                               f["yz"], step, self.nt, p)

    @staticmethod
    def _six(L):
        return tuple(L.f[n] for n in ("xx", "yy", "zz", "xy", "xz", "yz"))

    def close(self) -> None:
        for s in (self.sources, self.receivers, self.forces, self.sgt):
            if s is not None:
                s.close()
