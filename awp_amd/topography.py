"""Curvilinear topography of grid 0: the free-surface elevation, the
reference's binary topography file, and the device state (metric arrays and
the topo_t the launchers read).

Geometry (reference src/topography/geometry/geometry.c:188-210,
metrics/metrics.c): the grid follows z = g(r) f(x, y) with g(r) = r h on the
node levels r = 0 .. nz-2 and f = 1 + elevation / (h (nz-2)), so the top node
level sits at (nz-2) h + elevation above the bottom.  The elevation is given
on the f grid, topographic (i, j) in [-8, nx+8) x [-8, ny+8) (ngsl = 8 cells
beyond the interior on every side), as an (nx+16, ny+16) array, x first.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .grid import ALIGN, NGSL
from .refstructs import FGrid, GGrid, Topo

# Metric array names (f_grid_t / g_grid_t members without their d_ prefix).
# strain-rate arrays of the two-pass stress update (AWP_TOPO_NRATES)
TOPO_NRATES = 6
F_NAMES = ("f", "f_1", "f_2", "f_c", "f1_1", "f1_2", "f1_c", "f2_1", "f2_2", "f2_c")
G_NAMES = ("g", "g3", "g3_c", "g_c")


def f_shape(nx: int, ny: int) -> tuple[int, int]:
    """(rows, columns) of one f* metric array (metrics.c:18)."""
    return (nx + 4 + 2 * NGSL, ny + 4 + 2 * NGSL + 2 * ALIGN)


def g_size(nz: int) -> int:
    return nz + 2 * ALIGN


@dataclass
class Topography:
    elevation: np.ndarray  # (nx+16, ny+16) float64 metres, f grid [-8, n+8)^2

    @property
    def nx(self) -> int:
        return self.elevation.shape[0] - 2 * NGSL

    @property
    def ny(self) -> int:
        return self.elevation.shape[1] - 2 * NGSL

    def check(self, nx: int, ny: int) -> None:
        if self.elevation.shape != (nx + 2 * NGSL, ny + 2 * NGSL):
            raise ValueError(f"elevation {self.elevation.shape} does not cover the f grid of a "
                             f"{nx} x {ny} mesh ({nx + 2 * NGSL} x {ny + 2 * NGSL})")
        if not np.all(np.isfinite(self.elevation)):
            raise ValueError("elevation has non-finite values")

    @classmethod
    def flat(cls, nx: int, ny: int) -> "Topography":
        return cls(np.zeros((nx + 2 * NGSL, ny + 2 * NGSL)))

    @classmethod
    def gaussian_hills(cls, nx: int, ny: int, h: float, n: int = 6, seed: int = 3,
                       hmax: float = 1500.0, wmin: float = 2000.0,
                       wmax: float = 8000.0) -> "Topography":
        """Sum of n seeded Gaussian hills, heights up to hmax, widths (e-folding
        half-widths) in [wmin, wmax], centres inside the mesh.  The peak of the
        sum is clipped to hmax."""
        rng = np.random.default_rng(seed)
        x = (np.arange(nx + 2 * NGSL) - NGSL) * h
        y = (np.arange(ny + 2 * NGSL) - NGSL) * h
        X, Y = np.meshgrid(x, y, indexing="ij")
        e = np.zeros_like(X)
        for _ in range(n):
            cx = rng.uniform(0.15, 0.85) * nx * h
            cy = rng.uniform(0.15, 0.85) * ny * h
            wx = rng.uniform(wmin, wmax)
            wy = rng.uniform(wmin, wmax)
            amp = rng.uniform(0.3, 1.0) * hmax
            e += amp * np.exp(-((X - cx) / wx) ** 2 - ((Y - cy) / wy) ** 2)
        if e.max() > hmax:
            e *= hmax / e.max()
        return cls(e)

    # -- the reference's topography file (serial_reader.c:10-62) --
    @classmethod
    def read(cls, path: str | Path, nx: int, ny: int) -> "Topography":
        """int32 gnx, gny, padding, then (gnx+2p) x (gny+2p) float32, y
        fastest.  One rank reads rows and columns 0 .. n+15 of the padded
        global array into the f grid [-8, n+8) (so a file with padding p > 8
        shifts by p-8 cells, as the reference's reader does)."""
        raw = Path(path).read_bytes()
        gnx, gny, pad = np.frombuffer(raw[:12], dtype=np.int32)
        if gnx != nx or gny != ny or pad < NGSL:
            raise ValueError(f"topography file is {gnx} x {gny} (padding {pad}); "
                             f"the mesh is {nx} x {ny} and needs padding >= {NGSL}")
        gmx, gmy = gnx + 2 * pad, gny + 2 * pad
        data = np.frombuffer(raw[12:12 + 4 * gmx * gmy], dtype=np.float32)
        if data.size != gmx * gmy:
            raise ValueError("topography file is truncated")
        data = data.reshape(gmx, gmy)
        return cls(data[:nx + 2 * NGSL, :ny + 2 * NGSL].astype(np.float64))

    def write(self, path: str | Path, padding: int = NGSL) -> None:
        if padding != NGSL:
            raise ValueError("write() stores padding ngsl = 8, the layout read() restores")
        hdr = np.array([self.nx, self.ny, padding], dtype=np.int32).tobytes()
        Path(path).write_bytes(hdr + np.ascontiguousarray(self.elevation, np.float32).tobytes())

    def f_input(self, dtype=np.float32) -> np.ndarray:
        """The f metric array holding the elevation on [-8, n+8)^2 (input of
        geom_custom), zeros elsewhere."""
        f = np.zeros(f_shape(self.nx, self.ny), dtype=dtype)
        r0 = 2                      # topographic i = -8 is row i + ngsl + 2
        c0 = 2 + ALIGN              # topographic j = -8 is column j + align + ngsl + 2
        f[r0:r0 + self.nx + 2 * NGSL, c0:c0 + self.ny + 2 * NGSL] = self.elevation
        return f


# ---------------------------------------------------------------------------
def f_grid(arrays: dict, nx: int, ny: int, h: float) -> FGrid:
    """An f_grid_t over device arrays {name: tensor}, with its sizes, bounds
    and offsets as metrics_init_f sets them (src/topography/metrics/
    metrics.c:14-33)."""
    fg = FGrid()
    rows, cols = f_shape(nx, ny)
    fg.size[:] = [nx, ny, 1]
    fg.mem[:] = [rows, cols, 1]
    fg.bounds_x[:] = [-NGSL, nx + NGSL]
    fg.bounds_y[:] = [-NGSL, ny + NGSL]
    fg.bounds_stress_x[:] = [-NGSL // 2, nx + NGSL // 2]
    fg.bounds_stress_y[:] = [-NGSL // 2, ny + NGSL // 2]
    fg.offset[:] = [NGSL + 2, NGSL + ALIGN + 2, 0]
    fg.hi = float(np.float32(1.0 / h))
    fg.line, fg.slice = 1, cols
    for n in F_NAMES:
        setattr(fg, "d_" + n, arrays[n].data_ptr())
    return fg


def g_grid(arrays: dict, nz: int, h: float) -> GGrid:
    """A g_grid_t over device arrays {name: tensor} (metrics.c:204-221)."""
    gg = GGrid()
    gg.size, gg.mem, gg.line, gg.offset = nz, nz + 2 * ALIGN, 1, ALIGN
    gg.hi = float(np.float32(1.0 / h))
    for n in G_NAMES:
        setattr(gg, "d_" + n, arrays[n].data_ptr())
    return gg


# topo_t member <- (level dict, key)
_BIND = (("rho", "m", "d1"), ("lami", "m", "lam"), ("mui", "m", "mu"), ("qpi", "m", "qp"),
         ("qsi", "m", "qs"), ("ww", "m", "ww"), ("wwo", "m", "wwo"), ("vx1", "m", "vx1"),
         ("vx2", "m", "vx2"), ("coeff", "m", "coeff"), ("dcrjx", "m", "dcrjx"),
         ("dcrjy", "m", "dcrjy"), ("dcrjz", "m", "dcrjz"))
_FIELDS = ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6", "u1", "v1",
           "w1")


class TopoDevice:
    """Metric arrays in HBM, built by the library's setup kernels, and the
    topo_t of one grid-0 level (level.f / level.m tensors).  two_pass: the
    stress update runs in two passes over six strain-rate arrays
    (awp_topo_stress_interior_rates_H), else in one (topo_stress_interior_H).
    The two-pass update takes the level's relaxation texture as the solver set
    it up (level.texmode "rule" or "table" with the 8-entry level.tex:
    awp_topo_stress_interior_tex_H); the one-pass update reads the arrays."""

    def __init__(self, lib, level, topo: Topography, device, stream: int,
                 two_pass: bool = True):
        import torch
        d = level.dims
        topo.check(d.nxt, d.nyt)
        self.lib = lib
        self.level = level
        # scratch of the two-pass stress update: AWP_TOPO_NRATES strain-rate
        # arrays
        self.rates = (torch.empty((TOPO_NRATES,) + tuple(d.shape), dtype=torch.float32,
                                  device=device)
                      if two_pass else None)
        f32 = dict(dtype=torch.float32, device=device)
        self.mf = {n: torch.zeros(f_shape(d.nxt, d.nyt), **f32) for n in F_NAMES}
        self.mg = {n: torch.zeros(g_size(d.nzt), **f32) for n in G_NAMES}
        self.mf["f"].copy_(torch.from_numpy(topo.f_input(np.float32)))
        DH, DT = float(level.medium.DH), float(level.medium.DT)
        self.fg = f_grid(self.mf, d.nxt, d.nyt, DH)
        self.gg = g_grid(self.mg, d.nzt, DH)
        lib.awp_topo_geom_custom(self.mf["f"].data_ptr(), d.nxt, d.nyt, d.nzt, DH, stream)
        lib.awp_topo_metrics_f(C.byref(self.fg), d.nxt, d.nyt, DH, stream)
        lib.awp_topo_geom_no_stretching(self.mg["g"].data_ptr(), d.nzt, DH, stream)
        lib.awp_topo_metrics_g(C.byref(self.gg), d.nzt, DH, stream)
        self.T = Topo()
        T = self.T
        T.use = 1
        T.x_rank_l = T.x_rank_r = T.y_rank_b = T.y_rank_f = -1
        T.nx, T.ny, T.nz = d.nxt, d.nyt, d.nzt
        T.px = T.py = 1
        T.mx, T.my, T.mz = d.nxt + 4 + 2 * NGSL, d.nyt + 4 + 2 * NGSL, d.nzt + 2 * ALIGN
        T.line, T.slice = d.yline, d.slice
        T.gridsize = d.size
        # topo_t.dth = dt/h in _prec, timestep = 1 (topography.c:71-72)
        T.dth = float(np.float32(DT) / np.float32(DH))
        T.timestep = 1.0
        T.gridspacing = DH
        T.metrics_f = self.fg
        T.metrics_g = self.gg
        self.bind()
        lib.topo_set_constants(C.byref(T))

    def bind(self) -> None:
        """(Re)point the topo_t at the level's current tensors."""
        L, T = self.level, self.T
        for member, where, key in _BIND:
            setattr(T, member, getattr(L, where)[key].data_ptr())
        for name in _FIELDS:
            setattr(T, name, L.f[name].data_ptr())

    def velocity(self, stream: int) -> None:
        self.T.stream_i = stream
        self.lib.topo_velocity_interior_H(C.byref(self.T))

    def stress(self, stream: int) -> None:
        self.T.stream_i = stream
        mode = getattr(self.level, "texmode", "arrays")
        if self.rates is not None and mode in ("rule", "table"):
            self.lib.awp_topo_stress_interior_tex_H(C.byref(self.T), self.rates.data_ptr(),
                                                    C.addressof(self.level.tex),
                                                    int(mode == "rule"))
        elif self.rates is not None:
            self.lib.awp_topo_stress_interior_rates_H(C.byref(self.T), self.rates.data_ptr())
        else:
            self.lib.topo_stress_interior_H(C.byref(self.T))

    def velocity_s16(self, stream: int, fmt: int, scales) -> None:
        self.T.stream_i = stream
        self.lib.awp_topo_velocity_s16(C.byref(self.T), fmt, scales)

    def stress_s16(self, stream: int, fmt: int, scales) -> None:
        self.T.stream_i = stream
        self.lib.awp_topo_stress_s16(C.byref(self.T), fmt, scales,
                                     self.rates.data_ptr() if self.rates is not None else None)

    def metric_tensors(self) -> dict:
        """The metric arrays on the device, by name (f .. f2_c, g .. g_c)."""
        return {**self.mf, **self.mg}

    def metrics(self) -> dict:
        """Host copies of the metric arrays."""
        import torch
        torch.cuda.synchronize()
        out = {n: t.cpu().numpy() for n, t in self.mf.items()}
        out.update({n: t.cpu().numpy() for n, t in self.mg.items()})
        return out
