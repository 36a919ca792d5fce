"""Helpers for the curvilinear-topography tests: direct calls of the oracle's
topography functions on host arrays, metric construction, staggered-grid
positions and linear fields in physical coordinates."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

from awp_amd.grid import ALIGN, NGSL, Dims
from awp_amd.topography import F_NAMES, G_NAMES, Topography, f_shape, g_size

ROOT = Path(__file__).resolve().parents[1]


def parse_tables(path: Path) -> dict:
    """name -> numpy array of every `static constexpr double NAME[..]` table."""
    txt = path.read_text()
    out = {}
    for m in re.finditer(r"static constexpr double (\w+)((?:\[\d+\])+) = \{(.*?)\};", txt, re.S):
        dims = [int(x) for x in re.findall(r"\d+", m.group(2))]
        nums = np.array([float(x) for x in re.findall(r"-?\d+\.\d+", m.group(3))])
        out[m.group(1)] = nums.reshape(dims)
    return out


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data


class TopoOracle:
    """Oracle metric arrays and direct kernel calls for one grid."""

    def __init__(self, lib, dims: Dims, DH: float, topo: Topography | None = None,
                 dtype=np.float64, g: np.ndarray | None = None, f: np.ndarray | None = None):
        self.lib = lib.load() if hasattr(lib, "load") else lib
        lib = self.lib
        self.d = dims
        self.DH = DH
        self.dtype = np.dtype(dtype)
        self.suf = "f32" if self.dtype == np.float32 else "f64"
        self.mf = {n: np.zeros(f_shape(dims.nxt, dims.nyt), dtype=dtype) for n in F_NAMES}
        self.mg = {n: np.zeros(g_size(dims.nzt), dtype=dtype) for n in G_NAMES}
        if f is not None:
            self.mf["f"][...] = f
        else:
            topo = topo if topo is not None else Topography.flat(dims.nxt, dims.nyt)
            self.mf["f"][...] = topo.f_input(dtype)
            getattr(lib, f"awp_oracle_topo_geom_{self.suf}")(_p(self.mf["f"]), dims.nxt,
                                                             dims.nyt, dims.nzt, DH)
        self.build_f()
        fill = 1
        if g is not None:
            self.mg["g"][...] = g
            fill = 0
        getattr(lib, f"awp_oracle_topo_metrics_g_{self.suf}")(
            _p(self.mg["g"]), _p(self.mg["g_c"]), _p(self.mg["g3"]), _p(self.mg["g3_c"]),
            dims.nzt, DH, fill)
        self.pf = (C.c_void_p * 10)(*[_p(self.mf[n]) for n in F_NAMES])
        self.pg = (C.c_void_p * 4)(*[_p(self.mg[n]) for n in G_NAMES])

    def build_f(self):
        d = self.d
        getattr(self.lib, f"awp_oracle_topo_metrics_f_{self.suf}")(
            (C.c_void_p * 10)(*[_p(self.mf[n]) for n in F_NAMES]), d.nxt, d.nyt, self.DH)

    # metric accessors in topographic coordinates
    def F(self, name, i, j):
        return self.mf[name][np.asarray(i) + NGSL + 2, np.asarray(j) + ALIGN + NGSL + 2]

    def Gz(self, name, k):
        return self.mg[name][np.asarray(k) + ALIGN]

    def velocity(self, f, m, DT, a=1.0, bounds=None):
        d = self.d
        nu = self.dtype.type(DT) / self.dtype.type(self.DH)
        bi, bj, ei, ej = bounds if bounds is not None else (0, 0, d.nxt, d.nyt)
        getattr(self.lib, f"awp_oracle_topo_vel_{self.suf}")(
            _p(f["u1"]), _p(f["v1"]), _p(f["w1"]), _p(m["dcrjx"]), _p(m["dcrjy"]),
            _p(m["dcrjz"]), self.pf, self.pg, _p(m["d1"]), _p(f["xx"]), _p(f["xy"]),
            _p(f["xz"]), _p(f["yy"]), _p(f["yz"]), _p(f["zz"]), a, float(nu), d.nxt, d.nyt,
            d.nzt, bi, bj, ei, ej)

    def stress_bounds(self):
        d = self.d
        return 14, d.nxt + 6, 14, d.nyt + 6

    def stress(self, f, m, DT):
        d = self.d
        s_i, e_i, s_j, e_j = self.stress_bounds()
        getattr(self.lib, f"awp_oracle_topo_str_{self.suf}")(
            *(_p(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                 "r5", "r6", "u1", "v1", "w1")),
            self.pf, self.pg, _p(m["lam"]), _p(m["mu"]), _p(m["qp"]), _p(m["coeff"]),
            _p(m["qs"]), _p(m["dcrjx"]), _p(m["dcrjy"]), _p(m["dcrjz"]), _p(m["vx1"]),
            _p(m["vx2"]), _p(m["ww"]), _p(m["wwo"]), d.nxt, d.nyt, d.nzt, self.DH, DT, s_i,
            e_i, s_j, e_j)


# Staggered positions in units of h (topographic coordinates): horizontal
# metric of each field, and whether its z level is a node (g) or a cell
# centre (g_c) value.
POSITION = {
    "u1": ("f_1", "c"), "v1": ("f_2", "c"), "w1": ("f_c", "n"),
    "xx": ("f_c", "c"), "yy": ("f_c", "c"), "zz": ("f_c", "c"),
    "xy": ("f", "c"), "xz": ("f_1", "n"), "yz": ("f_2", "n"),
}
OFFSET = {"f": (0.0, 0.0), "f_1": (0.0, -0.5), "f_2": (0.5, 0.0), "f_c": (0.5, -0.5)}


def physical_xyz(T: TopoOracle, field: str):
    """Physical (X, Y, Z) of every array element of a field (Z from the
    metrics: z = g(r) f(x, y) at the field's staggered position)."""
    d = T.d
    sh = d.shape
    I = np.arange(sh[0]) - (NGSL + 2)
    J = np.arange(sh[1]) - (NGSL + 2)
    K = np.arange(sh[2]) - ALIGN
    fname, zkind = POSITION[field]
    ox, oy = OFFSET[fname]
    X = (I + ox) * T.DH
    Y = (J + oy) * T.DH
    fm = T.mf[fname][I + NGSL + 2][:, J + ALIGN + NGSL + 2]
    gk = T.mg["g_c" if zkind == "c" else "g"][K + ALIGN]
    Z = fm[:, :, None] * gk[None, None, :]
    return (np.broadcast_to(X[:, None, None], sh), np.broadcast_to(Y[None, :, None], sh), Z)
