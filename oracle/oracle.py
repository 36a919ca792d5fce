"""Python driver of the CPU oracle (oracle/awp_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

OracleSolver mirrors awp_amd.solver.Solver on host numpy arrays in float32
(the reference's single-precision arithmetic) or float64 (the FP64 oracle),
in the pmcl3d per-step order (reference src/awp/pmcl3d.c:1628-2487).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from awp_amd.stepping import PmclStepOrder

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libawp_oracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        P, I, D = C.c_void_p, C.c_int, C.c_double
        for suf in ("f32", "f64"):
            getattr(_lib, f"awp_oracle_dvelcx_{suf}").argtypes = [P] * 13 + [I, I, I, D, D, I, I, I, I]
            getattr(_lib, f"awp_oracle_dstrqc_{suf}").argtypes = (
                [P] * 28 + [I, I, I, D, D] + [I] * 9)
            getattr(_lib, f"awp_oracle_fstr_{suf}").argtypes = [P] * 3 + [I] * 7
            getattr(_lib, f"awp_oracle_addsrc_{suf}").argtypes = (
                [I, I, I, P, I] + [P] * 12 + [I, I, I, D, D])
        for suf in ("f32", "f64"):
            getattr(_lib, f"awp_oracle_drprecpc_calc_{suf}").argtypes = (
                [P] * 13 + [I, I, I, D, D, I, I, I, I, P, D])
            getattr(_lib, f"awp_oracle_drprecpc_app_{suf}").argtypes = (
                [P] * 8 + [I, I, I, D, I, I, I, I, P])
        for suf in ("f32", "f64"):
            getattr(_lib, f"awp_oracle_dvelc2_{suf}").argtypes = [P] * 13 + [I, I, I, D, D]
            getattr(_lib, f"awp_oracle_dstrqc2_{suf}").argtypes = (
                [P] * 27 + [I, I, I, D, D, I, I, I, I])
            getattr(_lib, f"awp_oracle_intp3d_{suf}").argtypes = [P, P] + [I] * 6
            getattr(_lib, f"awp_oracle_swap3_{suf}").argtypes = [P, P] + [I] * 6
        _lib.awp_oracle_rotate_principal_f64.argtypes = [D, D, P, P]
        _lib.awp_oracle_dstrqc_parts_f32.argtypes = [P] * 28 + [I, I, I, D, D] + [I] * 10
        _lib.awp_oracle_s16_encode.argtypes = [P, P, C.c_longlong, I, C.c_float]
        _lib.awp_oracle_s16_decode.argtypes = [P, P, C.c_longlong, I, C.c_float]
        _lib.awp_oracle_dstrqc_new_f32.argtypes = [P] * 28 + [I, I, I, D, D] + [I] * 9
        for suf in ("f32", "f64"):
            T = C.c_float if suf == "f32" else C.c_double
            getattr(_lib, f"awp_oracle_topo_geom_{suf}").argtypes = [P, I, I, I, D]
            getattr(_lib, f"awp_oracle_topo_metrics_f_{suf}").argtypes = [P, I, I, D]
            getattr(_lib, f"awp_oracle_topo_metrics_g_{suf}").argtypes = [P] * 4 + [I, D, I]
            getattr(_lib, f"awp_oracle_topo_vel_{suf}").argtypes = (
                [P] * 15 + [T, T] + [I] * 7)
            getattr(_lib, f"awp_oracle_topo_str_{suf}").argtypes = (
                [P] * 29 + [I, I, I, D, D, I, I, I, I])
        _lib.awp_oracle_set_threads.argtypes = [I]
        _lib.awp_oracle_get_threads.restype = I
    return _lib


def set_threads(n: int) -> None:
    load().awp_oracle_set_threads(int(n))


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


FIELDS = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
          "r1", "r2", "r3", "r4", "r5", "r6")
# the oracle's nine-field pack order for the grid coupling
PACK9 = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")


def _as_list(x, n):
    if isinstance(x, (list, tuple)):
        if len(x) != n:
            raise ValueError(f"expected {n} per-grid entries, got {len(x)}")
        return list(x)
    return [x] + [None] * (n - 1)


class _Level:
    def __init__(self, medium, sources, plastic, dtype):
        self.medium = medium
        self.dims = d = medium.dims
        self.f = {n: np.zeros(d.shape, dtype=dtype) for n in FIELDS}
        self.m = {}
        for n in ("d1", "lam", "mu", "qp", "qs", "vx1", "vx2", "wwo", "coeff", "lam_mu",
                  "dcrjx", "dcrjy", "dcrjz"):
            self.m[n] = np.ascontiguousarray(getattr(medium, n), dtype=dtype)
        self.m["ww"] = np.ascontiguousarray(medium.ww, dtype=np.int32)
        self.DH, self.DT = float(medium.DH), float(medium.DT)
        self.src = None
        if sources is not None and sources.npsrc > 0:
            self.src = dict(psrc=np.ascontiguousarray(sources.psrc(d).reshape(-1), dtype=np.int32),
                            cols=[np.ascontiguousarray(c, dtype=dtype) for c in sources.columns()],
                            npsrc=sources.npsrc, nst=sources.nst)
        self.plastic = plastic
        if plastic is not None:
            for n in ("sigma2", "cohes", "phi"):
                self.m[n] = np.ascontiguousarray(getattr(plastic, n), dtype=dtype)
            self.f["yldfac"] = np.ones(d.shape, dtype=dtype)
            self.f["neta"] = np.zeros(d.shape, dtype=dtype)
            self.pl = np.ascontiguousarray(plastic.params.packed())

    def pack(self, names=PACK9):
        arr = (C.c_void_p * 9)(*[_p(self.f[n]) for n in names])
        return arr


class TopoHost:
    """Metric arrays of grid 0 built by the oracle's restatement of
    geom_custom, metrics_build_f and metrics_build_g."""

    def __init__(self, lib, topo, dims, DH: float, DT: float, dtype):
        from awp_amd.topography import F_NAMES, G_NAMES, f_shape, g_size
        topo.check(dims.nxt, dims.nyt)
        self.lib = lib
        self.suf = "f32" if np.dtype(dtype) == np.float32 else "f64"
        self.dims = dims
        self.DH, self.DT = DH, DT
        self.mf = {n: np.zeros(f_shape(dims.nxt, dims.nyt), dtype=dtype) for n in F_NAMES}
        self.mg = {n: np.zeros(g_size(dims.nzt), dtype=dtype) for n in G_NAMES}
        self.mf["f"][...] = topo.f_input(dtype)
        getattr(lib, f"awp_oracle_topo_geom_{self.suf}")(_p(self.mf["f"]), dims.nxt, dims.nyt,
                                                         dims.nzt, DH)
        getattr(lib, f"awp_oracle_topo_metrics_f_{self.suf}")(
            (C.c_void_p * 10)(*[_p(self.mf[n]) for n in F_NAMES]), dims.nxt, dims.nyt, DH)
        getattr(lib, f"awp_oracle_topo_metrics_g_{self.suf}")(
            _p(self.mg["g"]), _p(self.mg["g_c"]), _p(self.mg["g3"]), _p(self.mg["g3_c"]),
            dims.nzt, DH, 1)
        self.pf = (C.c_void_p * 10)(*[_p(self.mf[n]) for n in F_NAMES])
        self.pg = (C.c_void_p * 4)(*[_p(self.mg[n]) for n in G_NAMES])
        # topo_t.dth = dt/h and timestep = 1 in _prec (topography.c:71-72)
        T = np.float32 if self.suf == "f32" else np.float64
        self.nu = float(T(DT) / T(DH))

    def velocity(self, f, m):
        d = self.dims
        getattr(self.lib, f"awp_oracle_topo_vel_{self.suf}")(
            _p(f["u1"]), _p(f["v1"]), _p(f["w1"]), _p(m["dcrjx"]), _p(m["dcrjy"]),
            _p(m["dcrjz"]), self.pf, self.pg, _p(m["d1"]), _p(f["xx"]), _p(f["xy"]),
            _p(f["xz"]), _p(f["yy"]), _p(f["yz"]), _p(f["zz"]), 1.0, self.nu, d.nxt, d.nyt,
            d.nzt, 0, 0, d.nxt, d.nyt)

    def stress_bounds(self):
        """Array-index stress bounds of one rank without neighbours
        (topography.c:377-420 plus the ngsl+2 shift)."""
        d = self.dims
        half, shift = NGSL_ORACLE // 2, NGSL_ORACLE + 2
        return half + shift, d.nxt - half + shift, half + shift, d.nyt - half + shift

    def stress(self, f, m):
        d = self.dims
        s_i, e_i, s_j, e_j = self.stress_bounds()
        getattr(self.lib, f"awp_oracle_topo_str_{self.suf}")(
            *(_p(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                 "r5", "r6", "u1", "v1", "w1")),
            self.pf, self.pg, _p(m["lam"]), _p(m["mu"]), _p(m["qp"]), _p(m["coeff"]),
            _p(m["qs"]), _p(m["dcrjx"]), _p(m["dcrjy"]), _p(m["dcrjz"]), _p(m["vx1"]),
            _p(m["vx2"]), _p(m["ww"]), _p(m["wwo"]), d.nxt, d.nyt, d.nzt, self.DH, self.DT,
            s_i, e_i, s_j, e_j)


NGSL_ORACLE = 8


class OracleSolver(PmclStepOrder):
    """medium: one Medium or a per-grid list (grid 0 on top); sources and
    plastic: one entry (grid 0) or per-grid lists; receivers on grid 0;
    topography: curvilinear grid 0 (no plasticity), above any Cartesian
    grids of a discontinuous mesh."""

    def __init__(self, medium, sources=None, receivers=None, nt_record: int = 0,
                 dtype=np.float32, variant: str = "dstrqc", NPC: int = 0, plastic=None,
                 topography=None, kinematic=None):
        self.lib = load()
        self.dtype = np.dtype(dtype)
        self.suf = "f32" if self.dtype == np.float32 else "f64"
        if variant != "dstrqc" and self.suf != "f32":
            raise ValueError("dstrqc_new literal variant exists in float32 only")
        self.variant = variant
        self.NPC = NPC
        media = list(medium) if isinstance(medium, (list, tuple)) else [medium]
        G = self.ngrids = len(media)
        srcs = _as_list(sources, G)
        pls = _as_list(plastic, G) if plastic is not None else [None] * G
        self.plastic = pls if plastic is not None else None
        self.levels = [_Level(media[p], srcs[p], pls[p], self.dtype) for p in range(G)]
        d = self.levels[0].dims
        self.rec = None
        if receivers is not None and len(receivers.xy) > 0:
            self.rec = dict(idx=receivers.array_index(d),
                            out=np.zeros((max(nt_record, 1), len(receivers.xy), 3),
                                         dtype=self.dtype))
        self.cur_step = 0
        self.topo = None
        if topography is not None:
            if plastic is not None:
                raise ValueError("topography runs without plasticity")
            L0 = self.levels[0]
            self.topo = TopoHost(self.lib, topography, L0.dims, L0.DH, L0.DT, self.dtype)
        self.kinematic = None
        if kinematic is not None:
            if self.dtype != np.float32:
                raise ValueError("the kinematic-source oracle is single precision")
            kinematic.check(self.levels[0].dims)
            self.kinematic = dict(
                fault=kinematic, psrc=np.ascontiguousarray(kinematic.cells, np.int32),
                cols=kinematic.columns(), mom=np.zeros(kinematic.n, np.float32),
                filt=np.zeros(kinematic.n * (kinematic.order + 1), np.float64))

    @property
    def dims(self):
        return self.levels[0].dims

    @property
    def medium(self):
        return self.levels[0].medium

    @property
    def f(self):
        return self.levels[0].f

    @property
    def m(self):
        return self.levels[0].m

    def grid_dims(self, p):
        return self.levels[p].dims

    def source_steps(self, p):
        s = self.levels[p].src
        return s["nst"] if s is not None else 0

    def velocity(self, p: int = 0) -> None:
        if p == 0 and self.topo is not None:
            self.topo.velocity(self.levels[0].f, self.levels[0].m)
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        s_i, e_i = d.i_range
        getattr(self.lib, f"awp_oracle_dvelcx_{self.suf}")(
            *(_p(f[n]) for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")),
            _p(m["dcrjx"]), _p(m["dcrjy"]), _p(m["dcrjz"]), _p(m["d1"]), d.nxt, d.nyt, d.nzt,
            L.DH, L.DT, s_i, e_i, p, self.ngrids)

    def velocity_dm(self, p: int) -> None:
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        getattr(self.lib, f"awp_oracle_dvelc2_{self.suf}")(
            *(_p(f[n]) for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")),
            _p(m["dcrjx"]), _p(m["dcrjy"]), _p(m["dcrjz"]), _p(m["d1"]), d.nxt, d.nyt, d.nzt,
            L.DH, L.DT)

    def interpolate(self, p: int) -> None:
        lo, hi = self.levels[p], self.levels[p - 1]
        getattr(self.lib, f"awp_oracle_intp3d_{self.suf}")(
            lo.pack(), hi.pack(), lo.dims.nxt, lo.dims.nyt, lo.dims.nzt, hi.dims.nxt,
            hi.dims.nyt, hi.dims.nzt)

    def swap(self, p: int) -> None:
        hi, lo = self.levels[p], self.levels[p + 1]
        getattr(self.lib, f"awp_oracle_swap3_{self.suf}")(
            lo.pack(), hi.pack(), lo.dims.nxt, lo.dims.nyt, lo.dims.nzt, hi.dims.nxt,
            hi.dims.nyt, hi.dims.nzt)

    def stress(self, p: int = 0) -> None:
        d = self.levels[p].dims
        (s_i, e_i), (s_j, e_j) = d.i_range, d.j_range
        self.stress_region(p, s_i, e_i, s_j, e_j)

    def stress_region(self, p, s_i, e_i, s_j, e_j) -> None:
        if p == 0 and self.topo is not None:
            self.topo.stress(self.levels[0].f, self.levels[0].m)
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        if s_i > e_i or s_j > e_j:
            return
        name = (f"awp_oracle_dstrqc_{self.suf}" if self.variant == "dstrqc"
                else "awp_oracle_dstrqc_new_f32")
        getattr(self.lib, name)(
            *(_p(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                 "r5", "r6", "u1", "v1", "w1")),
            _p(m["lam"]), _p(m["mu"]), _p(m["qp"]), _p(m["coeff"]), _p(m["qs"]),
            _p(m["dcrjx"]), _p(m["dcrjy"]), _p(m["dcrjz"]), _p(m["lam_mu"]), _p(m["vx1"]),
            _p(m["vx2"]), _p(m["ww"]), _p(m["wwo"]), d.nxt, d.nyt, d.nzt, L.DH, L.DT,
            d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, p)

    def stress_dm(self, p, s_i, e_i, s_j, e_j) -> None:
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        getattr(self.lib, f"awp_oracle_dstrqc2_{self.suf}")(
            *(_p(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                 "r5", "r6", "u1", "v1", "w1")),
            _p(m["lam"]), _p(m["mu"]), _p(m["qp"]), _p(m["qs"]), _p(m["dcrjx"]),
            _p(m["dcrjy"]), _p(m["dcrjz"]), _p(m["coeff"]), _p(m["vx1"]), _p(m["vx2"]),
            _p(m["ww"]), _p(m["wwo"]), d.nxt, d.nyt, d.nzt, L.DH, L.DT, s_i, e_i, s_j, e_j)

    def plastic_calc(self, p, xs, xe, ys, ye) -> None:
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        getattr(self.lib, f"awp_oracle_drprecpc_calc_{self.suf}")(
            *(_p(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")), _p(m["mu"]), _p(m["d1"]),
            _p(m["sigma2"]), _p(f["yldfac"]), _p(m["cohes"]), _p(m["phi"]), _p(f["neta"]),
            d.nxt, d.nyt, d.nzt, L.DH, L.DT, xs, xe, ys, ye, _p(L.pl),
            float(L.plastic.params.Tv))

    def plastic_app(self, p, xs, xe, ys, ye) -> None:
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        getattr(self.lib, f"awp_oracle_drprecpc_app_{self.suf}")(
            *(_p(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")), _p(m["sigma2"]),
            _p(f["yldfac"]), d.nxt, d.nyt, d.nzt, L.DH, xs, xe, ys, ye, _p(L.pl))

    def add_source(self, p: int, sample_step: int) -> None:
        L = self.levels[p]
        if L.src is None:
            return
        f, s, d = L.f, L.src, L.dims
        getattr(self.lib, f"awp_oracle_addsrc_{self.suf}")(
            sample_step, s["nst"], 3, _p(s["psrc"]), s["npsrc"], *(_p(c) for c in s["cols"]),
            _p(f["xx"]), _p(f["yy"]), _p(f["zz"]), _p(f["xy"]), _p(f["yz"]), _p(f["xz"]),
            d.nxt, d.nyt, d.nzt, L.DH, L.DT)

    def add_kinematic(self, p: int) -> None:
        from oracle import kin
        K, L = self.kinematic, self.levels[p]
        f, F = L.f, K["fault"]
        kin.addkinsrc(self.cur_step, K["psrc"], L.m["mu"], K["cols"][:5],
                      [f[n] for n in ("xx", "yy", "zz", "xy", "yz", "xz")], K["mom"],
                      K["filt"] if F.order else None, F.b, F.a, np.float32(L.DT),
                      np.float32(L.DH), L.dims)

    def free_surface(self) -> None:
        d, f = self.dims, self.f
        s_i, e_i = d.i_range
        s_j, e_j = d.j_range
        getattr(self.lib, f"awp_oracle_fstr_{self.suf}")(
            _p(f["zz"]), _p(f["xz"]), _p(f["yz"]), d.nxt, d.nyt, d.nzt, s_i, e_i, s_j, e_j)

    def record(self, slot: int) -> None:
        if self.rec is None or slot >= self.rec["out"].shape[0]:
            return
        idx = self.rec["idx"]
        f = self.f
        i, j, k = idx[:, 0], idx[:, 1], idx[:, 2]
        out = self.rec["out"][slot]
        out[:, 0] = f["u1"][i, j, k]
        out[:, 1] = f["v1"][i, j, k]
        out[:, 2] = f["w1"][i, j, k - 1]

    def run(self, nt: int) -> None:
        if self.cur_step == 0:
            self.start()
        for _ in range(nt):
            self.step()

    def traces(self) -> np.ndarray:
        return self.rec["out"] if self.rec is not None else np.zeros((0, 0, 3))


class OracleSolver16(OracleSolver):
    """The 16-bit storage mode (awp_amd.storage) on the oracle: the state is
    the stored 16-bit values; every launch decodes them to binary32, runs the
    FP32 oracle function and encodes the fields the launch writes.  The
    free-surface ghosts are a pass of their own before the stress sweep, as
    on the GPU, so the sweep reads the stored ghosts."""

    def __init__(self, medium, sources=None, receivers=None, nt_record: int = 0,
                 storage: str = "bf16", check_every: int | None = None, topography=None,
                 scaling: str = "auto"):
        from awp_amd.storage import FIELDS16, FORMATS, default_check_every, initial_scales
        super().__init__(medium, sources, receivers, nt_record, dtype=np.float32,
                         topography=topography)
        self.storage = storage
        self.fmt = FORMATS[storage]
        self.scaling = scaling
        self.check_every = (default_check_every(storage, scaling) if check_every is None
                            else int(check_every))
        self.fields16 = FIELDS16
        L = self.levels[0]
        L.q = {n: np.zeros(L.dims.shape, dtype=np.uint16) for n in FIELDS16}
        self.scales = initial_scales(storage, medium, sources, scaling)
        self.rescales = 0

    # codecs
    def _dec(self, names):
        L = self.levels[0]
        for n in names:
            self.lib.awp_oracle_s16_decode(_p(L.q[n]), _p(L.f[n]), L.f[n].size, self.fmt,
                                           float(self.scales[n]))

    def _enc(self, names):
        L = self.levels[0]
        for n in names:
            self.lib.awp_oracle_s16_encode(_p(L.f[n]), _p(L.q[n]), L.f[n].size, self.fmt,
                                           float(1.0 / self.scales[n]))

    def _absmax_all(self):
        L = self.levels[0]
        out = {}
        tmp = np.empty(L.dims.shape, dtype=np.float32)
        for n in self.fields16:
            self.lib.awp_oracle_s16_decode(_p(L.q[n]), _p(tmp), tmp.size, self.fmt, 1.0)
            out[n] = float(np.abs(tmp).max())
        return out

    def _rescale(self, name, shift):
        L = self.levels[0]
        tmp = np.empty(L.dims.shape, dtype=np.float32)
        self.lib.awp_oracle_s16_decode(_p(L.q[name]), _p(tmp), tmp.size, self.fmt,
                                       float(2.0 ** shift))
        self.lib.awp_oracle_s16_encode(_p(tmp), _p(L.q[name]), tmp.size, self.fmt, 1.0)

    def check_scales(self):
        from awp_amd.storage import ScaledStorage
        return ScaledStorage.check_scales(self)

    def step(self):
        super().step()
        if self.check_every and self.cur_step % self.check_every == 0:
            self.check_scales()

    # launches
    def velocity(self, p=0):
        self._dec(self.fields16[:9])
        super().velocity(p)
        self._enc(("u1", "v1", "w1"))

    def stress_region(self, p, s_i, e_i, s_j, e_j):
        if p == 0 and self.topo is not None:
            self._dec(self.fields16)
            self.topo.stress(self.levels[0].f, self.levels[0].m)
            self._enc(("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6"))
            return
        if s_i > e_i or s_j > e_j:
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        names = ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6", "u1",
                 "v1", "w1")
        for parts, out in ((1, ("u1", "v1", "w1")),
                           (2, ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                "r5", "r6"))):
            self._dec(self.fields16)
            self.lib.awp_oracle_dstrqc_parts_f32(
                *(_p(f[n]) for n in names), _p(m["lam"]), _p(m["mu"]), _p(m["qp"]),
                _p(m["coeff"]), _p(m["qs"]), _p(m["dcrjx"]), _p(m["dcrjy"]), _p(m["dcrjz"]),
                _p(m["lam_mu"]), _p(m["vx1"]), _p(m["vx2"]), _p(m["ww"]), _p(m["wwo"]), d.nxt,
                d.nyt, d.nzt, L.DH, L.DT, d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, p, parts)
            self._enc(out)

    def add_source(self, p, sample_step):
        if self.levels[p].src is None:
            return
        self._dec(("xx", "yy", "zz", "xy", "xz", "yz"))
        super().add_source(p, sample_step)
        self._enc(("xx", "yy", "zz", "xy", "xz", "yz"))

    def free_surface(self):
        self._dec(("zz", "xz", "yz"))
        super().free_surface()
        self._enc(("zz", "xz", "yz"))

    def record(self, slot):
        self._dec(("u1", "v1", "w1"))
        super().record(slot)

    def field(self, name):
        self._dec((name,))
        return self.levels[0].f[name].copy()
