"""One-GPU driver of the time-stepping path through the C ABI.

The per-step call order is pmcl3d's (awp_amd.stepping); every update runs in
libawp_hip.so.  Device memory is torch-allocated (plumbing only).  A run has
one grid, or a stack of grids coupled by the discontinuous-mesh kernels
(grid 0 on top, each lower grid 3x coarser).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .configs import PointSources, Receivers
from .grid import ALIGN, NGSL, Dims
from .media import Medium
from .plasticity import PlasticMedium
from .stepping import PmclStepOrder
from .storage import FIELDS16, FORMATS, ScaledStorage, default_check_every, initial_scales

FIELDS = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
          "r1", "r2", "r3", "r4", "r5", "r6")
# argument order of the nine-field packs of intp3d_H and swap_H
INTP_ORDER = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")
SWAP_ORDER = ("xx", "yy", "zz", "xy", "xz", "yz", "u1", "v1", "w1")


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def _as_list(x, n: int) -> list:
    if isinstance(x, (list, tuple)):
        if len(x) != n:
            raise ValueError(f"expected {n} per-grid entries, got {len(x)}")
        return list(x)
    return [x] + [None] * (n - 1)


def check_source_cells(dims: Dims, psrc: np.ndarray) -> None:
    """Source cells (tpsrc convention) must land in the padded grid."""
    idx = psrc[:, 0] + 1 + NGSL
    idy = psrc[:, 1] + 1 + NGSL
    idz = psrc[:, 2] + ALIGN - 1
    sh = dims.shape
    if (idx.min() < 0 or idx.max() >= sh[0] or idy.min() < 0 or idy.max() >= sh[1]
            or idz.min() < 0 or idz.max() >= sh[2]):
        raise ValueError("source cell outside the padded grid")


def check_receiver_cells(dims: Dims, idx: np.ndarray) -> None:
    sh = dims.shape
    if (idx[:, 0].min() < 0 or idx[:, 0].max() >= sh[0] or idx[:, 1].min() < 0
            or idx[:, 1].max() >= sh[1] or idx[:, 2].min() < 1 or idx[:, 2].max() >= sh[2]):
        raise ValueError("receiver cell outside the padded grid")


def check_dm_stack(media: Sequence[Medium]) -> None:
    """Grid p+1 must be 3x coarser than grid p with a third of its lateral
    cells (pmcl3d.c:370-377); the coupling kernels index on that."""
    for p in range(len(media) - 1):
        f, c = media[p].dims, media[p + 1].dims
        if f.nxt != 3 * c.nxt or f.nyt != 3 * c.nyt:
            raise ValueError(f"grid {p} ({f.nxt}x{f.nyt}) is not 3x grid {p + 1} "
                             f"({c.nxt}x{c.nyt})")
        if not np.isclose(media[p].DH * 3.0, media[p + 1].DH, rtol=1e-6):
            raise ValueError(f"grid {p + 1} spacing {media[p + 1].DH} is not 3x {media[p].DH}")
        if f.nzt < 9 or c.nzt < 4:
            raise ValueError("discontinuous mesh needs nzt >= 9 on fine and >= 4 on coarse grids")
        if media[p].DT != media[p + 1].DT:
            raise ValueError("all grids share one time step")


# the state awp_step reads from one set and writes to the other
STEP_FIELDS = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")


class _Level:
    """Device state of one grid: wavefield, medium, sources, plasticity."""

    def __init__(self, medium: Medium, sources: PointSources | None,
                 plastic: PlasticMedium | None, device: torch.device):
        self.dims: Dims = medium.dims
        self.medium = medium
        d = self.dims
        f32 = dict(dtype=torch.float32, device=device)
        self.f = {name: torch.zeros(d.shape, **f32) for name in FIELDS}
        self.m = {}
        for name in ("d1", "lam", "mu", "qp", "qs", "vx1", "vx2", "wwo", "coeff", "lam_mu",
                     "dcrjx", "dcrjy", "dcrjz"):
            self.m[name] = torch.from_numpy(
                np.ascontiguousarray(getattr(medium, name), dtype=np.float32)).to(device)
        self.m["ww"] = torch.from_numpy(np.ascontiguousarray(medium.ww, dtype=np.int32)).to(device)
        self.plastic = plastic
        if plastic is not None:
            for name in ("sigma2", "cohes", "phi"):
                self.m[name] = torch.from_numpy(
                    np.ascontiguousarray(getattr(plastic, name), dtype=np.float32)).to(device)
            self.f["yldfac"] = torch.ones(d.shape, **f32)
            self.f["neta"] = torch.zeros(d.shape, **f32)
        self.src = None
        if sources is not None and sources.npsrc > 0:
            psrc = sources.psrc(d)
            check_source_cells(d, psrc)
            self.src = dict(
                psrc=torch.from_numpy(psrc.reshape(-1).astype(np.int32)).to(device),
                cols=[torch.from_numpy(c.astype(np.float32)).to(device)
                      for c in sources.columns()],
                npsrc=sources.npsrc, nst=sources.nst)


class Solver(PmclStepOrder):
    """FP32 wavefield state and media, resident in HBM.

    medium: one Medium, or a list (grid 0 on top) for a discontinuous mesh.
    sources / plastic: one entry (grid 0) or a per-grid list.
    receivers: surface receivers on grid 0."""

    def __init__(self, medium: Medium | Sequence[Medium],
                 sources: PointSources | Sequence | None = None,
                 receivers: Receivers | None = None, nt_record: int = 0,
                 device: str = "cuda", variant: str = "dstrqc", NPC: int = 0,
                 plastic: PlasticMedium | Sequence | None = None, topography=None,
                 kinematic=None, pointio: dict | None = None, texture: str = "arrays",
                 fused: bool = False, dm_fused: bool = True):
        if not torch.cuda.is_available():
            raise RuntimeError("Solver needs a GPU: the HIP path has no CPU fallback")
        self.lib = _lib.load()
        media = list(medium) if isinstance(medium, (list, tuple)) else [medium]
        G = len(media)
        check_dm_stack(media)
        self.ngrids = G
        self.device = torch.device(device)
        self.variant = variant
        self.NPC = NPC
        srcs = _as_list(sources, G)
        pls = _as_list(plastic, G) if plastic is not None else [None] * G
        if plastic is not None and any(p is None for p in pls):
            raise ValueError("plasticity needs a plastic medium on every grid")
        self.plastic = pls if plastic is not None else None
        self.levels = [_Level(media[p], srcs[p], pls[p], self.device) for p in range(G)]
        # relaxation texture of the stress update: the per-cell arrays
        # (dstrqc_H), or the 8-entry table with the ww array ("table") or with
        # init_texture's parity rule ("rule") through awp_dstrqc_tex; "auto"
        # picks per grid the rule where ww follows it, else the table, else
        # the arrays (a coarse grid's top levels carry the fine grid's
        # texture, pmcl3d.c:1049-1058, so it breaks the rule there)
        if texture not in ("arrays", "table", "rule", "auto"):
            raise ValueError(f"unknown texture mode {texture!r}")
        if texture != "arrays" and variant != "dstrqc":
            raise ValueError("the texture table applies to the dstrqc variant")
        self.texture = texture
        # on a grid above a coarser one, dvelc2_H's level in the velocity
        # sweep (awp_dvelcx_dm) instead of a launch of its own
        self.dm_fused = bool(dm_fused)
        # fused: the whole NVE = 1 step of one grid in one sweep (awp_step),
        # from the state in f into a second state set; the two swap roles
        # every step
        self.fused = bool(fused)
        if self.fused and (G != 1 or plastic is not None or topography is not None
                           or variant != "dstrqc" or NPC != 0):
            raise ValueError("the fused step runs one Cartesian grid, NVE = 1, dstrqc, NPC = 0")
        for L in self.levels:
            L.texmode = "arrays"
        if texture != "arrays" or self.fused:
            from .media import texture_table
            for L in self.levels:
                if texture == "auto":
                    for mode in ("rule", "table"):
                        try:
                            tab = texture_table(L.medium, rule=mode == "rule")
                        except ValueError:
                            continue
                        L.texmode = mode
                        L.tex = (C.c_float * 24)(*tab.tolist())
                        break
                else:
                    tab = texture_table(L.medium, rule=texture == "rule")
                    # the fused step always takes the table, with ww unless the rule
                    L.texmode = "table" if texture == "arrays" else texture
                    L.tex = (C.c_float * 24)(*tab.tolist())
            if self.fused and self.levels[0].texmode == "arrays":
                raise ValueError("the fused step needs vx1, vx2, wwo as functions of ww")
        if self.fused:
            L = self.levels[0]
            self.f_next = {n: torch.empty_like(L.f[n]) for n in STEP_FIELDS}
        DH = (C.c_float * G)(*[m.DH for m in media])
        nxt = (C.c_int * G)(*[m.dims.nxt for m in media])
        nyt = (C.c_int * G)(*[m.dims.nyt for m in media])
        nzt = (C.c_int * G)(*[m.dims.nzt for m in media])
        if self.plastic is not None:
            pp = pls[0].params
            Rz, RzT = pp.rotation()
            Rzc = (C.c_float * 9)(*Rz.tolist())
            RzTc = (C.c_float * 9)(*RzT.tolist())
            fmaj, fmin = pp.fmajor, pp.fminor
        else:
            Rzc = RzTc = (C.c_float * 9)()
            fmaj = fmin = 0.0
        self.lib.SetDeviceConstValue(DH, C.c_float(media[0].DT), nxt, nyt, nzt, G, fmaj, fmin,
                                     Rzc, RzTc)
        _lib.check()
        # curvilinear topography replaces grid 0's velocity and stress updates
        # (pmcl3d.c:1661-1694, 1843-1856); the grids below it stay Cartesian
        # and couple to it through the discontinuous-mesh kernels
        self.topo = None
        if topography is not None:
            if plastic is not None:
                raise ValueError("topography runs without plasticity")
            from .topography import TopoDevice
            self.topo = TopoDevice(self.lib, self.levels[0], topography, self.device,
                                   self.stream)
            _lib.check()
        # receivers on grid 0
        self.rec = None
        d0 = self.levels[0].dims
        if receivers is not None and len(receivers.xy) > 0:
            idx = receivers.array_index(d0)
            check_receiver_cells(d0, idx)
            f32 = dict(dtype=torch.float32, device=self.device)
            self.rec = dict(idx=torch.from_numpy(idx.reshape(-1).copy()).to(self.device),
                            n=len(idx), out=torch.zeros((max(nt_record, 1), len(idx), 3), **f32))
        # kinematic finite fault on grid 0 (IFAULT = 5): replaces the point sources
        self.kinematic = None
        if kinematic is not None:
            kinematic.check(self.levels[0].dims)
            dev = dict(device=self.device)
            self.kinematic = dict(
                fault=kinematic,
                psrc=torch.from_numpy(np.ascontiguousarray(kinematic.cells, np.int32)
                                      .reshape(-1).copy()).to(self.device),
                cols=[torch.from_numpy(c).to(self.device) for c in kinematic.columns()],
                mom=torch.zeros(kinematic.n, dtype=torch.float32, **dev),
                filt=torch.zeros(kinematic.n * (kinematic.order + 1), dtype=torch.float64,
                                 **dev),
                b=None if kinematic.b is None else np.ascontiguousarray(kinematic.b, np.float64),
                a=None if kinematic.a is None else np.ascontiguousarray(kinematic.a, np.float64))
        # input-file point sets: pointio = {"nt": run length, "directory": ...,
        # "sources"/"receivers"/"forces"/"sgt": input-file names}
        self.pointio = None
        if pointio is not None:
            from .pointsets import PointIO, grids_init
            opts = dict(pointio)
            self.pointio_steps = int(opts.pop("nt"))
            stack = [grids_init(L.dims.nxt, L.dims.nyt, L.dims.nzt,
                                int(p == 0 and self.topo is not None), L.medium.DH)
                     for p, L in enumerate(self.levels)]
            f_host = self.topo.metrics()["f"] if self.topo is not None else None
            self.pointio = PointIO(stack, f_host=f_host, **opts)
        self.cur_step = 0

    # grid-0 views kept for single-grid callers
    @property
    def dims(self) -> Dims:
        return self.levels[0].dims

    @property
    def medium(self) -> Medium:
        return self.levels[0].medium

    @property
    def f(self) -> dict:
        return self.levels[0].f

    @property
    def m(self) -> dict:
        return self.levels[0].m

    @property
    def stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def grid_dims(self, p: int) -> Dims:
        return self.levels[p].dims

    def source_steps(self, p: int) -> int:
        s = self.levels[p].src
        return s["nst"] if s is not None else 0

    # -- the launchers --
    def velocity(self, p: int = 0) -> None:
        if p == 0 and self.topo is not None:
            self.topo.velocity(self.stream)
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        s_i, e_i = d.i_range
        fn = (self.lib.awp_dvelcx_dm if self.dm_fused and p < self.ngrids - 1
              else self.lib.dvelcx_H_opt)
        fn(_ptr(f["u1"]), _ptr(f["v1"]), _ptr(f["w1"]), _ptr(f["xx"]), _ptr(f["yy"]),
           _ptr(f["zz"]), _ptr(f["xy"]), _ptr(f["xz"]), _ptr(f["yz"]), _ptr(m["dcrjx"]),
           _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), _ptr(m["d1"]), d.nyt, d.nzt, self.stream, s_i,
           e_i, p, self.ngrids)

    def velocity_dm(self, p: int) -> None:
        if self.dm_fused and not (p == 0 and self.topo is not None):
            return  # done in the velocity sweep (awp_dvelcx_dm)
        # a curvilinear grid 0 has no Cartesian sweep to carry the level:
        # dvelc2_H runs on its own, as in pmcl3d
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        self.lib.dvelc2_H(*(_ptr(f[n]) for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz",
                                                 "yz")),
                          _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), _ptr(m["d1"]),
                          d.nxt, d.nyt, self.stream, p)

    def interpolate(self, p: int) -> None:
        """Coarse grid p onto the bottom of fine grid p-1 (intp3d_H)."""
        lo, hi = self.levels[p], self.levels[p - 1]
        self.lib.intp3d_H(*(_ptr(lo.f[n]) for n in INTP_ORDER),
                          *(_ptr(hi.f[n]) for n in INTP_ORDER), lo.dims.nxt, lo.dims.nyt, 0,
                          self.stream, p)

    def swap(self, p: int) -> None:
        """Fine grid p onto the top of coarse grid p+1 (swap_H)."""
        hi, lo = self.levels[p], self.levels[p + 1]
        self.lib.swap_H(*(_ptr(lo.f[n]) for n in SWAP_ORDER),
                        *(_ptr(hi.f[n]) for n in SWAP_ORDER), lo.dims.nxt, lo.dims.nyt,
                        None, None, None, None, 0, self.stream, p)

    def stress(self, p: int = 0) -> None:
        d = self.levels[p].dims
        (s_i, e_i), (s_j, e_j) = d.i_range, d.j_range
        self.stress_region(p, s_i, e_i, s_j, e_j)

    def stress_region(self, p: int, s_i: int, e_i: int, s_j: int, e_j: int) -> None:
        if p == 0 and self.topo is not None:
            # the curvilinear update covers its own bounds (topography.c:377-420)
            self.topo.stress(self.stream)
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        if s_i > e_i or s_j > e_j:
            return
        if L.texmode != "arrays":
            self.lib.awp_dstrqc_tex(
                *(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                       "r5", "r6", "u1", "v1", "w1")),
                _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["coeff"]), _ptr(m["qs"]),
                _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nyt, d.nzt, self.stream,
                _ptr(m["lam_mu"]), C.addressof(L.tex),
                _ptr(m["ww"]) if L.texmode == "table" else None,
                d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, p)
            return
        fn = self.lib.dstrqc_H if self.variant == "dstrqc" else self.lib.dstrqc_H_new
        fn(*(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                  "r5", "r6", "u1", "v1", "w1")),
           _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["coeff"]), _ptr(m["qs"]),
           _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nyt, d.nzt, self.stream,
           _ptr(m["lam_mu"]), _ptr(m["vx1"]), _ptr(m["vx2"]), _ptr(m["ww"]), _ptr(m["wwo"]),
           d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, p)

    def stress_dm(self, p: int, s_i: int, e_i: int, s_j: int, e_j: int) -> None:
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        if L.texmode != "arrays":
            self.lib.awp_dstrqc2_tex(
                *(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                       "r5", "r6", "u1", "v1", "w1")),
                _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["qs"]), _ptr(m["dcrjx"]),
                _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nxt, d.nyt, self.stream, _ptr(m["coeff"]),
                C.addressof(L.tex), _ptr(m["ww"]) if L.texmode == "table" else None,
                s_i, e_i, s_j, e_j, p)
            return
        self.lib.dstrqc2_H(*(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2",
                                                  "r3", "r4", "r5", "r6", "u1", "v1", "w1")),
                           _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["qs"]),
                           _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nxt, d.nyt,
                           self.stream, _ptr(m["coeff"]), _ptr(m["vx1"]), _ptr(m["vx2"]),
                           _ptr(m["ww"]), _ptr(m["wwo"]), s_i, e_i, s_j, e_j, p)

    def plastic_calc(self, p: int, xs: int, xe: int, ys: int, ye: int) -> None:
        L = self.levels[p]
        f, m, d = L.f, L.m, L.dims
        args = [_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")] + [
            _ptr(m["mu"]), _ptr(m["d1"]), _ptr(m["sigma2"]), _ptr(f["yldfac"]),
            _ptr(m["cohes"]), _ptr(m["phi"]), _ptr(f["neta"]), d.nzt, xs, xe, ys, ye,
            self.stream, p]
        if L.plastic.params.Tv > 0:
            self.lib.awp_drprecpc_calc_tv(*args, float(L.plastic.params.Tv))
        else:
            self.lib.drprecpc_calc_H_opt(*args)

    def plastic_app(self, p: int, xs: int, xe: int, ys: int, ye: int) -> None:
        L = self.levels[p]
        f, m, d = L.f, L.m, L.dims
        self.lib.drprecpc_app_H(*(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")),
                                _ptr(m["mu"]), _ptr(m["sigma2"]), _ptr(f["yldfac"]), d.nzt, xs,
                                xe, ys, ye, self.stream, p)

    def add_source(self, p: int, sample_step: int) -> None:
        L = self.levels[p]
        if L.src is None:
            return
        f, s = L.f, L.src
        axx, ayy, azz, axz, ayz, axy = s["cols"]
        self.lib.addsrc_H(sample_step, s["nst"], 3, _ptr(s["psrc"]), s["npsrc"], self.stream,
                          _ptr(axx), _ptr(ayy), _ptr(azz), _ptr(axz), _ptr(ayz), _ptr(axy),
                          _ptr(f["xx"]), _ptr(f["yy"]), _ptr(f["zz"]), _ptr(f["xy"]),
                          _ptr(f["yz"]), _ptr(f["xz"]), p)

    def add_kinematic(self, p: int) -> None:
        K, L = self.kinematic, self.levels[p]
        f = L.f
        # the filter is process-wide state in the library: set it for this fault
        if K["b"] is None:
            self.lib.SetDeviceFilterParameters(0, None, None)
        else:
            self.lib.SetDeviceFilterParameters(len(K["b"]) - 1, K["b"].ctypes.data,
                                               K["a"].ctypes.data)
        c = K["cols"]
        self.lib.addkinsrc_H(self.cur_step, 3, _ptr(K["psrc"]), K["fault"].n, self.stream,
                             _ptr(L.m["mu"]), *(_ptr(x) for x in c), _ptr(f["xx"]),
                             _ptr(f["yy"]), _ptr(f["zz"]), _ptr(f["xy"]), _ptr(f["yz"]),
                             _ptr(f["xz"]), _ptr(K["mom"]), _ptr(K["filt"]), p)

    def free_surface(self) -> None:
        d, f = self.dims, self.f
        s_i, e_i = d.i_range
        s_j, e_j = d.j_range
        self.lib.fstr_H(_ptr(f["zz"]), _ptr(f["xz"]), _ptr(f["yz"]), self.stream, s_i, e_i, s_j,
                        e_j)

    def record(self, slot: int) -> None:
        if self.rec is None or slot >= self.rec["out"].shape[0]:
            return
        f, r = self.f, self.rec
        self.lib.awp_receivers_gather(_ptr(f["u1"]), _ptr(f["v1"]), _ptr(f["w1"]),
                                      _ptr(r["idx"]), r["n"], _ptr(r["out"]), slot, self.stream,
                                      0)

    def start(self) -> None:
        if self.fused:
            # the second state set starts as a copy of the first: the step
            # leaves the cells it does not update as they are in both
            for n in STEP_FIELDS:
                self.f_next[n].copy_(self.levels[0].f[n])
        super().start()

    def step(self) -> None:
        if not self.fused:
            return super().step()
        # pmcl3d's NVE = 1 order on one grid: velocity, stress (with the
        # free-surface ghosts), sources, fstr, receivers
        self.cur_step += 1
        self.fused_step()
        self._sources()
        self.free_surface()
        self.record(self.cur_step - 1)
        if self.pointio is not None:
            self.pointio.output(self)

    def fused_step(self) -> None:
        """awp_step from the current state into the second set, then swap
        the two so that f holds the state after the step."""
        L = self.levels[0]
        d, f, m = L.dims, L.f, L.m
        a9 = (C.c_void_p * 9)(*[_ptr(f[n]) for n in STEP_FIELDS])
        b9 = (C.c_void_p * 9)(*[_ptr(self.f_next[n]) for n in STEP_FIELDS])
        r6 = (C.c_void_p * 6)(*[_ptr(f[n]) for n in ("r1", "r2", "r3", "r4", "r5", "r6")])
        med = (C.c_void_p * 5)(*[_ptr(m[n]) for n in ("d1", "lam", "mu", "qp", "qs")])
        self.lib.awp_step(a9, b9, r6, med, _ptr(m["coeff"]), C.addressof(L.tex),
                          _ptr(m["ww"]) if L.texmode != "rule" else None,
                          _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]),
                          _ptr(m["lam_mu"]), d.nxt, d.nyt, d.nzt, self.stream)
        for n in STEP_FIELDS:
            f[n], self.f_next[n] = self.f_next[n], f[n]

    def run(self, nt: int, check_every: int = 0) -> None:
        if self.cur_step == 0:
            self.start()
        for n in range(nt):
            self.step()
            if check_every and (n + 1) % check_every == 0:
                _lib.check()
        _lib.check()

    def traces(self) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.rec["out"].cpu().numpy() if self.rec is not None else np.zeros((0, 0, 3))

    def field(self, name: str, p: int = 0) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.levels[p].f[name].cpu().numpy()


class Solver16(ScaledStorage, Solver):
    """Single-grid NVE = 1 run with the wavefield and memory variables stored
    in 16 bits (storage "bf16" or "fp16") with per-field power-of-two scales
    (awp_amd.storage); arithmetic in binary32, medium in binary32."""

    def __init__(self, medium: Medium, sources: PointSources | None = None,
                 receivers: Receivers | None = None, nt_record: int = 0, device: str = "cuda",
                 storage: str = "bf16", check_every: int | None = None, topography=None,
                 scaling: str = "auto", texture: str = "arrays"):
        if isinstance(medium, (list, tuple)):
            raise ValueError("16-bit storage runs one grid")
        if storage not in FORMATS:
            raise ValueError(f"storage must be one of {sorted(FORMATS)}")
        # texture: as for Solver; "table" / "rule" take the Cartesian stress
        # update through awp_dstrqc_tex_s16 (the curvilinear one reads the
        # per-cell arrays)
        super().__init__(medium, sources, receivers, nt_record, device, topography=topography,
                         texture=texture)
        self.storage = storage
        self.fmt = FORMATS[storage]
        self.scaling = scaling
        self.check_every = (default_check_every(storage, scaling) if check_every is None
                            else int(check_every))
        L = self.levels[0]
        for name in FIELDS16:
            L.f[name] = torch.zeros(L.dims.shape, dtype=torch.int16, device=self.device)
        self.scales = initial_scales(storage, medium, sources, scaling)
        self._absbuf = torch.zeros(len(FIELDS16), dtype=torch.float32, device=self.device)
        self.rescales = 0
        self._sync_scales()
        if self.topo is not None:
            self.topo.bind()

    def _sync_scales(self) -> None:
        self._sc = (C.c_float * len(FIELDS16))(*[self.scales[f] for f in FIELDS16])

    def _absmax_all(self) -> dict:
        L = self.levels[0]
        n = L.dims.size
        base = self._absbuf.data_ptr()
        for idx, name in enumerate(FIELDS16):
            self.lib.awp_s16_absmax(_ptr(L.f[name]), n, self.fmt, base + 4 * idx, self.stream)
        _lib.check()
        vals = self._absbuf.cpu().numpy()
        return {name: float(vals[idx]) for idx, name in enumerate(FIELDS16)}

    def _rescale(self, name: str, shift: int) -> None:
        L = self.levels[0]
        self.lib.awp_s16_rescale(_ptr(L.f[name]), L.dims.size, self.fmt, int(shift), self.stream)

    def check_scales(self) -> list:
        changed = super().check_scales()
        if changed:
            self._sync_scales()
        return changed

    def step(self) -> None:
        super().step()
        if self.check_every and self.cur_step % self.check_every == 0:
            self.check_scales()

    def velocity(self, p: int = 0) -> None:
        if p == 0 and self.topo is not None:
            self.topo.velocity_s16(self.stream, self.fmt, self._sc)
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        s_i, e_i = d.i_range
        self.lib.awp_dvelcx_s16(*(_ptr(f[n]) for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy",
                                                       "xz", "yz")),
                                _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]),
                                _ptr(m["d1"]), d.nyt, d.nzt, self.stream, s_i, e_i, p,
                                self.ngrids, self.fmt, self._sc)

    def stress_region(self, p: int, s_i: int, e_i: int, s_j: int, e_j: int) -> None:
        if p == 0 and self.topo is not None:
            self.topo.stress_s16(self.stream, self.fmt, self._sc)
            return
        L = self.levels[p]
        d, f, m = L.dims, L.f, L.m
        if s_i > e_i or s_j > e_j:
            return
        if L.texmode != "arrays":
            self.lib.awp_dstrqc_tex_s16(
                *(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                       "r5", "r6", "u1", "v1", "w1")),
                _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["coeff"]), _ptr(m["qs"]),
                _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nyt, d.nzt, self.stream,
                _ptr(m["lam_mu"]), C.addressof(L.tex),
                _ptr(m["ww"]) if L.texmode == "table" else None,
                d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, p, self.fmt, self._sc)
            return
        self.lib.awp_dstrqc_s16(
            *(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                   "r5", "r6", "u1", "v1", "w1")),
            _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["coeff"]), _ptr(m["qs"]),
            _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nyt, d.nzt, self.stream,
            _ptr(m["lam_mu"]), _ptr(m["vx1"]), _ptr(m["vx2"]), _ptr(m["ww"]), _ptr(m["wwo"]),
            d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, p, self.fmt, self._sc)

    def add_source(self, p: int, sample_step: int) -> None:
        L = self.levels[p]
        if L.src is None:
            return
        f, s = L.f, L.src
        axx, ayy, azz, axz, ayz, axy = s["cols"]
        self.lib.awp_addsrc_s16(sample_step, s["nst"], 3, _ptr(s["psrc"]), s["npsrc"],
                                self.stream, _ptr(axx), _ptr(ayy), _ptr(azz), _ptr(axz),
                                _ptr(ayz), _ptr(axy), _ptr(f["xx"]), _ptr(f["yy"]),
                                _ptr(f["zz"]), _ptr(f["xy"]), _ptr(f["yz"]), _ptr(f["xz"]), p,
                                self.fmt, self._sc)

    def free_surface(self) -> None:
        d, f = self.dims, self.f
        s_i, e_i = d.i_range
        s_j, e_j = d.j_range
        self.lib.awp_fstr_s16(_ptr(f["zz"]), _ptr(f["xz"]), _ptr(f["yz"]), self.stream, s_i, e_i,
                              s_j, e_j, self.fmt, self._sc)

    def record(self, slot: int) -> None:
        if self.rec is None or slot >= self.rec["out"].shape[0]:
            return
        f, r = self.f, self.rec
        self.lib.awp_receivers_gather_s16(_ptr(f["u1"]), _ptr(f["v1"]), _ptr(f["w1"]),
                                          _ptr(r["idx"]), r["n"], _ptr(r["out"]), slot,
                                          self.stream, 0, self.fmt, self._sc)

    def velocity_dm(self, p):
        raise NotImplementedError("16-bit storage runs one grid")

    stress_dm = interpolate = swap = plastic_calc = plastic_app = velocity_dm

    def raw(self, name: str) -> np.ndarray:
        """Stored 16-bit values of a field (uint16)."""
        torch.cuda.synchronize(self.device)
        return self.levels[0].f[name].cpu().numpy().view(np.uint16)

    def field(self, name: str, p: int = 0) -> np.ndarray:
        L = self.levels[p]
        if name not in FIELDS16:
            return super().field(name, p)
        out = torch.empty(L.dims.shape, dtype=torch.float32, device=self.device)
        self.lib.awp_s16_decode(_ptr(L.f[name]), _ptr(out), L.dims.size, self.fmt,
                                float(self.scales[name]), self.stream)
        _lib.check()
        torch.cuda.synchronize(self.device)
        return out.cpu().numpy()
