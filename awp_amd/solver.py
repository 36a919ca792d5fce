"""One-GPU driver of the time-stepping path through the C ABI.

The per-step call order is pmcl3d's (awp_amd.stepping); every update runs in
libawp_hip.so.  Device memory is torch-allocated (plumbing only).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .configs import PointSources, Receivers
from .grid import ALIGN, NGSL, Dims
from .media import Medium
from .plasticity import PlasticMedium
from .stepping import PmclStepOrder

FIELDS = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
          "r1", "r2", "r3", "r4", "r5", "r6")


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class Solver(PmclStepOrder):
    """FP32 wavefield state and medium of one grid, resident in HBM."""

    def __init__(self, medium: Medium, sources: PointSources | None = None,
                 receivers: Receivers | None = None, nt_record: int = 0,
                 device: str = "cuda", variant: str = "dstrqc", NPC: int = 0,
                 plastic: PlasticMedium | None = None):
        if not torch.cuda.is_available():
            raise RuntimeError("Solver needs a GPU: the HIP path has no CPU fallback")
        self.lib = _lib.load()
        self.dims: Dims = medium.dims
        self.medium = medium
        self.device = torch.device(device)
        self.variant = variant
        self.NPC = NPC
        d = self.dims
        f32 = dict(dtype=torch.float32, device=self.device)
        self.f = {name: torch.zeros(d.shape, **f32) for name in FIELDS}
        self.m = {}
        for name in ("d1", "lam", "mu", "qp", "qs", "vx1", "vx2", "wwo"):
            self.m[name] = torch.from_numpy(
                np.ascontiguousarray(getattr(medium, name), dtype=np.float32)).to(self.device)
        self.m["ww"] = torch.from_numpy(np.ascontiguousarray(medium.ww, dtype=np.int32)).to(self.device)
        for name in ("coeff", "lam_mu", "dcrjx", "dcrjy", "dcrjz"):
            self.m[name] = torch.from_numpy(
                np.ascontiguousarray(getattr(medium, name), dtype=np.float32)).to(self.device)
        DH = (C.c_float * 1)(medium.DH)
        nxt = (C.c_int * 1)(d.nxt)
        nyt = (C.c_int * 1)(d.nyt)
        nzt = (C.c_int * 1)(d.nzt)
        self.plastic = plastic
        if plastic is not None:
            pp = plastic.params
            Rz, RzT = pp.rotation()
            Rzc = (C.c_float * 9)(*Rz.tolist())
            RzTc = (C.c_float * 9)(*RzT.tolist())
            fmaj, fmin = pp.fmajor, pp.fminor
            for name in ("sigma2", "cohes", "phi"):
                self.m[name] = torch.from_numpy(
                    np.ascontiguousarray(getattr(plastic, name), dtype=np.float32)).to(self.device)
            self.f["yldfac"] = torch.ones(d.shape, **f32)
            self.f["neta"] = torch.zeros(d.shape, **f32)
        else:
            Rzc = RzTc = (C.c_float * 9)()
            fmaj = fmin = 0.0
        self.lib.SetDeviceConstValue(DH, C.c_float(medium.DT), nxt, nyt, nzt, 1, fmaj, fmin, Rzc,
                                     RzTc)
        _lib.check()
        # sources
        self.src = None
        if sources is not None and sources.npsrc > 0:
            psrc = sources.psrc(d)
            self._check_sources(psrc)
            cols = sources.columns()
            self.src = dict(
                psrc=torch.from_numpy(psrc.reshape(-1).copy()).to(self.device),
                cols=[torch.from_numpy(c.astype(np.float32)).to(self.device) for c in cols],
                npsrc=sources.npsrc, nst=sources.nst)
        # receivers
        self.rec = None
        if receivers is not None and len(receivers.xy) > 0:
            idx = receivers.array_index(d)
            self._check_receivers(idx)
            self.rec = dict(idx=torch.from_numpy(idx.reshape(-1).copy()).to(self.device),
                            n=len(idx), out=torch.zeros((max(nt_record, 1), len(idx), 3), **f32))
        self.cur_step = 0

    # -- validation of host-built index lists before they reach a kernel --
    def _check_sources(self, psrc: np.ndarray) -> None:
        d = self.dims
        idx = psrc[:, 0] + 1 + NGSL
        idy = psrc[:, 1] + 1 + NGSL
        idz = psrc[:, 2] + ALIGN - 1
        sh = d.shape
        if (idx.min() < 0 or idx.max() >= sh[0] or idy.min() < 0 or idy.max() >= sh[1]
                or idz.min() < 0 or idz.max() >= sh[2]):
            raise ValueError("source cell outside the padded grid")

    def _check_receivers(self, idx: np.ndarray) -> None:
        sh = self.dims.shape
        if (idx[:, 0].min() < 0 or idx[:, 0].max() >= sh[0] or idx[:, 1].min() < 0
                or idx[:, 1].max() >= sh[1] or idx[:, 2].min() < 1 or idx[:, 2].max() >= sh[2]):
            raise ValueError("receiver cell outside the padded grid")

    @property
    def stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    # -- the launchers --
    def velocity(self) -> None:
        d, f, m = self.dims, self.f, self.m
        s_i, e_i = d.i_range
        self.lib.dvelcx_H_opt(_ptr(f["u1"]), _ptr(f["v1"]), _ptr(f["w1"]), _ptr(f["xx"]),
                              _ptr(f["yy"]), _ptr(f["zz"]), _ptr(f["xy"]), _ptr(f["xz"]),
                              _ptr(f["yz"]), _ptr(m["dcrjx"]), _ptr(m["dcrjy"]),
                              _ptr(m["dcrjz"]), _ptr(m["d1"]), d.nyt, d.nzt, self.stream, s_i,
                              e_i, 0, 1)

    def stress(self) -> None:
        d = self.dims
        (s_i, e_i), (s_j, e_j) = d.i_range, d.j_range
        self.stress_region(s_i, e_i, s_j, e_j)

    def stress_region(self, s_i: int, e_i: int, s_j: int, e_j: int) -> None:
        d, f, m = self.dims, self.f, self.m
        if s_i > e_i or s_j > e_j:
            return
        fn = self.lib.dstrqc_H if self.variant == "dstrqc" else self.lib.dstrqc_H_new
        fn(*(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                                  "r5", "r6", "u1", "v1", "w1")),
           _ptr(m["lam"]), _ptr(m["mu"]), _ptr(m["qp"]), _ptr(m["coeff"]), _ptr(m["qs"]),
           _ptr(m["dcrjx"]), _ptr(m["dcrjy"]), _ptr(m["dcrjz"]), d.nyt, d.nzt, self.stream,
           _ptr(m["lam_mu"]), _ptr(m["vx1"]), _ptr(m["vx2"]), _ptr(m["ww"]), _ptr(m["wwo"]),
           d.nxt, self.NPC, 0, 0, s_i, e_i, s_j, e_j, 0)

    def plastic_calc(self, xs: int, xe: int, ys: int, ye: int) -> None:
        f, m, d = self.f, self.m, self.dims
        args = [_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")] + [
            _ptr(m["mu"]), _ptr(m["d1"]), _ptr(m["sigma2"]), _ptr(f["yldfac"]),
            _ptr(m["cohes"]), _ptr(m["phi"]), _ptr(f["neta"]), d.nzt, xs, xe, ys, ye,
            self.stream, 0]
        if self.plastic.params.Tv > 0:
            self.lib.awp_drprecpc_calc_tv(*args, float(self.plastic.params.Tv))
        else:
            self.lib.drprecpc_calc_H_opt(*args)

    def plastic_app(self, xs: int, xe: int, ys: int, ye: int) -> None:
        f, m, d = self.f, self.m, self.dims
        self.lib.drprecpc_app_H(*(_ptr(f[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")),
                                _ptr(m["mu"]), _ptr(m["sigma2"]), _ptr(f["yldfac"]), d.nzt, xs,
                                xe, ys, ye, self.stream, 0)

    def add_source(self, sample_step: int) -> None:
        if self.src is None:
            return
        f, s = self.f, self.src
        axx, ayy, azz, axz, ayz, axy = s["cols"]
        self.lib.addsrc_H(sample_step, s["nst"], 3, _ptr(s["psrc"]), s["npsrc"], self.stream,
                          _ptr(axx), _ptr(ayy), _ptr(azz), _ptr(axz), _ptr(ayz), _ptr(axy),
                          _ptr(f["xx"]), _ptr(f["yy"]), _ptr(f["zz"]), _ptr(f["xy"]),
                          _ptr(f["yz"]), _ptr(f["xz"]), 0)

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

    def field(self, name: str) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.f[name].cpu().numpy()
