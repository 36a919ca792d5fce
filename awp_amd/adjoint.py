"""Material gradient of a receiver misfit by the hand-written discrete adjoint.

J(lam, mu, rho) = 1/2 sum_{t, receivers, components} (d - d_obs)^2, d the FP32
receiver traces of one grid's NVE = 1 run (pmcl3d step order) and J in
binary64.  lam and mu are the compliance arrays the kernels read (1/lambda,
1/mu), rho the density, all on the padded grid; gradients in the moduli
follow from dJ/dlambda = -lam^2 dJ/dlam (moduli_gradients).

The forward keeps the full state every `checkpoint_every` steps (the memory
variables make the step irreversible); the backward sweep recomputes one
segment at a time and runs the adjoint steps of the segment in reverse
(awp_adjoint_step).  The recompute steps out of place (awp_dvelcx_oop,
awp_dstrqc_oop) from ring slot n-1 into slot n, so every velocity and stress
state of the segment stays in HBM without copies; the memory variables are
updated in place.  All of it runs in libawp_hip.so; torch holds the device
buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .grid import ALIGN, NGSL
from .media import Medium
from .solver import FIELDS, Solver

VEL = ("u1", "v1", "w1")
STRESS = ("xx", "yy", "zz", "xy", "xz", "yz")
ADJ_ORDER = STRESS + ("r1", "r2", "r3", "r4", "r5", "r6") + VEL


def _pp(ts):
    return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


class AdjointProblem:
    """One grid, NVE = 1: forward with checkpoints, misfit, adjoint sweep."""

    def __init__(self, medium: Medium, sources, receivers, nt: int, observed: np.ndarray,
                 checkpoint_every: int = 50, device: str = "cuda",
                 resident_steps: int | None = None, hbm_reserve_gib: float = 16.0):
        self.solver = Solver(medium, sources, receivers, nt_record=nt, device=device)
        if self.solver.rec is None:
            raise ValueError("the misfit needs receivers")
        self.lib = self.solver.lib
        self.dev = self.solver.device
        self.nt = int(nt)
        self.K = max(1, int(checkpoint_every))
        self.medium = medium
        d = medium.dims
        self.dims = d
        obs = np.ascontiguousarray(observed, dtype=np.float32)
        if obs.shape != tuple(self.solver.rec["out"].shape):
            raise ValueError(f"observed traces {obs.shape} != {tuple(self.solver.rec['out'].shape)}")
        self.obs = torch.from_numpy(obs).to(self.dev)
        f32 = dict(dtype=torch.float32, device=self.dev)
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.res = torch.zeros(self.obs.shape, **f64)
        self.J = torch.zeros(1, **f64)
        self.work = torch.zeros(256, **f64)
        sh = d.shape
        self.adj = {n: torch.zeros(sh, **f32) for n in FIELDS}
        self.scratch = [torch.zeros(sh, **f32) for _ in range(9)]
        self.g32 = [torch.zeros(sh, **f32) for _ in range(8)]
        self.g64 = [torch.zeros(sh, **f64) for _ in range(8)]
        l2 = (d.nxt + 20, d.nyt + 20)
        self.gl32 = torch.zeros(l2, **f32)
        self.gl64 = torch.zeros(l2, **f64)
        self.col = torch.zeros((3,) + l2, **f32)
        self.checkpoints = {}
        self.forward_done = False
        if self.solver.variant != "dstrqc":
            raise ValueError("the adjoint transposes the production stress update (dstrqc_H)")
        # The out-of-place recompute keeps the cells no kernel writes at their
        # initial zeros; a source there would accumulate in place forward but
        # not in the ring, so sources must sit in the stress update region.
        if self.solver.levels[0].src is not None:
            psrc = self.solver.levels[0].src["psrc"].cpu().numpy().reshape(-1, 3)
            (xs, xe), (ys, ye) = d.i_range, d.j_range
            ii, jj = psrc[:, 0] + 1 + NGSL, psrc[:, 1] + 1 + NGSL  # tpsrc -> array index
            kk = psrc[:, 2] + ALIGN - 1
            ok = ((ii >= xs) & (ii <= xe) & (jj >= ys) & (jj <= ye) & (kk >= ALIGN + 3)
                  & (kk <= d.ktop))
            if not ok.all():
                raise ValueError("adjoint sources must lie in the stress update region")
        # Velocity + stress states in HBM.  The last segments of the run keep
        # every state from the forward pass (slots[n - n_tail] for step n >=
        # n_tail), so the backward sweep recomputes only the segments before
        # n_tail; once those resident segments are done, slots[0..K] serve
        # as the recompute ring of the others.  resident_steps bounds the
        # tail (None: as many whole segments as the free HBM holds, less
        # hbm_reserve_gib; 0: recompute every segment).
        slot_bytes = len(VEL + STRESS) * int(np.prod(sh)) * 4
        if resident_steps is None:
            free, _ = torch.cuda.mem_get_info(self.dev)
            avail = free - int(hbm_reserve_gib * 2**30)
            resident_steps = max(0, avail // slot_bytes - 1)
        nseg = (self.nt + self.K - 1) // self.K
        tail_segs = min(nseg, max(0, int(resident_steps)) // self.K)
        # n_tail: the first step of the resident segments (a segment start)
        self.n_tail = (nseg - tail_segs) * self.K if tail_segs < nseg else 0
        if tail_segs == 0:
            self.n_tail = self.nt + 1  # nothing resident
        nslots = max(self.K + 1, self.nt - self.n_tail + 1 if tail_segs else 0)
        self.slots = [{x: torch.zeros(sh, **f32) for x in VEL + STRESS} for _ in range(nslots)]
        self.ring = self.slots[:self.K + 1]

    # -- medium --
    def set_medium(self, lam, mu, rho) -> None:
        m = self.solver.m
        for name, a in (("lam", lam), ("mu", mu), ("d1", rho)):
            t = torch.as_tensor(a)
            m[name].copy_(t.to(device=self.dev, dtype=torch.float32))
        # lam_mu = (1/lam) / (2/mu + 1/lam) at the surface (pmcl3d.c:941-952)
        d = self.dims
        top = d.ktop
        (xs, xe), (ys, ye) = d.i_range, d.j_range
        lam64 = m["lam"][xs:xe + 1, ys:ye + 1, top].double()
        mu64 = m["mu"][xs:xe + 1, ys:ye + 1, top].double()
        t_xl = 1.0 / lam64
        lm = torch.zeros_like(m["lam_mu"])
        lm[xs:xe + 1, ys:ye + 1] = (t_xl / (2.0 / mu64 + t_xl)).float()
        m["lam_mu"].copy_(lm)

    def _pointers(self):
        m = self.solver.m
        med = _pp([m["d1"], m["lam"], m["mu"], m["qp"], m["qs"], m["vx1"], m["vx2"], m["wwo"],
                   m["lam_mu"]])
        return med

    # -- forward --
    def _reset(self) -> None:
        s = self.solver
        for t in s.f.values():
            t.zero_()
        s.rec["out"].zero_()
        s.cur_step = 0

    def _save(self, n: int) -> None:
        f = self.solver.f
        self.checkpoints[n] = [f[x].clone() for x in FIELDS]

    def forward(self) -> float:
        """Checkpointed forward run and the misfit (binary64)."""
        s = self.solver
        self._reset()
        self.checkpoints = {}
        s.start()
        if self.n_tail > 0:
            self._save(0)
        for n in range(1, min(self.n_tail, self.nt + 1)):
            s.step()
            if n % self.K == 0 and n < min(self.n_tail, self.nt):
                self._save(n)
        if self.n_tail <= self.nt:
            # the resident tail: the state at n_tail into slot 0, then every
            # step out of place into the next slot, recording the receivers
            # from it
            if self.n_tail > 0:
                s.step()
            for x in VEL + STRESS:
                self.slots[0][x].copy_(s.f[x])
            rec = s.rec
            for n in range(self.n_tail + 1, self.nt + 1):
                fout = self.slots[n - self.n_tail]
                self._recompute_step(self.slots[n - 1 - self.n_tail], fout)
                if n - 1 < rec["out"].shape[0]:
                    self.lib.awp_receivers_gather(fout["u1"].data_ptr(), fout["v1"].data_ptr(),
                                                  fout["w1"].data_ptr(), rec["idx"].data_ptr(),
                                                  rec["n"], rec["out"].data_ptr(), n - 1,
                                                  s.stream, 0)
        self.lib.awp_misfit(s.rec["out"].data_ptr(), self.obs.data_ptr(), self.obs.numel(),
                            self.res.data_ptr(), self.J.data_ptr(), self.work.data_ptr(),
                            s.stream)
        _lib.check()
        self.forward_done = True
        return float(self.J.item())

    def traces(self) -> np.ndarray:
        return self.solver.traces()

    def _recompute_step(self, fin: dict, fout: dict) -> None:
        """One forward step (the pmcl3d order of Solver.step for one grid,
        NVE = 1) from state fin into fout; memory variables in place."""
        s, lib = self.solver, self.lib
        L = s.levels[0]
        d, m, st = L.dims, L.m, s.stream
        (s_i, e_i), (s_j, e_j) = d.i_range, d.j_range
        s.cur_step += 1
        p = lambda t: t.data_ptr()  # noqa: E731
        lib.awp_dvelcx_oop(*(p(fout[x]) for x in VEL), *(p(fin[x]) for x in VEL),
                           *(p(fin[x]) for x in STRESS), p(m["dcrjx"]), p(m["dcrjy"]),
                           p(m["dcrjz"]), p(m["d1"]), d.nyt, d.nzt, st, s_i, e_i, 0, 1)
        s0 = _pp([fin[x] for x in STRESS])
        lib.awp_dstrqc_oop(*(p(fout[x]) for x in STRESS), s0,
                           *(p(L.f[x]) for x in ("r1", "r2", "r3", "r4", "r5", "r6")),
                           *(p(fout[x]) for x in VEL), p(m["lam"]), p(m["mu"]), p(m["qp"]),
                           p(m["coeff"]), p(m["qs"]), p(m["dcrjx"]), p(m["dcrjy"]),
                           p(m["dcrjz"]), d.nyt, d.nzt, st, p(m["lam_mu"]), p(m["vx1"]),
                           p(m["vx2"]), p(m["ww"]), p(m["wwo"]), d.nxt, s.NPC, 0, 0, s_i, e_i,
                           s_j, e_j, 0)
        src = L.src
        nst = s.source_steps(0)
        if nst and s.cur_step < nst:
            axx, ayy, azz, axz, ayz, axy = src["cols"]
            lib.addsrc_H(s.cur_step + 1, nst, 3, p(src["psrc"]), src["npsrc"], st, p(axx),
                         p(ayy), p(azz), p(axz), p(ayz), p(axy), p(fout["xx"]), p(fout["yy"]),
                         p(fout["zz"]), p(fout["xy"]), p(fout["yz"]), p(fout["xz"]), 0)
        lib.fstr_H(p(fout["zz"]), p(fout["xz"]), p(fout["yz"]), st, s_i, e_i, s_j, e_j)

    # -- backward --
    def backward(self):
        """(dJ/dlam, dJ/dmu, dJ/drho) as binary64 device tensors on the padded
        grid (lam, mu: compliances)."""
        if not self.forward_done:
            raise RuntimeError("run forward() first")
        s = self.solver
        lib = self.lib
        st = s.stream
        for t in list(self.adj.values()) + self.scratch + self.g32 + self.g64:
            t.zero_()
        self.gl32.zero_()
        self.gl64.zero_()
        med = self._pointers()
        m = s.m
        adj = _pp([self.adj[n] for n in ADJ_ORDER])
        scratch = _pp(self.scratch)
        g32 = _pp(self.g32)
        rec = s.rec
        npts = rec["n"]
        starts = sorted(set(self.checkpoints) |
                        set(range(max(self.n_tail, 0), self.nt, self.K)))
        for n0 in reversed(starts):
            n1 = min(n0 + self.K, self.nt)
            if n0 >= self.n_tail:
                # resident from the forward pass
                ring = self.slots[n0 - self.n_tail:n1 - self.n_tail + 1]
            else:
                # recompute the segment from its checkpoint into the ring
                ring = self.ring
                for x, t in zip(FIELDS, self.checkpoints[n0]):
                    (ring[0][x] if x in ring[0] else s.f[x]).copy_(t)
                s.cur_step = n0
                for n in range(n0 + 1, n1 + 1):
                    self._recompute_step(ring[n - 1 - n0], ring[n - n0])
            for n in range(n1, n0, -1):
                res_n = self.res[n - 1]
                lib.awp_adjoint_inject(self.adj["u1"].data_ptr(), self.adj["v1"].data_ptr(),
                                       self.adj["w1"].data_ptr(), rec["idx"].data_ptr(), npts,
                                       res_n.data_ptr(), st)
                fv = _pp([ring[n - n0][x] for x in VEL])
                fs = _pp([ring[n - 1 - n0][x] for x in STRESS])
                lib.awp_adjoint_step(adj, fv, fs, med, m["ww"].data_ptr(), m["coeff"].data_ptr(),
                                     m["dcrjx"].data_ptr(), m["dcrjy"].data_ptr(),
                                     m["dcrjz"].data_ptr(), scratch, g32,
                                     self.gl32.data_ptr(), self.col.data_ptr(), st)
            for a, b in zip(self.g32, self.g64):
                lib.awp_adjoint_flush(a.data_ptr(), b.data_ptr(), a.numel(), st)
            lib.awp_adjoint_flush(self.gl32.data_ptr(), self.gl64.data_ptr(), self.gl32.numel(), st)
            _lib.check()
        # the adjoint state at step 0 complete (the last step's pending gather)
        lib.awp_adjoint_settle(adj, scratch, st)
        sh = self.dims.shape
        f64 = dict(dtype=torch.float64, device=self.dev)
        glam, gmu, grho = (torch.zeros(sh, **f64) for _ in range(3))
        lib.awp_adjoint_map(_pp(self.g64), self.gl64.data_ptr(), med, m["ww"].data_ptr(),
                            m["coeff"].data_ptr(), m["dcrjx"].data_ptr(), m["dcrjy"].data_ptr(),
                            m["dcrjz"].data_ptr(), glam.data_ptr(), gmu.data_ptr(),
                            grho.data_ptr(), st)
        _lib.check()
        return glam, gmu, grho


def moduli_gradients(lam, mu, glam, gmu):
    """dJ/dlambda, dJ/dmu from the compliance gradients (zero where the
    compliance is zero, i.e. outside the cells the kernels read)."""
    lam = torch.as_tensor(lam, dtype=torch.float64, device=glam.device)
    mu = torch.as_tensor(mu, dtype=torch.float64, device=gmu.device)
    return -lam * lam * glam, -mu * mu * gmu


class WaveformMisfit(torch.autograd.Function):
    """J = misfit(lam, mu, rho) with the adjoint as its backward.

    Inputs are float64 tensors of the padded grid shape (lam = 1/lambda,
    mu = 1/mu, rho); the problem object carries sources, receivers, the
    observed traces and the checkpoint interval."""

    @staticmethod
    def forward(ctx, lam, mu, rho, problem: AdjointProblem):
        problem.set_medium(lam.detach(), mu.detach(), rho.detach())
        J = problem.forward()
        ctx.problem = problem
        ctx.devices = (lam.device, mu.device, rho.device)
        return torch.tensor(J, dtype=torch.float64, device=lam.device)

    @staticmethod
    def backward(ctx, gJ):
        glam, gmu, grho = ctx.problem.backward()
        dl, dm, dr = ctx.devices
        g = gJ.to(torch.float64)
        return ((g.to(glam.device) * glam).to(dl), (g.to(gmu.device) * gmu).to(dm),
                (g.to(grho.device) * grho).to(dr), None)


def misfit(lam, mu, rho, problem: AdjointProblem):
    return WaveformMisfit.apply(lam, mu, rho, problem)
