"""The input-file point sets (receivers, sources, forces, SGT) driven on the
CPU oracle, for GPU parity tests of awp_amd.pointsets.PointIO.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
The mapping of points onto the grid stack is the product's host code
(awp_amd.pointsets.PointSet, built without device memory); the Lagrange
coefficients come from the oracle's restatement of cuinterp_lagrange_h and
the sampling and injection from its restatements of cuinterp_interp_H and
cusource_add_*_H (oracle/awp_oracle_interp.cpp).  Instead of step buffers and
files, receivers record every sampled step, and sources take sample
step // stride of each series from the input's data files, which is what the
reference's buffers deliver at that step (src/topography/sources/
source.c:496-517, src/buffers/buffer.c).
"""
from __future__ import annotations

import numpy as np

from awp_amd.pointsets import FORCE_WEIGHTS, NODE, PointSet, Sgt, Sources, X, Y, Z
from awp_amd.inputs import read_input
from oracle import interp as oi


class _Set:
    def __init__(self, inp, suffix, grid_type, stack, f_host):
        self.ps = PointSet(inp, suffix, grid_type, stack, f_host, upload=False)
        self.coef = []
        for j in range(self.ps.ngrids):
            if not self.ps.lengths[j]:
                self.coef.append(None)
                continue
            x, y, z = self.ps.coords[j]
            err, lx, ly, lz, ix, iy, iz = oi.lagrange_h(x, y, z, self.ps.xq[j], self.ps.yq[j],
                                                        self.ps.zq[j], inp.degree)
            self.coef.append((lx, ly, lz, ix, iy, iz))
        self.series = None

    def sample(self, field, j, out):
        if self.coef[j] is not None:
            oi.interp(out, np.ascontiguousarray(field.reshape(-1)), self.coef[j],
                      self.ps.global_indices[j], self.ps.grids[j])

    def load_series(self):
        """Every local source's `steps` samples, in local order."""
        inp = self.ps.inp
        raw = np.fromfile(self.ps.filename, dtype=np.float32).reshape(inp.length, inp.steps)
        self.series = raw[self.ps.indices]

    def add(self, kind, out, step, h, dt, j, **kw):
        inp = self.ps.inp
        if self.coef[j] is None or step % inp.stride:
            return
        vals = np.ascontiguousarray(self.series[:, step // inp.stride])
        flat = out.reshape(-1)
        oi.source_add(kind, flat, vals, self.coef[j], self.ps.global_indices[j],
                      self.ps.grids[j], h, dt, **kw)


class OraclePointIO:
    """PointIO's call order (pmcl3d.c:1560-1608, 1987-2006, 2243-2249) on
    the oracle's host arrays.  recorded[name][step] holds the samples of
    receiver component `name` (x y z, Gxx .. Gyz) at each sampled step, in
    local receiver order; material[name] the material samples."""

    def __init__(self, stack, sources=None, receivers=None, forces=None, sgt=None,
                 f_host=None):
        self.stack = stack
        self.sources = None
        if sources:
            inp = read_input(sources)
            self.sources = [_Set(inp, s, t, stack, f_host) for s, t in Sources.COMPONENTS]
        self.receivers = None
        if receivers:
            inp = read_input(receivers)
            self.receivers = [(n, _Set(inp, n, t, stack, f_host))
                              for n, t in (("x", X), ("y", Y), ("z", Z))]
        self.forces = None
        if forces:
            inp = read_input(forces)
            self.forces = [_Set(inp, s, t, stack, f_host)
                           for s, t in (("fx", X), ("fy", Y), ("fz", Z))]
        self.sgt = None
        if sgt:
            inp = read_input(sgt)
            self.sgt = [(n, _Set(inp, n, t, stack, f_host)) for n, t in Sgt.COMPONENTS]
            self.mat = _Set(inp, "", NODE, stack, f_host)
        self.recorded: dict = {}
        self.material: dict = {}

    def start(self, solver, nt):
        if self.sgt is not None:
            buf = np.zeros(self.mat.ps.length, np.float32)
            for p in range(solver.ngrids):
                m = solver.levels[p].m
                for name, fld in (("d1", m["lam"]), ("lami", m["lam"]), ("mui", m["mu"])):
                    self.mat.sample(fld, p, buf)
                    if p + 1 == solver.ngrids:
                        self.material[name] = buf.copy()
        self._output(solver, 0)
        for group in (self.sources, self.forces):
            for s in group or ():
                s.load_series()

    def inject(self, solver):
        step = solver.cur_step
        topo = getattr(solver, "topo", None)
        first = 0
        if topo is not None:
            L = solver.levels[0]
            d, h, dt = L.dims, L.medium.DH, L.medium.DT
            mets = {**topo.mf, **topo.mg}
            if self.forces is not None:
                for s, fld, (qw, fn, gn) in zip(self.forces, ("u1", "v1", "w1"), FORCE_WEIGHTS):
                    s.add("force", L.f[fld], step, h, dt, 0, quad_weight=float(np.float32(qw)),
                          f=mets[fn], dg=mets[gn], d1=L.m["d1"], ny=d.nyt, nz=d.nzt)
            if self.sources is not None:
                jac = (("f_c", "g3_c"), ("f_c", "g3_c"), ("f_c", "g3_c"), ("f", "g3_c"),
                       ("f_1", "g3"), ("f_2", "g3"))
                for s, fld, (fn, gn) in zip(self.sources, ("xx", "yy", "zz", "xy", "xz", "yz"),
                                            jac):
                    s.add("curvilinear", L.f[fld], step, h, dt, 0, f=mets[fn], dg=mets[gn],
                          ny=d.nyt)
            first = 1
        if self.sources is not None:
            for p in range(first, solver.ngrids):
                L = solver.levels[p]
                for s, fld in zip(self.sources, ("xx", "yy", "zz", "xy", "xz", "yz")):
                    s.add("cartesian", L.f[fld], step, L.medium.DH, L.medium.DT, p)

    def output(self, solver):
        self._output(solver, solver.cur_step)

    def _output(self, solver, step):
        groups = []
        if self.receivers is not None:
            groups += [(n, s, f) for (n, s), f in zip(self.receivers, ("u1", "v1", "w1"))]
        if self.sgt is not None:
            groups += [(n, s, f) for (n, s), f in zip(self.sgt, ("xx", "yy", "zz", "xy", "xz",
                                                                "yz"))]
        for name, s, fld in groups:
            inp = s.ps.inp
            if step % inp.stride:
                continue
            buf = np.zeros(s.ps.length, np.float32)
            for p in range(solver.ngrids):
                src = solver.levels[0].f["xy"] if name == "Gxy" else solver.levels[p].f[fld]
                s.sample(src, p, buf)
            self.recorded.setdefault(name, {})[step] = buf
