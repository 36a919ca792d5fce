"""The per-step call order of the pmcl3d time loop on one rank
(reference src/awp/pmcl3d.c:1628-2487), shared by every driver.

A driver supplies the launch primitives (velocity, stress_region,
add_source, plastic_calc, plastic_app, free_surface, record); this mixin
sequences them exactly as pmcl3d does:

  NVE = 1 (visco-elastic):
    dvelcx_H_opt  i in [xvs, xve]                                  :1690
    dstrqc_H      inner x strip, then the left/right strips        :1845, :1904-1919
    addsrc_H      sample cur_step, while cur_step < NST            :2007-2013
    fstr_H                                                         :2243
    receivers                                                      :2246
  NVE = 3 (visco-elasto-plastic) adds the region split:
    dstrqc_H      inner x strip outside [yls2, yre2], then left/right  :1858-1877
    addsrc_H
    drprecpc_calc front, back; dstrqc_H on the inner box; drprecpc_calc
    left, right, inside; drprecpc_app on everything               :2109-2209
    fstr_H, receivers
On one rank the NVE = 1 strips cover the stress region exactly once, so a
single dstrqc_H over all of it is the same update.  In the NVE = 3 order the
inner box's stress update runs after the source injection, and the
drprecpc_calc regions overlap (neta grows more than once in the overlaps):
both are pmcl3d's behaviour and are kept.
"""
from __future__ import annotations

from .grid import NGSL
from .plasticity import nve3_regions


class PmclStepOrder:
    cur_step: int
    plastic = None

    def start(self) -> None:
        """Initial source sample before the loop (pmcl3d.c:1175-1186)."""
        self.add_source(1)

    def _source(self) -> None:
        if self.src is not None and self.cur_step < self.src["nst"]:
            self.add_source(self.cur_step + 1)

    def step(self) -> None:
        self.cur_step += 1
        self.velocity()
        d = self.dims
        if self.plastic is None:
            (s_i, e_i), (s_j, e_j) = d.i_range, d.j_range
            self.stress_region(s_i, e_i, s_j, e_j)
            self._source()
        else:
            R = nve3_regions(d)
            self.stress_region(R["xss2"], R["xse2"], R["yls"], R["yls2"] - 1)
            self.stress_region(R["xss2"], R["xse2"], R["yre2"] + 1, R["yre"])
            self.stress_region(R["xss1"], R["xse1"], R["yls"], R["yre"])
            self.stress_region(R["xss3"], R["xse3"], R["yls"], R["yre"])
            self._source()
            xl, xr, yl, yr = R["xls"], R["xre"], R["yls"], R["yre"]
            self.plastic_calc(xl, xr, yl, yl + NGSL)
            self.plastic_calc(xl, xr, yr - NGSL, yr)
            self.stress_region(R["xss2"], R["xse2"], R["yls2"], R["yre2"])
            self.plastic_calc(xl, xl + NGSL, yl, yr)
            self.plastic_calc(xr - NGSL, xr, yl, yr)
            self.plastic_calc(xl + NGSL, xr - NGSL, yl + NGSL, yr - NGSL)
            self.plastic_app(xl, xr, yl, yr)
        self.free_surface()
        self.record(self.cur_step - 1)
