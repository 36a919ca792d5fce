"""The per-step call order of the pmcl3d time loop on one rank
(reference src/awp/pmcl3d.c:1628-2487), shared by every driver.

A driver holds ngrids grids, grid 0 on top (finest), each coarser grid 3x the
spacing of the one above it (pmcl3d.c:370-377), and supplies the launch
primitives per grid p: velocity, velocity_dm, interpolate, swap,
stress_region, stress_dm, add_source, plastic_calc, plastic_app, plus
free_surface and record on grid 0.  This mixin sequences them as pmcl3d does:

  dvelcx_H_opt   every grid, i in [xvs, xve]                          :1690
  dvelc2_H       every grid above the lowest                          :1727-1730
  intp3d_H       coarse p -> fine p-1                                  :1765-1770
  swap_H         fine p -> coarse p+1                                  :1802-1806
  dstrqc_H       NVE = 1: the whole stress region of every grid        :1836-1852
                 NVE = 3: the inner x strip outside [yls2, yre2]      :1858-1877
                 then the left/right strips of every grid              :1895-1920
  dstrqc2_H      every grid above the lowest, [xss1, xse3] x [yls, yre] :1936-1941
  input-file sources (sources_add_*), forces on a topography grid 0     :1987-2006
  addsrc_H       every grid with sources, while cur_step < NST         :2007-2015
  intp3d_H, swap_H again                                               :2045-2092
  NVE = 3: drprecpc_calc front/back; dstrqc_H on the inner box;
           drprecpc_calc left/right, inside; drprecpc_app              :2109-2209
  fstr_H on grid 0, receivers; input-file receivers and SGT output      :2243-2249
On one rank the NVE = 1 strips cover the stress region exactly once, so a
single dstrqc_H over all of it is the same update.  In the NVE = 3 order the
inner box's stress update runs after the source injection and after the
grid coupling, and the drprecpc_calc regions overlap (neta grows more than
once in the overlaps): all of it is pmcl3d's behaviour and is kept.
"""
from __future__ import annotations

from .grid import NGSL
from .plasticity import nve3_regions


class PmclStepOrder:
    cur_step: int
    ngrids: int = 1
    plastic = None  # None, or one plastic medium per grid

    kinematic = None  # a KinematicFault on grid 0 (IFAULT = 5), or None
    # input-file sources, forces, receivers and SGT output
    # (awp_amd.pointsets.PointIO), or None
    pointio = None
    pointio_steps = 0   # the run length the point-set file names are sized for

    def start(self) -> None:
        """Initial source sample before the loop (pmcl3d.c:1175-1186; only
        for point sources, IFAULT < 4); the input-file point sets sample
        step 0 and read their first source data (pmcl3d.c:1560-1608)."""
        if self.pointio is not None:
            self.pointio.start(self, self.pointio_steps)
        if self.kinematic is not None:
            return
        for p in range(self.ngrids):
            self.add_source(p, 1)

    def _sources(self) -> None:
        if self.pointio is not None:
            # sources_read / forces_add / sources_add_* (pmcl3d.c:1987-2006)
            self.pointio.inject(self)
        if self.kinematic is not None:
            # IFAULT = 5: the subfaults of grid 0 every step (pmcl3d.c:2025-2034)
            self.add_kinematic(0)
            return
        for p in range(self.ngrids):
            nst = self.source_steps(p)
            if nst and self.cur_step < nst:
                self.add_source(p, self.cur_step + 1)

    def _couple(self) -> None:
        G = self.ngrids
        for p in range(1, G):
            self.interpolate(p)
        for p in range(G - 1):
            self.swap(p)

    def step(self) -> None:
        self.cur_step += 1
        G = self.ngrids
        for p in range(G):
            self.velocity(p)
        for p in range(G - 1):
            self.velocity_dm(p)
        self._couple()
        regions = [nve3_regions(self.grid_dims(p)) for p in range(G)]
        if self.plastic is None:
            for p in range(G):
                d = self.grid_dims(p)
                (s_i, e_i), (s_j, e_j) = d.i_range, d.j_range
                self.stress_region(p, s_i, e_i, s_j, e_j)
        else:
            for p, R in enumerate(regions):
                self.stress_region(p, R["xss2"], R["xse2"], R["yls"], R["yls2"] - 1)
                self.stress_region(p, R["xss2"], R["xse2"], R["yre2"] + 1, R["yre"])
            for p, R in enumerate(regions):
                self.stress_region(p, R["xss1"], R["xse1"], R["yls"], R["yre"])
                self.stress_region(p, R["xss3"], R["xse3"], R["yls"], R["yre"])
        for p in range(G - 1):
            R = regions[p]
            self.stress_dm(p, R["xss1"], R["xse3"], R["yls"], R["yre"])
        self._sources()
        self._couple()
        if self.plastic is not None:
            for p, R in enumerate(regions):
                xl, xr, yl, yr = R["xls"], R["xre"], R["yls"], R["yre"]
                self.plastic_calc(p, xl, xr, yl, yl + NGSL)
                self.plastic_calc(p, xl, xr, yr - NGSL, yr)
            for p, R in enumerate(regions):
                self.stress_region(p, R["xss2"], R["xse2"], R["yls2"], R["yre2"])
            for p, R in enumerate(regions):
                xl, xr, yl, yr = R["xls"], R["xre"], R["yls"], R["yre"]
                self.plastic_calc(p, xl, xl + NGSL, yl, yr)
                self.plastic_calc(p, xr - NGSL, xr, yl, yr)
            for p, R in enumerate(regions):
                xl, xr, yl, yr = R["xls"], R["xre"], R["yls"], R["yre"]
                self.plastic_calc(p, xl + NGSL, xr - NGSL, yl + NGSL, yr - NGSL)
            for p, R in enumerate(regions):
                self.plastic_app(p, R["xls"], R["xre"], R["yls"], R["yre"])
        self.free_surface()
        self.record(self.cur_step - 1)
        if self.pointio is not None:
            # receivers_write, sgt_write (pmcl3d.c:2245-2249)
            self.pointio.output(self)
