"""Kinematic finite-fault sources (pmcl3d's IFAULT = 5): the subfault table
the addkinsrc_H launcher reads (src/awp/kernel.cu:3777-3858).

Each subfault j sits at one cell and has two parameter rows in the six
source arrays: row 2j holds the source-time-function type (1 Brune, 2 Liu),
slip, rupture time, rise time and area in axx, ayy, azz, axz, ayz; row 2j+1
holds strike, dip and rake (degrees) in axx, ayy, azz.  An optional IIR filter
(b, a; SetDeviceFilterParameters) is applied to every source time function.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class KinematicFault:
    cells: np.ndarray      # (n, 3) int: x, y (1-based), z (1 = the free surface)
    stf_type: np.ndarray   # 1 Brune, 2 Liu
    slip: np.ndarray       # m
    ruptime: np.ndarray    # s
    risetime: np.ndarray   # s
    area: np.ndarray       # m^2
    strike: np.ndarray     # degrees
    dip: np.ndarray
    rake: np.ndarray
    b: np.ndarray | None = None
    a: np.ndarray | None = None

    @property
    def n(self) -> int:
        return len(self.cells)

    @property
    def order(self) -> int:
        return 0 if self.b is None else len(self.b) - 1

    def columns(self) -> list[np.ndarray]:
        """axx, ayy, azz, axz, ayz, axy, each 2n floats."""
        n = self.n
        cols = [np.zeros(2 * n, np.float32) for _ in range(6)]
        for c, v in enumerate((self.stf_type, self.slip, self.ruptime, self.risetime,
                               self.area)):
            cols[c][0::2] = v
        for c, v in enumerate((self.strike, self.dip, self.rake)):
            cols[c][1::2] = v
        return cols

    def check(self, dims) -> None:
        c = np.asarray(self.cells)
        if c.ndim != 2 or c.shape[1] != 3:
            raise ValueError("cells must be (n, 3)")
        if (c[:, 0].min() < 1 or c[:, 0].max() > dims.nxt or c[:, 1].min() < 1 or
                c[:, 1].max() > dims.nyt or c[:, 2].min() < 1 or c[:, 2].max() > dims.nzt):
            raise ValueError("a subfault cell lies outside the interior")
        if self.order and (self.a is None or len(self.a) != len(self.b) or self.order > 19):
            raise ValueError("the filter needs a and b of equal length, order <= 19")
