"""Padded-grid layout of pmcl3d (reference src/awp/pmcl3d.c:1154-1166,
src/awp/grid.c:14-38, include/awp/pmcl3d_cons.h:16-23).

Each 3-D field is a C-ordered array of shape (nxt+20, nyt+20, nzt+64) with k
contiguous.  Interior cells: i in [10, nxt+9], j in [10, nyt+9],
k in [32, nzt+31]; k = nzt+31 is the free surface, depth grows as k falls.
"""
from __future__ import annotations

from dataclasses import dataclass

NGSL = 8
NGSL2 = 16
ALIGN = 32
MAXGRIDS = 10

VEL_FIELDS = ("u1", "v1", "w1")
STRESS_FIELDS = ("xx", "yy", "zz", "xy", "xz", "yz")
MEMORY_FIELDS = ("r1", "r2", "r3", "r4", "r5", "r6")
MEDIUM_FIELDS = ("d1", "lam", "mu", "qp", "qs", "vx1", "vx2", "wwo")  # + int ww


@dataclass(frozen=True)
class Dims:
    nxt: int
    nyt: int
    nzt: int

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.nxt + 4 + NGSL2, self.nyt + 4 + NGSL2, self.nzt + 2 * ALIGN)

    @property
    def yline(self) -> int:
        return self.nzt + 2 * ALIGN

    @property
    def slice(self) -> int:
        return (self.nyt + 4 + NGSL2) * self.yline

    @property
    def size(self) -> int:
        a, b, c = self.shape
        return a * b * c

    @property
    def ncells(self) -> int:
        return self.nxt * self.nyt * self.nzt

    # pmcl3d's bounds on one rank (pmcl3d.c:572-640 with no neighbours)
    @property
    def i_range(self) -> tuple[int, int]:
        return (2 + NGSL, self.nxt + NGSL + 1)

    @property
    def j_range(self) -> tuple[int, int]:
        return (2 + NGSL, self.nyt + NGSL + 1)

    @property
    def ktop(self) -> int:
        return self.nzt + ALIGN - 1

    def interior(self):
        """Slices selecting the interior cells of a padded array."""
        return (slice(2 + NGSL, self.nxt + NGSL + 2), slice(2 + NGSL, self.nyt + NGSL + 2),
                slice(ALIGN, self.nzt + ALIGN))

    def cell(self, x: int, y: int, z: int) -> tuple[int, int, int]:
        """Array index of a 1-based model cell (x, y) with z counted 1 at the
        surface, the convention of the source/receiver files
        (src/awp/source.c:179, kernel.cu:1796-1798)."""
        return (x + 1 + NGSL, y + 1 + NGSL, self.nzt + ALIGN - z)
