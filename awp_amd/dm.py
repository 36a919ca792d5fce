"""Host-side setup of a discontinuous-mesh stack: grid 0 on top, each lower
grid 3x coarser with a third of the lateral cells (src/awp/pmcl3d.c:370-377).

  * geometry -- the coarse top level nzt+align-1 carries the material of the
    fine level align+7 (src/awp/mesh.c:1404-1419: fine levels align+1, +4, +7
    map to coarse nzt+align-3, -2, -1), so the coarse grid's top surface lies
    (nzt_fine - 9) fine cells below the fine grid's surface;
  * checkmesh -- after each grid's medium is built, the three top coarse
    levels take the fine values at ih = 3(i-10)+10, jh = 3(j-10)+11 for
    d1, mu, lam, qp, qs (pmcl3d.c:889-898) and vx1, vx2, wwo, ww
    (pmcl3d.c:1049-1058), on the interior columns only (mesh.c:1404-1478);
  * the Cerjan sponge of grid p is ND*3^(ngrids-1-p) cells wide, and only
    the lowest grid has a bottom sponge (pmcl3d.c:1027-1029).
"""
from __future__ import annotations

from typing import Callable, Sequence

import numpy as np

from .grid import ALIGN, NGSL, Dims
from .media import Medium, build_medium

CHECKMESH_FIELDS = ("d1", "mu", "lam", "qp", "qs", "vx1", "vx2", "wwo", "ww")


def grid_factors(ngrids: int) -> list[int]:
    """grdfct: 3^(ngrids-1-p), the refinement of grid p over the lowest grid."""
    return [3 ** (ngrids - 1 - p) for p in range(ngrids)]


def stack_dims(nx: int, ny: int, nz: Sequence[int]) -> list[Dims]:
    """Dims of each grid from the lowest grid's lateral size and each grid's
    own nzt (top first)."""
    g = grid_factors(len(nz))
    return [Dims(nx * g[p], ny * g[p], nz[p]) for p in range(len(nz))]


def depth_offsets(dims: Sequence[Dims], DH: Sequence[float]) -> list[float]:
    """Depth of each grid's top surface [m]."""
    z0 = [0.0]
    for p in range(len(dims) - 1):
        z0.append(z0[p] + (dims[p].nzt - 9) * DH[p])
    return z0


def checkmesh(fine: Medium, coarse: Medium) -> int:
    """Copy the fine grid's values onto the coarse grid's three top levels
    (mesh.c:1404-1478); returns how many values changed."""
    dc = coarse.dims
    i = np.arange(2 + NGSL, dc.nxt + 2 + NGSL)
    j = np.arange(2 + NGSL, dc.nyt + 2 + NGSL)
    ih = (i - 2 - NGSL) * 3 + 2 + NGSL
    jh = (j - 2 - NGSL) * 3 + 3 + NGSL
    pairs = ((ALIGN + dc.nzt - 3, ALIGN + 1), (ALIGN + dc.nzt - 2, ALIGN + 4),
             (ALIGN + dc.nzt - 1, ALIGN + 7))
    changed = 0
    for name in CHECKMESH_FIELDS:
        lo, hi = getattr(coarse, name), getattr(fine, name)
        for kc, kh in pairs:
            src = hi[np.ix_(ih, jh, [kh])]
            dst = lo[np.ix_(i, j, [kc])]
            changed += int((dst != src).sum())
            lo[np.ix_(i, j, [kc])] = src
    return changed


def _interp_axis(F: np.ndarray, coords: np.ndarray, h: float, axis: int) -> np.ndarray:
    """Linear interpolation of lattice values (node n at (n+0.5)h) along one
    axis, clamped at the ends."""
    n = F.shape[axis]
    t = np.clip(coords / h - 0.5, 0.0, n - 1.0)
    n0 = np.minimum(np.floor(t).astype(np.int64), max(n - 2, 0))
    w = (t - n0).astype(F.dtype)
    a = np.take(F, n0, axis=axis)
    b = np.take(F, np.minimum(n0 + 1, n - 1), axis=axis)
    shp = [1] * F.ndim
    shp[axis] = -1
    w = w.reshape(shp)
    return a * (1 - w) + b * w


def sample_lattice(F: np.ndarray, h: float, x: np.ndarray, y: np.ndarray,
                   z: np.ndarray) -> np.ndarray:
    """Trilinear samples of a lattice field at the tensor grid x * y * z."""
    G = _interp_axis(F, z, h, 2)
    G = _interp_axis(G, x, h, 0)
    return _interp_axis(G, y, h, 1)


def build_stack(dims: Sequence[Dims], DH0: float, DT: float,
                model: Callable[[np.ndarray, np.ndarray, np.ndarray], tuple],
                *, ND: int = 20, ARBC: float = 0.92, EX: float = 0.0, FAC: float = 0.005,
                FP: float = 2.5) -> list[Medium]:
    """Media of every grid of the stack, sampled from one continuous model.

    dims: per grid, top first; DH0: spacing of grid 0; model(x, y, depth)
    with 1-D coordinate vectors [m] returns (vp, vs, rho, Qp, Qs) on their
    tensor grid."""
    G = len(dims)
    g = grid_factors(G)
    DH = [DH0 * g[0] / g[p] for p in range(G)]
    z0 = depth_offsets(dims, DH)
    media = []
    for p, d in enumerate(dims):
        x = (np.arange(d.nxt) + 0.5) * DH[p]
        y = (np.arange(d.nyt) + 0.5) * DH[p]
        z = z0[p] + (np.arange(d.nzt) + 0.5) * DH[p]
        vp, vs, rho, Qp, Qs = model(x, y, z)
        media.append(build_medium(d, DH[p], DT, vp, vs, rho, Qp, Qs, EX=EX, FAC=FAC, FP=FP,
                                  ND=ND * g[p], ARBC=ARBC, islowest=(p == G - 1)))
    for p in range(1, G):
        checkmesh(media[p - 1], media[p])
    return media
