"""Host-side setup of Drucker-Prager plasticity (NVE = 3).

Follows the reference:
  * lithostatic intermediate principal stress sigma2 -- src/awp/mesh.c:1293-1301;
  * Hoek-Brown friction angle and cohesion -- mesh.c:1139-1176 (hoek_brown) and
    the depth / fault-distance GSI taper of inidrpr_hoekbrown_light,
    mesh.c:1303-1356 (parameters as in nonlinear.dat, mesh.c:1214-1244);
  * principal-stress rotation Rz, RzT -- mesh.c:1370-1402;
  * no-yield zone around point sources (cohes = 1e18) -- src/awp/pmcl3d.c:965-989.
A constant-parameter variant (cohesion, friction coefficient) serves
workloads that prescribe Drucker-Prager parameters directly.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import ALIGN, NGSL, NGSL2, Dims


@dataclass
class PlasticParams:
    fmajor: float = 1.0
    fminor: float = 1.0
    # strike[1] is forced to 0 (mesh.c:1204); with dip (0, 90, 0) the frame is
    # orthonormal when strike[0] and strike[2] differ by 90 degrees
    strike: tuple = (0.0, 0.0, 90.0)
    dip: tuple = (0.0, 90.0, 0.0)        # mesh.c:1206-1208
    Tv: float = 0.0                      # > 0: fixed viscoplastic relaxation time

    def rotation(self) -> tuple[np.ndarray, np.ndarray]:
        """Rz, RzT as rotation_matrix builds them (mesh.c:1370-1402): angles in
        single precision, cosf/sinf products."""
        f = np.float32
        alpha = [f(s / 180.0 * np.pi) for s in self.strike]
        beta = [f(d / 180.0 * np.pi) for d in self.dip]
        Rz = np.empty(9, dtype=np.float32)
        for r in range(3):
            Rz[3 * r + 0] = np.cos(alpha[r]) * np.cos(beta[r])
            Rz[3 * r + 1] = np.sin(alpha[r]) * np.cos(beta[r])
            Rz[3 * r + 2] = np.sin(beta[r])
        RzT = Rz.reshape(3, 3).T.reshape(-1).copy()
        return Rz, RzT

    def packed(self) -> np.ndarray:
        """fmajor, fminor, Rz[9], RzT[9] as float64 (the oracle's layout)."""
        Rz, RzT = self.rotation()
        return np.concatenate([[self.fmajor, self.fminor], Rz, RzT]).astype(np.float64)


@dataclass
class PlasticMedium:
    sigma2: np.ndarray
    cohes: np.ndarray
    phi: np.ndarray
    params: PlasticParams = field(default_factory=PlasticParams)


def lithostatic_sigma2(dims: Dims, DH: float, d1: np.ndarray) -> np.ndarray:
    """sigma2 = -g * integral of rho over depth, half a cell at the surface, on
    i, j in [2, n+18), k from the surface + 1 down to align-2 (mesh.c:1293-1301)."""
    s = np.zeros(dims.shape)
    top = dims.ktop
    i0, i1 = 2, dims.nxt + 2 + NGSL2
    j0, j1 = 2, dims.nyt + 2 + NGSL2
    s[i0:i1, j0:j1, top] = -DH / 2.0 * d1[i0:i1, j0:j1, top] * 9.81
    s[i0:i1, j0:j1, top + 1] = s[i0:i1, j0:j1, top]
    for k in range(top - 1, ALIGN - 3, -1):
        s[i0:i1, j0:j1, k] = s[i0:i1, j0:j1, k + 1] - DH * d1[i0:i1, j0:j1, k] * 9.81
    return s


def hoek_brown_constants(GSI, mi, D=0.0):
    """mb, s, a of the generalised Hoek-Brown criterion
    sigma1 = sigma3 + sigma_ci (mb sigma3 / sigma_ci + s)^a (Hoek,
    Carranza-Torres and Corkum 2002; mesh.c:1148-1150)."""
    GSI = np.asarray(GSI, dtype=np.float64)
    mb = mi * np.exp((GSI - 100.0) / (28.0 - 14.0 * D))
    s = np.exp((GSI - 100.0) / (9.0 - 3.0 * D))
    a = 0.5 + 1.0 / 6.0 * (np.exp(-GSI / 15.0) - np.exp(-20.0 / 3.0))
    return mb, s, a


def hoek_brown_sigma_cm(sigma_ci, GSI, mi, D=0.0):
    """Global rock-mass strength sigma'_cm (mesh.c:1152-1153)."""
    mb, s, a = hoek_brown_constants(GSI, mi, D)
    return np.asarray(sigma_ci, dtype=np.float64) * (mb + 4.0 * s - a * (mb - 8.0 * s)) * (
        mb / 4.0 + s) ** (a - 1.0) / (2.0 * (1.0 + a) * (2.0 + a))


def hoek_brown(sigma_0, sigma_ci, GSI, mi, D=0.0, tunnel=1):
    """Generalised Hoek-Brown -> equivalent Mohr-Coulomb phi [deg], cohesion
    (Hoek et al. 2002; mesh.c:1139-1176).  Vectorised, float64."""
    sigma_0 = np.asarray(sigma_0, dtype=np.float64)
    GSI = np.asarray(GSI, dtype=np.float64)
    sigma_ci = np.asarray(sigma_ci, dtype=np.float64)
    mb, s, a = hoek_brown_constants(GSI, mi, D)
    sigma_cm = hoek_brown_sigma_cm(sigma_ci, GSI, mi, D)
    with np.errstate(divide="ignore", invalid="ignore"):
        if tunnel == 0:
            s3max = sigma_cm * 0.72 * (sigma_cm / sigma_0) ** (-0.91)
        else:
            s3max = sigma_cm * 0.47 * (sigma_cm / sigma_0) ** (-0.94)
        s3n = s3max / sigma_ci
        pw = (s + mb * s3n) ** (a - 1.0)
        tmp = 6.0 * a * mb * pw / (2.0 * (1.0 + a) * (2.0 + a) + 6 * a * mb * pw)
        phi = np.arcsin(tmp) / np.pi * 180.0
        cohes = (sigma_ci * ((1.0 + 2 * a) * s + (1.0 - a) * mb * s3n) * pw) / (
            (1.0 + a) * (2.0 + a) * np.sqrt(1.0 + (6 * a * mb * pw) / ((1.0 + a) * (2.0 + a))))
    return phi, cohes


@dataclass
class HoekBrownInput:
    """The parameters of nonlinear.dat (mesh.c:1214-1244)."""
    strike0: float
    strike2: float
    fmajor: float
    fminor: float
    sigma_ci_type: int
    sigma_ci: float
    GSI: float
    mi: float
    gsi100: float
    fltpos: int

    @classmethod
    def parse(cls, text: str) -> "HoekBrownInput":
        ln = [l.split() for l in text.strip().splitlines()]
        return cls(float(ln[0][0]), float(ln[0][1]), float(ln[1][0]), float(ln[1][1]),
                   int(ln[2][0]), float(ln[2][1]), float(ln[3][0]), float(ln[4][0]),
                   float(ln[5][0]), int(ln[6][0]))


def hoek_brown_medium(dims: Dims, DH: float, d1, lam, mu, hb: HoekBrownInput,
                      coords=(0, 0)) -> PlasticMedium:
    """inidrpr_hoekbrown_light (mesh.c:1178-1368) on the padded grid."""
    if hb.sigma_ci_type not in (0, 3, 12):
        raise ValueError("sigma_ci_type must be 0, 3 or 12")
    sigma2 = lithostatic_sigma2(dims, DH, d1)
    cohes = np.zeros(dims.shape)
    phi = np.zeros(dims.shape)
    sInd = dims.ktop
    i0, i1 = 2, dims.nxt + 2 + NGSL2
    j0, j1 = 2, dims.nyt + 2 + NGSL2
    ks = np.arange(ALIGN, sInd + 2)
    depth = (0.5 * DH + DH * (sInd - ks))[None, None, :]
    GSI_w = np.minimum(hb.GSI + (100.0 - hb.GSI) * (depth - 0.5 * DH) / hb.gsi100, 100.0)
    GSI_core = np.where(depth < 4000.0, hb.GSI,
                        np.where(depth <= 6000.0, hb.GSI + (depth - 4000.0) / 2000.0 * (GSI_w - hb.GSI),
                                 GSI_w))
    jj = np.arange(j0, j1)
    ypos = coords[1] * dims.nyt + jj - 2 - NGSL
    fltdist = np.abs(ypos.astype(np.float32) + np.float32(0.5) - np.float32(hb.fltpos)) * DH
    fd = fltdist[None, :, None]
    GSI_d = np.where(fd < 225.0, GSI_core,
                     np.where(fd <= 750.0, GSI_core + (GSI_w - GSI_core) * (fd - 225.0) / (750.0 - 225.0),
                              GSI_w))
    sl = (slice(i0, i1), slice(j0, j1), slice(ALIGN, sInd + 2))
    if hb.sigma_ci_type == 3:
        tvp = np.sqrt((1.0 / lam[sl] + 2.0 / mu[sl]) / d1[sl])
        itime = 1.0 / tvp / 3.28 * 1.0e6
        sci = 1.4138e7 * itime ** (-3.0) * 1.0e6
    elif hb.sigma_ci_type == 12:
        tvp = np.sqrt((1.0 / lam[sl] + 2.0 / mu[sl]) / d1[sl]) / 1.0e3
        sci = 0.77 * tvp ** 2.93 * 1.0e6
    else:
        sci = hb.sigma_ci
    GSI_full = np.broadcast_to(GSI_d, sigma2[sl].shape)
    ph, co = hoek_brown(-sigma2[sl], sci, GSI_full, hb.mi, 0.0, 1)
    tvs = np.sqrt(1.0 / (mu[sl] * d1[sl]))
    co = np.where(tvs > 2000.0, 1.0e12, co)
    ph = np.where(tvs < 750.0, 35.0, ph)
    co = np.where(tvs < 750.0, 0.0, co)
    cohes[sl] = co
    phi[sl] = ph
    params = PlasticParams(fmajor=hb.fmajor, fminor=hb.fminor,
                           strike=(hb.strike0, 0.0, hb.strike2))
    return PlasticMedium(sigma2=sigma2, cohes=cohes, phi=phi, params=params)


def constant_medium(dims: Dims, DH: float, d1, cohesion: float, friction: float,
                    params: PlasticParams | None = None) -> PlasticMedium:
    """Drucker-Prager with uniform cohesion [Pa] and friction coefficient
    tan(phi); phi is stored in degrees as the kernels expect
    (kernel.cu:1372)."""
    sigma2 = lithostatic_sigma2(dims, DH, d1)
    cohes = np.full(dims.shape, float(cohesion))
    phi = np.full(dims.shape, float(np.degrees(np.arctan(friction))))
    return PlasticMedium(sigma2=sigma2, cohes=cohes, phi=phi,
                         params=params if params is not None else PlasticParams())


def protect_sources(pm: PlasticMedium, dims: Dims, psrc: np.ndarray) -> None:
    """cohes = 1e18 in a 3 x 4 x 3 box around each source cell
    (pmcl3d.c:965-989: x-1..x+1, y-2..y+1, z-1..z+1)."""
    sh = dims.shape
    for x, y, z in psrc:
        idx, idy, idz = x + 1 + NGSL, y + 1 + NGSL, z + ALIGN - 1
        for xi in range(idx - 1, idx + 2):
            for yi in range(idy - 2, idy + 2):
                for zi in range(idz - 1, idz + 2):
                    if 0 <= xi < sh[0] and 0 <= yi < sh[1] and 0 <= zi < sh[2]:
                        pm.cohes[xi, yi, zi] = 1.0e18


def nve3_regions(dims: Dims) -> dict:
    """The strip decomposition pmcl3d uses for NVE = 3 on one rank
    (pmcl3d.c:572-640, 1858-1877, 2109-2209)."""
    xls, xre = dims.i_range
    yls, yre = dims.j_range
    xss1, xse1 = xls, NGSL + 3
    xss2, xse2 = NGSL + 4, dims.nxt + NGSL - 1
    xss3, xse3 = dims.nxt + NGSL, xre
    yls2 = int(yls + (yre - yls) * 0.25)
    if yls2 % 2 != 0:
        yls2 += 1
    yre2 = int(yls + (yre - yls) * 0.75)
    if yre2 % 2 == 0:
        yre2 -= 1
    yls2 = max(yls2, yls + NGSL + 2)
    yre2 = min(yre2, yre - NGSL - 2)
    return dict(xls=xls, xre=xre, yls=yls, yre=yre, xss1=xss1, xse1=xse1, xss2=xss2,
                xse2=xse2, xss3=xss3, xse3=xse3, yls2=yls2, yre2=yre2)
