"""Host-side medium setup: everything the time-stepping kernels read besides
the wavefield, built on the padded pmcl3d grid.

Follows the reference's setup code:
  * relaxation times and weights  -- src/awp/mesh.c:38-53 (taumin/taumax),
    :684-712 (tausub), :714-1097 (weights_sub), src/awp/pmcl3d.c:1036-1046
    (tau1 = 0.5(1-e^{-DT/tau}), tau2 = e^{-DT/tau});
  * the 2x2x2 texture of vx1/vx2/wwo/ww -- mesh.c:1103-1137 (init_texture);
  * Q -> velocity dispersion correction and compliance storage
    -- mesh.c:228-238 (Qs = 0.05 vs, Qp = 2 Qs for 3-variable media),
    :314-396, :445-466 (mu = 1/(rho vs^2), lam = 1/(rho(vp^2-2vs^2)),
    qp/qs = Q^-1 / FAC^EX), the k flip (nzt+align-1)-k;
  * ghost shells -- mesh.c:500-590;
  * lam_mu = lambda/(lambda+2mu) on the surface -- pmcl3d.c:941-952;
  * Cerjan sponge -- src/awp/cerjan.c:5-63.
Arithmetic is carried out in float64; the FP32 solver receives the rounded
arrays and the FP64 oracle the float64 ones.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import ALIGN, NGSL, NGSL2, Dims

# Coarse-grained relaxation weights w[ix][iy][iz] and the 16 Q(f) polynomial
# coefficients, per EX bin (src/awp/mesh.c:714-1088).  Keyed by the lower
# edge of the bin; EX < 0.01 is the constant-Q table.  Listed in the order
# w000, w100, w010, w110, w001, w101, w011, w111.
_WEIGHT_TABLE = {
    0.00: ((0.8867, 1.0440, 0.0423, 0.8110, 1.7275, 0.5615, 0.8323, 0.4641),
           (-27.5089, 7.4177, -82.8803, 13.1952, 72.0312, 0.5298, -34.1779, 6.0293,
            -52.2607, 8.1754, -1.6270, 4.6858, -27.7770, 6.2852, 14.6295, 3.8839)),
    0.01: ((0.3273, 1.0434, 0.044, 0.9393, 1.7268, 0.369, 0.8478, 0.4474),
           (7.3781, 4.1655, -83.1627, 13.1326, 69.0839, 0.4981, -37.6966, 5.5263,
            -51.4056, 8.1934, 13.1865, 3.4775, -36.1049, 7.2107, 12.3809, 3.6117)),
    0.15: ((0.001, 1.0349, 0.0497, 1.0407, 1.7245, 0.2005, 0.804, 0.4452),
           (31.8902, 1.6126, -83.2611, 13.0749, 65.485, 0.5118, -42.02, 5.0875,
            -49.2656, 8.1552, 25.7345, 2.2801, -40.8942, 7.9311, 7.0206, 3.4692)),
    0.25: ((0.001, 1.0135, 0.0621, 1.1003, 1.7198, 0.0918, 0.6143, 0.4659),
           (43.775, -0.1091, -83.1088, 13.0161, 60.9008, 0.592, -43.4857, 4.5869,
            -45.3315, 8.0252, 34.3571, 1.199, -41.4422, 8.399, -2.8772, 3.5323)),
    0.35: ((0.001, 0.9782, 0.082, 1.1275, 1.7122, 0.001, 0.4639, 0.509),
           (41.6858, -0.7344, -164.2252, 14.9961, 103.0301, -0.4199, -41.1157, 3.8266,
            -73.0432, 8.5857, 38.0868, 0.3937, -43.2133, 8.6747, 5.6362, 3.3287)),
    0.45: ((0.2073, 0.912, 0.1186, 1.081, 1.6984, 0.001, 0.1872, 0.6016),
           (20.0539, -0.4354, -81.6068, 12.8573, 45.9948, 1.1528, -23.07, 2.6719,
            -27.8961, 7.1927, 31.4788, -0.0434, -25.1661, 8.245, -45.2178, 4.8476)),
    0.55: ((0.3112, 0.8339, 0.1616, 1.0117, 1.6821, 0.0001, 0.0001, 0.7123),
           (8.0848, -0.1968, -79.9715, 12.7318, 35.7155, 1.68, -13.0365, 1.8101,
            -13.2235, 6.3697, 25.4548, -0.3947, -10.4478, 7.657, -75.9179, 6.1791)),
    0.65: ((0.1219, 0.001, 0.5084, 0.2999, 1.2197, 0.001, 0.001, 1.3635),
           (1.9975, 0.418, -76.6932, 11.3479, 40.7406, 1.9511, -2.7761, 0.5987,
            0.0, 0.0, 0.0, 0.0, 41.317, 2.1874, -88.8095, 11.0003)),
    0.75: ((0.0462, 0.001, 0.4157, 0.1585, 1.3005, 0.001, 0.001, 1.4986),
           (5.1672, 0.2129, -46.506, 11.7213, -5.8873, 1.4279, -8.2448, 0.3455,
            15.0254, -0.283, 0.0, 0.0, 58.975, 0.8131, -108.6828, 12.4362)),
    0.85: ((0.001, 0.001, 0.1342, 0.1935, 1.5755, 0.001, 0.001, 1.5297),
           (-0.8151, 0.1621, -61.9333, 12.5014, 0.0358, -0.0006, 0.0, 0.0,
            22.0291, -0.4022, 0.0, 0.0, 56.0043, 0.7978, -116.9175, 13.0244)),
}
_BIN_EDGES = (0.01, 0.15, 0.25, 0.35, 0.45, 0.55, 0.65, 0.75, 0.85, 0.95)
# order of the 8 listed weights as (ix, iy, iz)
_W_ORDER = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1))
# tausub's permutation (mesh.c:690-697): rank r of the 8 log-spaced times
_TAU_RANK = {(0, 0, 0): 1, (1, 0, 0): 6, (0, 1, 0): 7, (1, 1, 0): 4,
             (0, 0, 1): 8, (1, 0, 1): 3, (0, 1, 1): 2, (1, 1, 1): 5}


@dataclass
class QModel:
    """Relaxation spectrum of the coarse-grained memory variables."""
    tau: np.ndarray       # (2,2,2) relaxation times [s]
    weights: np.ndarray   # (2,2,2) weights, already divided by FAC^EX
    coeff: np.ndarray     # (16,) Q(f) coefficients
    taumin: float
    taumax: float
    facex: float
    w0: float             # 2*pi*FP


def tau_limits(EX: float, FAC: float) -> tuple[float, float]:
    """taumin, taumax (mesh.c:38-53)."""
    pi = 4.0 * np.arctan(1.0)
    taumax = 1.0 / (2 * pi * 0.01) * 1.0 * FAC
    taumin = None
    if 0.01 <= EX < 0.65:
        taumin = 1.0 / (2 * pi * 10.0) * 0.2 * FAC
    elif 0.65 <= EX < 0.85:
        taumin = 1.0 / (2 * pi * 12.0) * 0.5 * FAC
        taumax = 1.0 / (2 * pi * 0.08) * 2.0 * FAC
    elif 0.85 <= EX < 0.95:
        taumin = 1.0 / (2 * pi * 15.0) * 0.8 * FAC
        taumax = 1.0 / (2 * pi * 0.1) * 2.5 * FAC
    elif EX < 0.01:
        taumin = 1.0 / (2 * pi * 10.0) * 0.2 * FAC
    if taumin is None:
        raise ValueError(f"EX={EX} has no relaxation table (reference covers EX < 0.95)")
    return taumin, taumax


def tausub(taumin: float, taumax: float) -> np.ndarray:
    """8 relaxation times log-spaced in [taumin, taumax], placed on the 2x2x2
    checkerboard by tausub's permutation (mesh.c:684-712)."""
    tau = np.empty((2, 2, 2))
    for idx, r in _TAU_RANK.items():
        t = 2.0 * ((r - 0.5) / 8.0) - 1.0
        tau[idx] = np.exp(0.5 * (np.log(taumax * taumin) + np.log(taumax / taumin) * t))
    return tau


def weights_sub(EX: float, FAC: float) -> tuple[np.ndarray, np.ndarray]:
    """Weights (divided by FAC^EX) and coeff for EX (mesh.c:714-1097)."""
    if EX >= 0.95:
        raise ValueError(f"EX={EX}: the reference has no weight table for EX >= 0.95")
    key = 0.00 if EX < 0.01 else max(e for e in _WEIGHT_TABLE if e <= EX and e >= 0.01)
    wl, cl = _WEIGHT_TABLE[key]
    w = np.empty((2, 2, 2))
    for idx, v in zip(_W_ORDER, wl):
        w[idx] = v
    facex = FAC ** EX
    return w / facex, np.asarray(cl, dtype=np.float64)


def q_model(EX: float = 0.0, FAC: float = 0.005, FP: float = 2.5) -> QModel:
    taumin, taumax = tau_limits(EX, FAC)
    tau = tausub(taumin, taumax)
    w, c = weights_sub(EX, FAC)
    pi = 4.0 * np.arctan(1.0)
    return QModel(tau=tau, weights=w, coeff=c, taumin=taumin, taumax=taumax,
                  facex=FAC ** EX, w0=2 * pi * FP)


def dispersion_modulus(v: np.ndarray, rho: np.ndarray, Q: np.ndarray, qm: QModel) -> np.ndarray:
    """Unrelaxed modulus rho*v'^2 such that the memory-variable medium has phase
    velocity v at the reference frequency FP (mesh.c:314-390).  Q > 200 uses the
    low-loss real expansion, 0 < Q <= 200 the complex coarse-grained modulus,
    Q <= 0 (elastic) no correction."""
    v = np.asarray(v, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    Q = np.asarray(Q, dtype=np.float64)
    m = rho * v * v
    out = m.copy()
    w0 = qm.w0
    tau = qm.tau.reshape(-1)
    hi = Q > 200.0
    if np.any(hi):
        val0 = np.zeros(int(hi.sum()))
        for t, w in zip(tau, qm.weights.reshape(-1)):
            val0 += w / ((w0 * w0 * t * t + 1.0) * Q[hi] * qm.facex)
        out[hi] = m[hi] / (1.0 - val0)
    lo = (Q > 0.0) & ~hi
    if np.any(lo):
        q = Q[lo]
        val = np.zeros(q.shape, dtype=np.complex128)
        for n, t in enumerate(tau):
            wl = qm.coeff[2 * n] / (q * q) + qm.coeff[2 * n + 1] / q
            val += 1.0 / (1.0 - wl / (1.0 + 1j * w0 * t))
        val = 1.0 / val
        out[lo] = m[lo] / (8.0 * val.real)
    return out


@dataclass
class Medium:
    """All medium arrays on the padded grid of one mesh, float64."""
    dims: Dims
    DH: float
    DT: float
    d1: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    qp: np.ndarray
    qs: np.ndarray
    vx1: np.ndarray
    vx2: np.ndarray
    wwo: np.ndarray
    ww: np.ndarray          # int32
    coeff: np.ndarray       # (16,)
    lam_mu: np.ndarray      # (nxt+20, nyt+20)
    dcrjx: np.ndarray
    dcrjy: np.ndarray
    dcrjz: np.ndarray
    qmodel: QModel
    extra: dict = field(default_factory=dict)

    def arrays(self) -> dict:
        return {k: getattr(self, k) for k in ("d1", "lam", "mu", "qp", "qs", "vx1", "vx2",
                                              "wwo", "ww", "coeff", "lam_mu", "dcrjx",
                                              "dcrjy", "dcrjz")}


def _to_padded(dims: Dims, vol: np.ndarray, out: np.ndarray) -> None:
    """Place a model-ordered volume (x, y, z with z=0 at the surface) into the
    padded array with the k flip k = nzt+align-1-z (mesh.c:445-466)."""
    ii, jj, _ = dims.interior()
    out[ii, jj, ALIGN:dims.nzt + ALIGN] = vol[:, :, ::-1]


def _fill_shell(dims: Dims, a: np.ndarray) -> None:
    """One-cell nearest-value ghost shell around the interior box: x/y planes,
    the bottom plane k=align-1, the top plane k=nzt+align, their edges and
    corners (mesh.c:500-590)."""
    nx, ny, nz = dims.nxt, dims.nyt, dims.nzt
    i0, i1 = 2 + NGSL, nx + NGSL + 1
    j0, j1 = 2 + NGSL, ny + NGSL + 1
    k0, k1 = ALIGN, nz + ALIGN - 1
    a[i0 - 1, j0:j1 + 1, k0:k1 + 1] = a[i0, j0:j1 + 1, k0:k1 + 1]
    a[i1 + 1, j0:j1 + 1, k0:k1 + 1] = a[i1, j0:j1 + 1, k0:k1 + 1]
    a[i0 - 1:i1 + 2, j0 - 1, k0:k1 + 1] = a[i0 - 1:i1 + 2, j0, k0:k1 + 1]
    a[i0 - 1:i1 + 2, j1 + 1, k0:k1 + 1] = a[i0 - 1:i1 + 2, j1, k0:k1 + 1]
    a[i0 - 1:i1 + 2, j0 - 1:j1 + 2, k0 - 1] = a[i0 - 1:i1 + 2, j0 - 1:j1 + 2, k0]
    a[i0 - 1:i1 + 2, j0 - 1:j1 + 2, k1 + 1] = a[i0 - 1:i1 + 2, j0 - 1:j1 + 2, k1]


def texture(dims: Dims, qm: QModel, DT: float) -> tuple[np.ndarray, ...]:
    """vx1, vx2, wwo, ww on the stress region (init_texture, mesh.c:1103-1137):
    the parity counters itx (per i), ity (per j, running over all i) and itz
    (per k, running over all j and i, starting at (nzt-1)%2) pick one of the 8
    relaxation mechanisms per cell."""
    shape = dims.shape
    vx1 = np.zeros(shape)
    vx2 = np.zeros(shape)
    wwo = np.zeros(shape)
    ww = np.zeros(shape, dtype=np.int32)
    tau2 = np.exp(-DT / qm.tau)
    tau1 = 0.5 * (1.0 - tau2)
    (xls, xre), (yls, yre) = dims.i_range, dims.j_range
    NI, NJ, NK = xre - xls + 1, yre - yls + 1, dims.nzt
    ni = np.arange(NI)[:, None, None]
    nj = np.arange(NJ)[None, :, None]
    nk = np.arange(NK)[None, None, :]
    itx = (ni + 1) % 2
    ity = (ni * NJ + nj + 1) % 2
    itz = ((dims.nzt - 1) % 2 + (ni * NJ + nj) * NK + nk + 1) % 2
    itx, ity, itz = np.broadcast_arrays(itx, ity, itz)
    sl = (slice(xls, xre + 1), slice(yls, yre + 1), slice(ALIGN, ALIGN + dims.nzt))
    vx1[sl] = tau1[itx, ity, itz]
    vx2[sl] = tau2[itx, ity, itz]
    wwo[sl] = 8.0 * qm.weights[itx, ity, itz]
    ww[sl] = 1 + 4 * itx + 2 * ity + itz
    return vx1, vx2, wwo, ww


def texture_table(medium: "Medium", rule: bool = False) -> np.ndarray:
    """The 8-entry form of the relaxation texture: vx1[8], vx2[8], wwo[8]
    (float32) indexed by ww - 1, for awp_dstrqc_tex.  Raises ValueError
    unless every cell with mechanism ww holds those values in the per-cell
    arrays, and, with rule=True, unless ww itself follows init_texture's
    parity rule over the interior (mesh.c:1103-1137), which the kernel then
    evaluates instead of reading ww."""
    ww = np.asarray(medium.ww)
    tab = np.zeros(24, dtype=np.float32)
    for m in range(1, 9):
        sel = ww == m
        if not sel.any():
            continue
        for c, name in enumerate(("vx1", "vx2", "wwo")):
            v = np.asarray(getattr(medium, name), dtype=np.float32)[sel]
            if not np.all(v == v[0]):
                raise ValueError(f"{name} is not a function of the mechanism ww={m}")
            tab[8 * c + m - 1] = v[0]
    if rule:
        d = medium.dims
        (xls, xre), (yls, yre) = d.i_range, d.j_range
        ni = np.arange(xre - xls + 1)[:, None, None]
        nj = np.arange(yre - yls + 1)[None, :, None]
        nk = np.arange(d.nzt)[None, None, :]
        col = ni * (yre - yls + 1) + nj
        want = 1 + 4 * ((ni + 1) & 1) + 2 * ((col + 1) & 1) + ((d.nzt + col * d.nzt + nk) & 1)
        got = ww[xls:xre + 1, yls:yre + 1, ALIGN:ALIGN + d.nzt]
        if not np.array_equal(got, want):
            raise ValueError("ww does not follow init_texture's parity rule")
    return tab


def cerjan(dims: Dims, ND: int = 20, ARBC: float = 0.92, islowest: bool = True,
           NPC: int = 0, NX: int | None = None, NY: int | None = None,
           coords=(0, 0)) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Cerjan sponge factors (src/awp/cerjan.c:5-63): exp(-(alpha*(ND-n+1))^2)
    over ND cells on the x/y sides and (lowest grid only) the bottom, with
    alpha = sqrt(-ln ARBC)/ND."""
    nxt, nyt, nzt = dims.nxt, dims.nyt, dims.nzt
    NX = nxt if NX is None else NX
    NY = nyt if NY is None else NY
    dx = np.ones(nxt + 4 + NGSL2)
    dy = np.ones(nyt + 4 + NGSL2)
    dz = np.ones(nzt + 2 * ALIGN)
    alpha = float(np.float32(np.sqrt(-np.log(ARBC)) / ND))  # _prec alpha
    if NPC < 2:
        if nxt * coords[0] + 1 <= ND:
            for i in range(ND):
                n = i + 1
                dx[i + 2 + NGSL] *= np.exp(-((alpha * (ND - n + 1)) ** 2))
        if nxt * coords[0] + nxt >= NX - ND + 1:
            for i in range(nxt - ND, nxt):
                n = i + NX - nxt + 1
                dx[i + 2 + NGSL] *= np.exp(-((alpha * (ND - (NX - n))) ** 2))
        if nyt * coords[1] + 1 <= ND:
            for j in range(ND):
                n = j + 1
                dy[j + 2 + NGSL] *= np.exp(-((alpha * (ND - n + 1)) ** 2))
        if nyt * coords[1] + nyt >= NY - ND + 1:
            for j in range(nyt - ND, nyt):
                n = j + NY - nyt + 1
                dy[j + 2 + NGSL] *= np.exp(-((alpha * (ND - (NY - n))) ** 2))
    if islowest:
        for k in range(ND):
            n = k + 1
            dz[k + ALIGN] *= np.exp(-((alpha * (ND - n + 1)) ** 2))
    return dx, dy, dz


def build_medium(dims: Dims, DH: float, DT: float, vp: np.ndarray, vs: np.ndarray,
                 rho: np.ndarray, Qp: np.ndarray | None = None, Qs: np.ndarray | None = None,
                 *, EX: float = 0.0, FAC: float = 0.005, FP: float = 2.5, ND: int = 20,
                 ARBC: float = 0.92, sponge: bool = True, SoCalQ: bool = True,
                 islowest: bool = True, NPC: int = 0) -> Medium:
    """Medium arrays from model-ordered (x, y, z-down) vp, vs, rho and Q.

    Qp/Qs None: Qs = 0.05 vs, Qp = 2 Qs (mesh.c:228-238, 3-variable media).
    Q <= 0 everywhere gives the elastic medium (qp = qs = 0)."""
    vp = np.broadcast_to(np.asarray(vp, dtype=np.float64), (dims.nxt, dims.nyt, dims.nzt))
    vs = np.broadcast_to(np.asarray(vs, dtype=np.float64), vp.shape)
    rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), vp.shape)
    if Qs is None:
        Qs = 0.05 * vs
        Qp = 2.0 * Qs
    Qp = np.broadcast_to(np.asarray(Qp, dtype=np.float64), vp.shape)
    Qs = np.broadcast_to(np.asarray(Qs, dtype=np.float64), vp.shape)
    qm = q_model(EX, FAC, FP)
    vpc = np.sqrt(dispersion_modulus(vp, rho, Qp, qm) / rho)
    vsc = np.sqrt(dispersion_modulus(vs, rho, Qs, qm) / rho)
    Qp_eff = np.where(vsc == 0.0, 100.0, Qp)
    Qs_eff = np.where(vsc == 0.0, 50.0, Qs)
    if SoCalQ:  # mesh.c:379-383
        with np.errstate(divide="ignore", invalid="ignore"):
            vsc = np.where(vpc / vsc < 1.45, vpc / 1.45, vsc)
    shape = dims.shape
    mu = np.zeros(shape)
    lam = np.zeros(shape)
    d1 = np.zeros(shape)
    qp = np.zeros(shape)
    qs = np.zeros(shape)
    _to_padded(dims, 1.0 / (rho * vsc * vsc), mu)
    _to_padded(dims, 1.0 / (rho * (vpc * vpc - 2.0 * vsc * vsc)), lam)
    _to_padded(dims, rho, d1)
    elastic_cell = Qp_eff <= 0.0
    with np.errstate(divide="ignore"):
        qpinv = np.where(elastic_cell, 0.0, 1.0 / np.where(elastic_cell, 1.0, Qp_eff))
        qsinv = np.where(elastic_cell, 0.0, 1.0 / np.where(elastic_cell, 1.0, Qs_eff))
    _to_padded(dims, qpinv / qm.facex, qp)
    _to_padded(dims, qsinv / qm.facex, qs)
    for a in (lam, mu, d1):
        _fill_shell(dims, a)
    # qp, qs: only the top plane is extended (mesh.c:583-593)
    ii, jj, _ = dims.interior()
    qp[ii, jj, dims.nzt + ALIGN] = qp[ii, jj, dims.nzt + ALIGN - 1]
    qs[ii, jj, dims.nzt + ALIGN] = qs[ii, jj, dims.nzt + ALIGN - 1]
    vx1, vx2, wwo, ww = texture(dims, qm, DT)
    # lam_mu (pmcl3d.c:941-952), float temporaries as in the reference
    lam_mu = np.zeros((dims.nxt + 4 + NGSL2, dims.nyt + 4 + NGSL2))
    (xls, xre), (yls, yre) = dims.i_range, dims.j_range
    top = dims.ktop
    t_xl = 1.0 / lam[xls:xre + 1, yls:yre + 1, top]
    t_xl2m = 2.0 / mu[xls:xre + 1, yls:yre + 1, top] + t_xl
    lam_mu[xls:xre + 1, yls:yre + 1] = t_xl / t_xl2m
    if sponge:
        dx, dy, dz = cerjan(dims, ND=ND, ARBC=ARBC, islowest=islowest, NPC=NPC)
    else:
        dx = np.ones(dims.nxt + 4 + NGSL2)
        dy = np.ones(dims.nyt + 4 + NGSL2)
        dz = np.ones(dims.nzt + 2 * ALIGN)
    return Medium(dims=dims, DH=DH, DT=DT, d1=d1, lam=lam, mu=mu, qp=qp, qs=qs, vx1=vx1,
                  vx2=vx2, wwo=wwo, ww=ww, coeff=qm.coeff.copy(), lam_mu=lam_mu, dcrjx=dx,
                  dcrjy=dy, dcrjz=dz, qmodel=qm)
