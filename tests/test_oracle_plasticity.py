"""Pins the oracle's Drucker-Prager plasticity (reference src/awp/kernel.cu:
43-64, 1253-1422, 1650-1778) to closed forms.

* rotate_principal gives the tensor R^T diag((s2+pf)(fmajor, 1, fminor)) R - pf I:
  its eigenvalues and eigenvectors are fixed by construction.
* On a uniform stress state above yield with Tv -> 0 (exp(-DT/Tv) = 0), one
  drprecpc_calc + drprecpc_app pair returns the total stress exactly onto the
  yield surface sqrt(J2) = c cos(phi) - (sigma_m + pf) sin(phi), keeps the
  mean stress, and adds (1-f)/mu * sqrt(J2 + sum of shear^2) to neta.
* Below yield nothing changes and yldfac = 1.
"""
import numpy as np
import pytest

from awp_amd.grid import ALIGN, Dims
from awp_amd.plasticity import PlasticParams, nve3_regions


# pi/180 as the reference writes it: a single-precision literal (kernel.cu:1372)
DEG = float(np.float32(0.017453292519943295))


def _p(a):
    return a.ctypes.data


def test_rotate_principal_eigenstructure(oracle_lib):
    o = oracle_lib.load()
    pp = PlasticParams(fmajor=1.4, fminor=0.6, strike=(35.0, 0.0, 125.0))
    pl = np.ascontiguousarray(pp.packed())
    out = np.zeros(9)
    s2, pf = -2.0e7, 9.0e6
    o.awp_oracle_rotate_principal_f64(s2, pf, _p(pl), _p(out))
    T = out.reshape(3, 3) + pf * np.eye(3)
    Rz = pl[2:11].reshape(3, 3)
    # rows of Rz are the principal directions
    want = np.array([(s2 + pf) * pp.fmajor, s2 + pf, (s2 + pf) * pp.fminor])
    for r in range(3):
        np.testing.assert_allclose(T @ Rz[r], want[r] * Rz[r], rtol=1e-6, atol=1e-6 * abs(s2))
    np.testing.assert_allclose(T, T.T, atol=1e-9 * abs(s2))


def _uniform_state(dims, vals, sigma2, mu=1 / 3.0e10, rho=2600.0, cohes=4.0e6, phi=30.0):
    sh = dims.shape
    st = {n: np.full(sh, v) for n, v in vals.items()}
    med = {"mu": np.full(sh, mu), "d1": np.full(sh, rho), "sigma2": np.full(sh, sigma2),
           "cohes": np.full(sh, cohes), "phi": np.full(sh, phi),
           "yldfac": np.ones(sh), "neta": np.zeros(sh)}
    return st, med


def _calc_app(o, st, med, dims, DH, DT, pl, Tv, region):
    xs, xe, ys, ye = region
    o.awp_oracle_drprecpc_calc_f64(
        *(_p(st[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")), _p(med["mu"]), _p(med["d1"]),
        _p(med["sigma2"]), _p(med["yldfac"]), _p(med["cohes"]), _p(med["phi"]), _p(med["neta"]),
        dims.nxt, dims.nyt, dims.nzt, DH, DT, xs, xe, ys, ye, _p(pl), Tv)
    o.awp_oracle_drprecpc_app_f64(
        *(_p(st[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")), _p(med["sigma2"]),
        _p(med["yldfac"]), dims.nxt, dims.nyt, dims.nzt, DH, xs, xe, ys, ye, _p(pl))


def _initial(o, pl, s2, pf):
    out = np.zeros(9)
    o.awp_oracle_rotate_principal_f64(s2, pf, _p(pl), _p(out))
    return out.reshape(3, 3)


def _yield_factor(o, pl, dyn, s2, pf, c, phi):
    """Closed-form Drucker-Prager return factor for Tv -> 0 at one level:
    taulim / sqrt(J2) of the total stress (dynamic + initial)."""
    I = _initial(o, pl, s2, pf)
    S = np.array([[dyn["xx"], dyn["xy"], dyn["xz"]], [dyn["xy"], dyn["yy"], dyn["yz"]],
                  [dyn["xz"], dyn["yz"], dyn["zz"]]]) + I
    sm = np.trace(S) / 3.0
    D = S - sm * np.eye(3)
    q = np.sqrt(0.5 * np.sum(D * D))
    taulim = c * np.cos(phi * DEG) - (sm + pf) * np.sin(phi * DEG)
    return taulim / q, S, I, q, taulim


@pytest.mark.parametrize("fmaj,fmin,strike", [(1.0, 1.0, (0, 0, 90.0)), (1.3, 0.7, (40.0, 0, 130.0))])
def test_return_mapping_onto_yield_surface(oracle_lib, fmaj, fmin, strike):
    o = oracle_lib.load()
    dims, DH, DT = Dims(8, 8, 8), 100.0, 0.005
    pp = PlasticParams(fmajor=fmaj, fminor=fmin, strike=strike)
    pl = np.ascontiguousarray(pp.packed())
    s2 = -1.5e7
    dyn = dict(xx=6.0e6, yy=-4.5e6, zz=1.5e6, xy=7.5e6, xz=-3.0e6, yz=5.0e6)
    c, phi, mu_c = 4.0e6, 30.0, 1 / 3.0e10
    st, med = _uniform_state(dims, dyn, s2, mu=mu_c, cohes=c, phi=phi)
    _calc_app(o, st, med, dims, DH, DT, pl, 1e-30, (10, 17, 10, 17))
    i, j, k = 13, 13, ALIGN + 3
    pf0 = ((dims.ktop - k) * DH + DH * 0.5) * 9.81e3
    pf1 = ((dims.ktop - k - 1) * DH + DH * 0.5) * 9.81e3
    f0, S0, I0, q0, tl0 = _yield_factor(o, pl, dyn, s2, pf0, c, phi)
    f1, *_ = _yield_factor(o, pl, dyn, s2, pf1, c, phi)
    f1 = min(f1, 1.0)                                     # level k+1 may not yield
    assert q0 > tl0 > 0                                   # above yield
    # the yield factor puts the cell's total stress on the yield surface
    assert med["yldfac"][i, j, k] == pytest.approx(f0, rel=1e-12)
    # xy: total shear scaled by the level-k factor about the initial stress
    assert st["xy"][i, j, k] + I0[0, 1] == pytest.approx(f0 * S0[0, 1], rel=1e-12)
    # xz, yz sit half a cell up: the mean of the k and k+1 factors.  The
    # initial stress drprecpc_app removes there is rotated with the fluid
    # pressure of depth + DH (kernel.cu:1675-1682 compute depth_kp1 as
    # (srfpos-k+1)*DH), a reference quirk the oracle keeps.
    pfq = ((dims.ktop - k + 1) * DH + DH * 0.5) * 9.81e3
    Ik1 = _initial(o, pl, s2, pfq)
    for n, (a, b) in (("xz", (0, 2)), ("yz", (1, 2))):
        iist = 0.5 * (I0[a, b] + Ik1[a, b])
        assert st[n][i, j, k] + iist == pytest.approx(0.5 * (f0 + f1) * (dyn[n] + iist), rel=1e-12)
    # plastic strain increment |d eps_p| = (1-f) |S_dev| / mu
    D0 = S0 - np.trace(S0) / 3.0 * np.eye(3)
    want_eta = (1 - f0) * mu_c * np.sqrt(0.5 * np.sum(np.diag(D0) ** 2) +
                                         2 * (D0[0, 1] ** 2 + D0[0, 2] ** 2 + D0[1, 2] ** 2))
    assert med["neta"][i, j, k] == pytest.approx(want_eta, rel=1e-10)
    if fmaj == fmin == 1.0:
        # isotropic initial stress: the normal deviator is scaled by f0 and the
        # mean stress is kept, so sqrt(J2) lands exactly on taulim
        S1 = np.array([[st["xx"][i, j, k], st["xy"][i, j, k]],
                       [st["xy"][i, j, k], st["yy"][i, j, k]]])
        sm1 = (st["xx"][i, j, k] + st["yy"][i, j, k] + st["zz"][i, j, k]) / 3.0 + s2
        assert sm1 == pytest.approx(np.trace(S0) / 3.0, rel=1e-12)
        D1 = np.diag([st[n][i, j, k] + s2 - sm1 for n in ("xx", "yy", "zz")])
        np.testing.assert_allclose(np.diag(D1), f0 * np.diag(D0), rtol=1e-10)
        del S1


def test_viscoplastic_relaxation_factor(oracle_lib):
    """yldfac = r + (1 - r) exp(-DT/Tv) with r = taulim/sqrt(J2); Tv = DH/vs by
    default."""
    o = oracle_lib.load()
    dims, DH, DT = Dims(8, 8, 8), 50.0, 0.004
    pl = np.ascontiguousarray(PlasticParams().packed())
    s2 = -8.0e6
    dyn = dict(xx=0.0, yy=0.0, zz=0.0, xy=6.0e6, xz=0.0, yz=0.0)
    mu_c, rho = 1 / 2.0e10, 2500.0
    st, med = _uniform_state(dims, dyn, s2, mu=mu_c, rho=rho, cohes=1.0e6, phi=25.0)
    _calc_app(o, st, med, dims, DH, DT, pl, 0.0, (10, 17, 10, 17))
    k, i, j = ALIGN + 4, 13, 13
    pf = ((dims.ktop - k) * DH + DH * 0.5) * 9.81e3
    r, *_ = _yield_factor(o, pl, dyn, s2, pf, 1.0e6, 25.0)
    assert r < 1
    vs = np.sqrt(2.0e10 / rho)
    want = r + (1 - r) * np.exp(-DT / (DH / vs))
    assert med["yldfac"][i, j, k] == pytest.approx(want, rel=1e-12)


def test_below_yield_is_identity(oracle_lib):
    o = oracle_lib.load()
    dims, DH, DT = Dims(8, 8, 8), 100.0, 0.005
    pl = np.ascontiguousarray(PlasticParams(fmajor=1.1, fminor=0.9, strike=(20.0, 0, 110.0)).packed())
    dyn = dict(xx=1.0e5, yy=-2.0e5, zz=3.0e4, xy=1.0e5, xz=-5.0e4, yz=2.0e4)
    st, med = _uniform_state(dims, dyn, -2.0e7, cohes=5.0e7)
    before = {n: v.copy() for n, v in st.items()}
    _calc_app(o, st, med, dims, DH, DT, pl, 0.0, (10, 17, 10, 17))
    for n in before:
        assert np.array_equal(st[n], before[n]), n
    assert np.all(med["yldfac"] == 1.0) and not med["neta"].any()


def test_nve3_regions_cover_the_stress_region_once():
    dims = Dims(64, 48, 16)
    R = nve3_regions(dims)
    cover = np.zeros((dims.nxt + 20, dims.nyt + 20), int)
    for xs, xe, ys, ye in ((R["xss2"], R["xse2"], R["yls"], R["yls2"] - 1),
                           (R["xss2"], R["xse2"], R["yre2"] + 1, R["yre"]),
                           (R["xss1"], R["xse1"], R["yls"], R["yre"]),
                           (R["xss3"], R["xse3"], R["yls"], R["yre"]),
                           (R["xss2"], R["xse2"], R["yls2"], R["yre2"])):
        cover[xs:xe + 1, ys:ye + 1] += 1
    (xl, xr), (yl, yr) = dims.i_range, dims.j_range
    assert np.all(cover[xl:xr + 1, yl:yr + 1] == 1)
    assert cover.sum() == dims.nxt * dims.nyt
    assert R["yls2"] % 2 == 0 and R["yre2"] % 2 == 1


# ---------------------------------------------------------------------------
# Plasticity setup (mesh.c:1139-1402): lithostatic sigma2, Hoek-Brown
# equivalent Mohr-Coulomb parameters, principal-stress rotation.

def test_lithostatic_sigma2_closed_forms():
    """sigma2 at the cell centres is -g times the overburden: for constant
    density -rho g depth, depth = (ktop - k + 1/2) DH (the depth at which
    the kernels take the fluid pressure, kernel.cu:1294); for two layers the
    density of each cell holds from its centre up to the next centre above,
    so the interface sits at the centre of the deepest upper-layer cell.
    The ghost level above the surface repeats the surface value; columns
    i, j in [2, n + 18) and levels down to align - 2 are filled."""
    from awp_amd.grid import NGSL2
    from awp_amd.plasticity import lithostatic_sigma2
    dims = Dims(6, 5, 30)
    DH, g = 75.0, 9.81
    top = dims.ktop
    k = np.arange(ALIGN - 2, top + 1)
    depth = (top - k + 0.5) * DH
    i0, i1, j0, j1 = 2, dims.nxt + 2 + NGSL2, 2, dims.nyt + 2 + NGSL2
    rho = np.full(dims.shape, 2700.0)
    s = lithostatic_sigma2(dims, DH, rho)
    np.testing.assert_allclose(s[i0:i1, j0:j1, ALIGN - 2:top + 1], -2700.0 * g * depth
                               * np.ones((i1 - i0, j1 - j0, 1)), rtol=1e-13)
    assert np.array_equal(s[:, :, top + 1], s[:, :, top])
    assert not s[:i0].any() and not s[i1:].any() and not s[:, :j0].any()
    assert not s[:, j1:].any() and not s[:, :, :ALIGN - 2].any()
    # two layers: cells k > kI light, k <= kI dense
    kI = top - 11
    rho = np.where(np.arange(dims.shape[2]) > kI, 2200.0, 2900.0)[None, None, :] * np.ones(
        dims.shape)
    s = lithostatic_sigma2(dims, DH, rho)
    D = (top - (kI + 1) + 0.5) * DH          # centre of the deepest light cell
    want = -g * (2200.0 * np.minimum(depth, D) + 2900.0 * np.maximum(depth - D, 0.0))
    np.testing.assert_allclose(s[4, 7, ALIGN - 2:top + 1], want, rtol=1e-13)


def _mc_fit(sigma_ci, mb, s, a, lo, hi, n=400001):
    """Least-squares line sigma1 = A + B sigma3 through the Hoek-Brown
    envelope on [lo, hi]: the equivalent Mohr-Coulomb criterion of Hoek,
    Carranza-Torres and Corkum (2002); returns phi [deg], cohesion, A."""
    x = np.linspace(lo, hi, n)
    y = x + sigma_ci * (mb * x / sigma_ci + s) ** a
    B, A = np.polyfit(x, y, 1)
    sp = (B - 1.0) / (B + 1.0)
    return np.degrees(np.arcsin(sp)), A * (1.0 - sp) / (2.0 * np.sqrt(1.0 - sp * sp)), A


@pytest.mark.parametrize("GSI,mi,sigma_ci,sigma_0", [(50.0, 10.0, 50e6, 20e6),
                                                      (80.0, 25.0, 100e6, 5e6),
                                                      (30.0, 7.0, 30e6, 50e6),
                                                      (100.0, 12.0, 80e6, 10e6)])
def test_hoek_brown_is_the_least_squares_mohr_coulomb_fit(GSI, mi, sigma_ci, sigma_0):
    """The closed forms of phi' and c' (mesh.c:1139-1176) are the
    least-squares Mohr-Coulomb line through the Hoek-Brown envelope
    sigma1 = sigma3 + sigma_ci (mb sigma3/sigma_ci + s)^a over
    sigma_t = -s sigma_ci/mb < sigma3 < sigma3max, with the published tunnel
    relation sigma3max = 0.47 sigma_cm (sigma_cm/sigma_0)^-0.94; and the
    rock-mass strength sigma'_cm is the intercept of the same fit over
    sigma_t < sigma3 < sigma_ci/4.  GSI = 100 gives mb = mi, s = 1, a = 1/2
    (intact rock).  The fits are evaluated numerically, independently of the
    closed forms."""
    from awp_amd.plasticity import hoek_brown, hoek_brown_constants, hoek_brown_sigma_cm
    mb, s, a = hoek_brown_constants(GSI, mi)
    if GSI == 100.0:
        assert (mb, s, a) == (mi, 1.0, 0.5)
    st = -s * sigma_ci / mb
    scm = float(hoek_brown_sigma_cm(sigma_ci, GSI, mi))
    _, _, A = _mc_fit(sigma_ci, mb, s, a, st, sigma_ci / 4.0)
    assert abs(scm / A - 1.0) < 1e-5, (scm, A)
    s3max = scm * 0.47 * (scm / sigma_0) ** (-0.94)
    phi_fit, c_fit, _ = _mc_fit(sigma_ci, mb, s, a, st, s3max)
    phi, c = hoek_brown(sigma_0, sigma_ci, GSI, mi, 0.0, 1)
    assert abs(float(phi) / phi_fit - 1.0) < 1e-5, (phi, phi_fit)
    assert abs(float(c) / c_fit - 1.0) < 1e-4, (c, c_fit)
    # the slope relation (tunnel = 0) takes 0.72 and -0.91
    s3max = scm * 0.72 * (scm / sigma_0) ** (-0.91)
    phi_fit, c_fit, _ = _mc_fit(sigma_ci, mb, s, a, st, s3max)
    phi, c = hoek_brown(sigma_0, sigma_ci, GSI, mi, 0.0, 0)
    assert abs(float(phi) / phi_fit - 1.0) < 1e-5 and abs(float(c) / c_fit - 1.0) < 1e-4


@pytest.mark.parametrize("s0", [0.0, 35.0, -120.0, 200.0])
def test_rotation_frame(s0):
    """rotation_matrix (mesh.c:1370-1402) with the reference's dips
    (0, 90, 0) and strike (s0, 0, s0 + 90): row 0 is the horizontal unit
    vector at azimuth s0 from x (the major principal direction), row 1 the
    vertical (the intermediate principal stress is vertical, mesh.c:1204),
    row 2 the horizontal normal; the frame is orthonormal to single
    precision and RzT is its transpose."""
    pp = PlasticParams(strike=(s0, 0.0, s0 + 90.0))
    Rz, RzT = pp.rotation()
    R = Rz.reshape(3, 3).astype(np.float64)
    a = np.radians(s0)
    np.testing.assert_allclose(R[0], [np.cos(a), np.sin(a), 0.0], atol=2e-7)
    np.testing.assert_allclose(R[1], [0.0, 0.0, 1.0], atol=2e-7)
    np.testing.assert_allclose(R[2], [-np.sin(a), np.cos(a), 0.0], atol=2e-7)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=5e-7)
    assert np.array_equal(RzT.reshape(3, 3), Rz.reshape(3, 3).T)
    # a strike pair that is not 90 degrees apart is not a rotation
    Rb = PlasticParams(strike=(s0, 0.0, s0 + 60.0)).rotation()[0].reshape(3, 3)
    assert np.abs(Rb @ Rb.T - np.eye(3)).max() > 0.4
