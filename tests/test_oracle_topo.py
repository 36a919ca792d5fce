"""Pinning of the curvilinear-topography oracle (CPU).

- Coefficient tables: every boundary row is exact on low-degree polynomials
  at the node / cell-centre positions of the grid in z (the metric and stress
  tables to degree 2, the interpolations to degree 1); the velocity tables,
  which fold in the traction-free penalty, are exact on polynomials that vanish
  at the surface; the stencils are the same numbers in the oracle's and the
  library's copies.
- Metric builders: constant, linear and quadratic f and g give the exact
  interpolants and derivatives at the staggered positions (the reference's
  tests/topography/metrics/test_metrics.c cases).
- Kernels: on flat topography (f = 1, g linear) both updates equal the
  Cartesian oracle kernels in the interior (the reference's
  tests/test_attenuation.cu invariant); on a tilted surface, one update of
  fields linear in the physical coordinates (x, y, z = g(r) f(x, y)) gives the
  exact strain-rate / stress-divergence closed form, so the metric terms
  (J^-1, f1_*, f2_*, g_c) are right.
"""
import numpy as np
import pytest

from awp_amd.grid import ALIGN, NGSL, Dims
from awp_amd.topography import Topography

from _topo import ROOT, TopoOracle, parse_tables, physical_xyz

NZ = 40


def _grids(nz):
    zn = np.arange(nz, dtype=float)
    zn[nz - 1] = np.nan                     # ghost node
    zc = np.arange(nz) - 0.5
    zc[0], zc[nz - 1] = 0.0, nz - 2         # boundary points of the cell grid
    return zn, zc


def _exact_degrees(table, src, dst, deriv, top, nz, zero_at=None, maxdeg=4):
    ok_deg = []
    for deg in range(maxdeg + 1):
        ok = True
        for k in range(table.shape[0]):
            ro = (nz - 1 - k) if top else k
            zo = dst[ro]
            if np.isnan(zo):
                assert np.all(table[k] == 0.0)  # a ghost row is all zero
                continue
            acc = 0.0
            for m in range(table.shape[1]):
                if table[k, m] == 0.0:
                    continue
                z = src[(nz - 1 - m) if top else m]
                assert not np.isnan(z), "a table column weights the ghost node"
                P = z ** deg if zero_at is None else (z - zero_at) * z ** deg
                acc += table[k, m] * P
            if zero_at is None:
                want = (deg * zo ** (deg - 1) if deg else 0.0) if deriv else zo ** deg
            else:
                want = ((zo ** deg + deg * (zo - zero_at) * zo ** (deg - 1)) if deg else 1.0) \
                    if deriv else (zo - zero_at) * zo ** deg
            if abs(acc - want) > 1e-10 * max(1.0, abs(want)):
                ok = False
        if ok:
            ok_deg.append(deg)
    return ok_deg


@pytest.fixture(scope="module")
def tables():
    return parse_tables(ROOT / "oracle" / "topo_tables.h")


def test_library_tables_are_the_oracle_tables(tables):
    lib = parse_tables(ROOT / "awp_amd" / "csrc" / "awp_topo_coef.h")
    for name, t in tables.items():
        assert np.array_equal(lib[name], t), name
        # signed zeros too (the stress kernels' centre weight is -0)
        assert np.array_equal(np.signbit(lib[name]), np.signbit(t)), name


def test_interior_stencils(tables):
    p4 = np.array([-1, 9, 9, -1]) / 16.0
    d4 = np.array([1 / 24, -9 / 8, 9 / 8, -1 / 24])
    comp = np.convolve(d4, p4)  # derivative of the 4-point interpolant
    for name in ("TV_DHPZ4", "TV_DPHZ4", "TS_PHDZ4", "TS_PDHZ4"):
        assert np.allclose(tables[name], comp, rtol=0, atol=1e-15), name
    # degree-3 exactness of the staggered derivative and interpolation
    x = np.arange(4) - 1.5
    for deg in range(4):
        assert abs(d4 @ x ** deg - (deg * 0.0 ** (deg - 1) if deg else 0.0)) < 1e-14
        assert abs(p4 @ x ** deg - (0.0 if deg else 1.0)) < 1e-14


def test_boundary_tables_polynomial_exactness(tables):
    zn, zc = _grids(NZ)
    T = tables
    cases = [  # name, source grid, destination grid, derivative, top rows, need degree
        ("TM_PHZL", zn, zc, False, False, 1), ("TM_PHZR", zn, zc, False, True, 1),
        ("TM_DZL", zc, zn, True, False, 2), ("TM_DZR", zc, zn, True, True, 2),
        ("TM_DHZL", zn, zc, True, False, 2), ("TM_DHZR", zn, zc, True, True, 2),
        ("TS_DZ4R", zc, zn, True, True, 2), ("TS_DHZ4R", zn, zc, True, True, 2),
        ("TS_PHDZ4R", zc, zc, True, True, 2), ("TS_PDHZ4R", zn, zn, True, True, 2),
    ]
    for name, src, dst, der, top, need in cases:
        degs = _exact_degrees(T[name], src, dst, der, top, NZ)
        assert list(range(need + 1)) == degs[:need + 1], (name, degs)
    zb = NZ - 2  # the free surface: last node = top cell-centre point
    for name, src, dst, need in (("TV_DHZ4R", zn, zc, 1), ("TV_DZ4R", zc, zn, 1),
                                 ("TV_DHPZ4R", zc, zc, 0), ("TV_DPHZ4R", zn, zn, 0)):
        degs = _exact_degrees(T[name], src, dst, True, True, NZ, zero_at=zb)
        assert list(range(need + 1)) == degs[:need + 1], (name, degs)
    # the velocity rows differ from the plain derivative rows only by the
    # penalty on the surface value
    dv, ds = T["TV_DHZ4R"], T["TS_DHZ4R"]
    assert np.array_equal(dv[1:], ds[1:])
    assert np.array_equal(np.delete(dv[0], 1), np.delete(ds[0], 1))
    assert dv[0, 1] == -ds[0, 1]


def _right_block(blk: dict, nrows: int, ncols: int) -> np.ndarray:
    """Dense top-rows operator from an SBP right-boundary block: the listed
    rows, then the last (interior) row shifted by one column per row."""
    op, idx = np.array(blk["op_right"]), np.array(blk["idx_right"])
    M = np.zeros((nrows, ncols))
    for r in range(op.shape[0]):
        M[r, idx[r]] += op[r]
    last = np.nonzero(op[-1])[0]
    for r in range(op.shape[0], nrows):
        c = last + (r - op.shape[0] + 1)
        if c.max() < ncols:
            M[r, c] = op[-1, last]
    return M


def test_velocity_composite_top_rows_are_products_of_pinned_factors(tables):
    """dhpz4r and dphz4r (optimized_velocity.cu:1056, 1110) are z-derivatives
    of z-interpolants, D(P(.), interp="first") in the reference's generator
    (tools/kernel_generation/scheme.py:150-163).  Each row is the product of
    the one-sided derivative (with the traction penalty) and the
    interpolation between the node and cell-centre grids.  The factors are
    pinned on their own: TV_DHZ4R / TV_DZ4R by exactness on polynomials
    vanishing at the surface, TM_PHZR and the node interpolation Pp (a
    fixture from the generator's coefficient files) by exactness to
    degree 1, and the product reproduces every coefficient of both
    composite tables to rounding."""
    import json
    T = tables
    blocks = json.loads((ROOT / "tests" / "golden" / "sbp_right_blocks.json").read_text())
    # the kernel tables are the generator's operators
    assert np.allclose(T["TV_DHZ4R"], _right_block(blocks["Dm_bc42"], 6, 8), rtol=0, atol=1e-15)
    assert np.allclose(T["TV_DZ4R"], _right_block(blocks["Dp_bc42"], 6, 7), rtol=0, atol=1e-15)
    assert np.allclose(T["TM_PHZR"], _right_block(blocks["Pm42"], 6, 8), rtol=0, atol=1e-15)
    # Pp: cell-centre values -> node values, exact on degree 0 and 1
    zn, zc = _grids(NZ)
    Pp = _right_block(blocks["Pp42"], 12, 12)
    degs = _exact_degrees(Pp[:6, :7], zc, zn, False, True, NZ)
    assert degs[:2] == [0, 1], degs
    # composites = products of the pinned factors
    DhP = _right_block({"op_right": T["TV_DHZ4R"].tolist(),
                        "idx_right": [list(range(8))] * 6}, 12, 12) @ Pp
    assert np.abs(DhP[:6, :9] - T["TV_DHPZ4R"]).max() < 1e-15
    assert not np.any(DhP[:6, 9:])
    Pm = _right_block({"op_right": T["TM_PHZR"].tolist(),
                       "idx_right": [list(range(8))] * 6}, 12, 12)
    DP = _right_block({"op_right": T["TV_DZ4R"].tolist(),
                       "idx_right": [list(range(7))] * 6}, 12, 12) @ Pm
    assert np.abs(DP[:6, :9] - T["TV_DPHZ4R"]).max() < 1e-15
    assert not np.any(DP[:6, 9:])
    # and the generator's own composite files hold the same numbers
    assert np.abs(np.array(blocks["DmPp_bc42"]["op_right"]) - T["TV_DHPZ4R"]).max() < 1e-15
    assert np.abs(np.array(blocks["DpPm_bc42"]["op_right"]) - T["TV_DPHZ4R"]).max() < 1e-15


def test_composite_product_check_catches_a_planted_coefficient(tables):
    """A change of three composite coefficients that keeps the row exact on
    (z - z_s) (which the exactness test cannot see) breaks the product
    identity."""
    import json
    T = tables
    blocks = json.loads((ROOT / "tests" / "golden" / "sbp_right_blocks.json").read_text())
    bad = T["TV_DHPZ4R"].copy()
    bad[2, 1] += 1e-3   # second difference over equally spaced cell centres:
    bad[2, 2] -= 2e-3   # orthogonal to constants and to linear functions
    bad[2, 3] += 1e-3
    Pp = _right_block(blocks["Pp42"], 12, 12)
    DhP = _right_block({"op_right": T["TV_DHZ4R"].tolist(),
                        "idx_right": [list(range(8))] * 6}, 12, 12) @ Pp
    zn, zc = _grids(NZ)
    zb = NZ - 2
    assert _exact_degrees(bad, zc, zc, True, True, NZ, zero_at=zb)[:1] == [0]
    assert np.abs(DhP[:6, :9] - bad).max() > 1e-4


# ---------------------------------------------------------------------------
def _dims():
    return Dims(24, 20, NZ)


def test_metrics_f_closed_forms(oracle_lib):
    """Interpolants of constant and linear f, derivatives of linear and
    quadratic f at the staggered positions (f_1: (i, j-1/2), f_2: (i+1/2, j),
    f_c: (i+1/2, j-1/2)); hi = 1/h scales the derivatives."""
    d = _dims()
    h = 50.0
    I = np.arange(d.nxt + 20) - (NGSL + 2)
    J = np.arange(d.nyt + 84) - (ALIGN + NGSL + 2)
    X, Y = np.meshgrid(I * h, J * h, indexing="ij")
    ii, jj = np.meshgrid(np.arange(-6, d.nxt + 6), np.arange(-6, d.nyt + 6), indexing="ij")
    for fx, grad in (
            (lambda x, y: 3.0 + 0 * x, lambda x, y: (0 * x, 0 * x)),
            (lambda x, y: 1.0 + 2e-4 * x - 3e-4 * y, lambda x, y: (2e-4 + 0 * x, -3e-4 + 0 * x)),
            (lambda x, y: 1.0 + 1e-7 * x * x + 2e-7 * y * y - 1e-7 * x * y,
             lambda x, y: (2e-7 * x - 1e-7 * y, 4e-7 * y - 1e-7 * x))):
        T = TopoOracle(oracle_lib, d, h, f=fx(X, Y))
        for name, (ox, oy) in (("f_1", (0.0, -0.5)), ("f_2", (0.5, 0.0)), ("f_c", (0.5, -0.5))):
            got = T.F(name, ii, jj)
            xs, ys = (ii + ox) * h, (jj + oy) * h
            want = fx(xs, ys)
            # the 4-point interpolation is exact to degree 3 in each direction
            assert np.allclose(got, want, rtol=1e-13, atol=1e-13), name
        for name, comp, (ox, oy) in (("f1_1", 0, (0.0, -0.5)), ("f1_2", 0, (0.5, 0.0)),
                                     ("f1_c", 0, (0.5, -0.5)), ("f2_1", 1, (0.0, -0.5)),
                                     ("f2_2", 1, (0.5, 0.0)), ("f2_c", 1, (0.5, -0.5))):
            # derivatives of interpolants read two cells further out: check
            # where all inputs were built
            sel = (ii >= -4) & (ii < d.nxt + 4) & (jj >= -4) & (jj < d.nyt + 4)
            got = T.F(name, ii, jj)[sel]
            want = grad((ii + ox) * h, (jj + oy) * h)[comp][sel]
            assert np.allclose(got, want, rtol=1e-11, atol=1e-16), name


def test_metrics_g_closed_forms(oracle_lib):
    """g linear (the geometry the reference builds) and quadratic: g_c at the
    cell centres, g3 and g3_c the exact derivatives (hi = 1/h scaling), g3 = 1
    on the top ghost."""
    d = _dims()
    h = 50.0
    nz = d.nzt
    zn, zc = _grids(nz)
    k = np.arange(nz)
    for gfun, dg in ((lambda r: 7.0 + 3.0 * r, lambda r: 3.0 + 0 * r),
                     (lambda r: 2.0 + r + 0.05 * r * r, lambda r: 1.0 + 0.1 * r)):
        g = np.zeros(nz + 2 * ALIGN)
        g[ALIGN:ALIGN + nz - 1] = gfun(zn[:nz - 1])
        T = TopoOracle(oracle_lib, d, h, g=g)
        gc = T.Gz("g_c", k)
        assert np.allclose(gc, gfun(zc), rtol=1e-12, atol=1e-12) or dg(0.0) != 3.0
        g3 = T.Gz("g3", k)
        assert np.allclose(g3[:nz - 1] * h, dg(zn[:nz - 1]), rtol=1e-10, atol=1e-10) \
            or dg(0.0) != 3.0
        assert g3[nz - 1] == 1.0
        g3c = T.Gz("g3_c", k)
        assert np.allclose(g3c * h, dg(zc), rtol=1e-10, atol=1e-10)
    # the reference's own geometry: g(r) = r h below the ghost
    T = TopoOracle(oracle_lib, d, h)
    assert np.array_equal(T.Gz("g", k[:nz - 1]), k[:nz - 1] * h)
    assert T.Gz("g", nz - 1) == 0.0
    assert np.allclose(T.Gz("g_c", k), zc * h, rtol=1e-13, atol=1e-11)
    assert np.allclose(T.Gz("g3", k), 1.0, rtol=1e-13)
    assert np.allclose(T.Gz("g3_c", k), 1.0, rtol=1e-13)


def test_geom_custom(oracle_lib):
    """f = 1 + elevation / (h (nz - 2)) on [-8, n+8)^2."""
    d = _dims()
    h = 100.0
    topo = Topography.gaussian_hills(d.nxt, d.nyt, h, n=3, seed=4, hmax=900.0, wmin=300.0,
                                     wmax=900.0)
    T = TopoOracle(oracle_lib, d, h, topo=topo)
    ii, jj = np.meshgrid(np.arange(-8, d.nxt + 8), np.arange(-8, d.nyt + 8), indexing="ij")
    assert np.allclose(T.F("f", ii, jj), 1.0 + topo.elevation / (h * (d.nzt - 2)), rtol=1e-15)


# ---------------------------------------------------------------------------
def _homogeneous(d, DT, DH, lam=2.0e10, mu=1.5e10, rho=2500.0, dtype=np.float64,
                 random_q=False, seed=0):
    sh = d.shape
    m = dict(d1=np.full(sh, rho, dtype), lam=np.full(sh, 1.0 / lam, dtype),
             mu=np.full(sh, 1.0 / mu, dtype), qp=np.zeros(sh, dtype), qs=np.zeros(sh, dtype),
             vx1=np.zeros(sh, dtype), vx2=np.zeros(sh, dtype), wwo=np.ones(sh, dtype),
             ww=np.ones(sh, np.int32), coeff=np.zeros(16, dtype),
             dcrjx=np.ones(sh[0], dtype), dcrjy=np.ones(sh[1], dtype), dcrjz=np.ones(sh[2], dtype))
    return m


FIELDS = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
          "r1", "r2", "r3", "r4", "r5", "r6")


def _zeros(d, dtype=np.float64):
    return {n: np.zeros(d.shape, dtype) for n in FIELDS}


def _tilted():
    d = Dims(24, 20, NZ)
    DH, DT = 50.0, 0.002
    e = Topography.flat(d.nxt, d.nyt)
    ii, jj = np.meshgrid(np.arange(-8, d.nxt + 8), np.arange(-8, d.nyt + 8), indexing="ij")
    e.elevation[...] = 0.12 * ii * DH - 0.07 * jj * DH + 300.0
    return d, DH, DT, e


def test_stress_rates_of_a_linear_velocity_field(oracle_lib):
    """Velocities linear in physical (x, y, z) on a tilted surface: one stress
    update from zero stress gives DT (lambda tr(grad v) I + mu (grad v +
    grad v^T)) at every updated point, the top six levels included."""
    d, DH, DT, e = _tilted()
    T = TopoOracle(oracle_lib, d, DH, topo=e)
    lam, mu = 2.0e10, 1.5e10
    m = _homogeneous(d, DT, DH, lam, mu)
    rng = np.random.default_rng(5)
    Gv = rng.uniform(-1e-3, 1e-3, (3, 3))  # dv_m / dx_n
    f = _zeros(d)
    for c, name in enumerate(("u1", "v1", "w1")):
        X, Y, Z = physical_xyz(T, name)
        f[name][...] = Gv[c, 0] * X + Gv[c, 1] * Y + Gv[c, 2] * Z
    T.stress(f, m, DT)
    tr = np.trace(Gv)
    want = {"xx": lam * tr + 2 * mu * Gv[0, 0], "yy": lam * tr + 2 * mu * Gv[1, 1],
            "zz": lam * tr + 2 * mu * Gv[2, 2], "xy": mu * (Gv[0, 1] + Gv[1, 0]),
            "xz": mu * (Gv[0, 2] + Gv[2, 0]), "yz": mu * (Gv[1, 2] + Gv[2, 1])}
    s_i, e_i, s_j, e_j = T.stress_bounds()
    # from level 4 up: the 7-point stencil at level 3 reaches the bottom
    # boundary point of the cell grid (z = 0, not -1/2).  xz and yz live on
    # the nodes, whose top level nz-1 is the ghost (its z rows are zero; the
    # free-surface pass sets xz = yz = 0 there).
    scale = max(abs(v) for v in want.values()) * DT
    for name, w in want.items():
        top = d.nzt - 1 if name in ("xz", "yz") else d.nzt
        box = (slice(s_i, e_i), slice(s_j, e_j), slice(ALIGN + 4, ALIGN + top))
        got = f[name][box]
        err = np.abs(got - DT * w).max() / scale
        assert err < 1e-9, (name, err)
    # outside the stress bounds nothing is written
    assert not f["xx"][:s_i].any() and not f["xx"][e_i:].any()


def test_velocity_rates_of_a_linear_stress_field(oracle_lib):
    """Stresses linear in physical (x, y, z): one velocity update from rest
    gives DT div(sigma) / rho at the interior levels."""
    d, DH, DT, e = _tilted()
    T = TopoOracle(oracle_lib, d, DH, topo=e)
    rho = 2500.0
    m = _homogeneous(d, DT, DH, rho=rho)
    rng = np.random.default_rng(6)
    S = rng.uniform(-1e3, 1e3, (6, 3))  # d sigma_c / dx_n, c in xx yy zz xy xz yz
    f = _zeros(d)
    for c, name in enumerate(("xx", "yy", "zz", "xy", "xz", "yz")):
        X, Y, Z = physical_xyz(T, name)
        f[name][...] = S[c, 0] * X + S[c, 1] * Y + S[c, 2] * Z + 1.0e5 * (c + 1)
    T.velocity(f, m, DT)
    div = {"u1": S[0, 0] + S[3, 1] + S[4, 2], "v1": S[3, 0] + S[1, 1] + S[5, 2],
           "w1": S[4, 0] + S[5, 1] + S[2, 2]}
    box = (slice(NGSL + 2, NGSL + 2 + d.nxt), slice(NGSL + 2, NGSL + 2 + d.nyt),
           slice(ALIGN + 4, ALIGN + d.nzt - 6))
    scale = max(abs(v) for v in div.values()) * DT / rho
    for name, w in div.items():
        err = np.abs(f[name][box] - DT * w / rho).max() / scale
        assert err < 1e-9, (name, err)


def test_velocity_top_rows_traction_free_field(oracle_lib):
    """Stresses c (z - z_s), zero on the surface z_s.  Flat surface: the
    one-sided rows of the top six levels (the traction penalty folded in) are
    exact.  Tilted surface: the metric terms there differentiate the quadratic
    g (z - z_s), which the boundary closure resolves to its own accuracy
    (measured 4e-3 of the rate) while the interior stays exact."""
    rho = 2500.0
    for tilt in (False, True):
        d, DH, DT, e = _tilted()
        if not tilt:
            e = Topography.flat(d.nxt, d.nyt)
        T = TopoOracle(oracle_lib, d, DH, topo=e)
        m = _homogeneous(d, DT, DH, rho=rho)
        G = (d.nzt - 2) * DH
        a, b = (0.12, -0.07) if tilt else (0.0, 0.0)
        fgrad = (a / (DH * (d.nzt - 2)), b / (DH * (d.nzt - 2)))  # df/dx, df/dy
        Dn = (-G * fgrad[0], -G * fgrad[1], 1.0)  # gradient of z - G f(x, y)
        rng = np.random.default_rng(7)
        c6 = rng.uniform(-1e3, 1e3, 6)
        f = _zeros(d)
        for c, name in enumerate(("xx", "yy", "zz", "xy", "xz", "yz")):
            X, Y, Z = physical_xyz(T, name)
            off = 300.0 if tilt else 0.0
            fx = 1.0 + (a * X + b * Y + off) / (DH * (d.nzt - 2))
            f[name][...] = c6[c] * (Z - G * fx)
        T.velocity(f, m, DT)
        C = {"u1": (0, 3, 4), "v1": (3, 1, 5), "w1": (4, 5, 2)}
        scale = np.abs(c6).max() * DT / rho
        for name, comps in C.items():
            w = sum(c6[comps[n]] * Dn[n] for n in range(3))
            # w1 lives on the nodes: its top level nz-1 is the ghost
            top = d.nzt - 1 if name == "w1" else d.nzt
            box = (slice(NGSL + 2, NGSL + 2 + d.nxt), slice(NGSL + 2, NGSL + 2 + d.nyt))
            err_int = np.abs(f[name][box + (slice(ALIGN + 4, ALIGN + d.nzt - 6),)]
                             - DT * w / rho).max() / scale
            err_top = np.abs(f[name][box + (slice(ALIGN + d.nzt - 6, ALIGN + top),)]
                             - DT * w / rho).max() / scale
            assert err_int < 1e-9, (tilt, name, err_int)
            assert err_top < (1e-2 if tilt else 1e-9), (tilt, name, err_top)


def _random_state(d, rng, dtype):
    f = {n: rng.standard_normal(d.shape).astype(dtype) for n in FIELDS}
    return f


def test_flat_topography_equals_cartesian(oracle_lib):
    """f = 1, g linear, h = 1, dt = 0.05/3 (tests/test_attenuation.cu:37-42):
    the curvilinear stress (with Q) and velocity updates equal dstrqc_new and
    dvelcx on random fields in the interior, to single-precision rounding;
    the velocity increments agree where density does not vary along x."""
    import ctypes as C
    from awp_amd import configs
    from oracle.oracle import load
    lib = load()
    d = Dims(24, 20, 32)
    DH, DT = 1.0, 0.05 / 3.0
    w = configs.anelastic(nx=d.nxt, ny=d.nyt, nz=d.nzt, nt=4, DH=DH, DT=DT)
    med = w.medium()
    m = {n: np.ascontiguousarray(getattr(med, n), np.float32)
         for n in ("d1", "lam", "mu", "qp", "qs", "vx1", "vx2", "wwo", "coeff", "lam_mu",
                   "dcrjx", "dcrjy", "dcrjz")}
    m["ww"] = np.ascontiguousarray(med.ww, np.int32)
    rng = np.random.default_rng(11)
    f0 = _random_state(d, rng, np.float32)
    T = TopoOracle(oracle_lib, d, DH, dtype=np.float32)
    ft = {n: a.copy() for n, a in f0.items()}
    fc = {n: a.copy() for n, a in f0.items()}
    T.stress(ft, m, DT)
    p = lambda a: a.ctypes.data
    s_i, e_i, s_j, e_j = T.stress_bounds()
    lib.awp_oracle_dstrqc_new_f32(
        *(p(fc[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5",
                             "r6", "u1", "v1", "w1")),
        p(m["lam"]), p(m["mu"]), p(m["qp"]), p(m["coeff"]), p(m["qs"]), p(m["dcrjx"]),
        p(m["dcrjy"]), p(m["dcrjz"]), p(m["lam_mu"]), p(m["vx1"]), p(m["vx2"]), p(m["ww"]),
        p(m["wwo"]), d.nxt, d.nyt, d.nzt, DH, DT, d.nxt, 0, 0, 0, s_i, e_i - 1, s_j, e_j - 1, 1)
    box = (slice(s_i, e_i), slice(s_j, e_j), slice(ALIGN + 3, ALIGN + d.nzt - 6))
    tot = 0.0
    for n in ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6"):
        a, b = ft[n][box].astype(np.float64), fc[n][box].astype(np.float64)
        tot += np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
    assert tot < 1e-5, tot
    # velocity: from zero velocities, so the compared values are the
    # increments DT * div(sigma) / rho themselves
    def vel_increments(m):
        ft = {n: a.copy() for n, a in f0.items()}
        fc = {n: a.copy() for n, a in f0.items()}
        for f in (ft, fc):
            for n in ("u1", "v1", "w1"):
                f[n][...] = 0
        T.velocity(ft, m, DT)
        lib.awp_oracle_dvelcx_f32(
            *(p(fc[n]) for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")),
            p(m["dcrjx"]), p(m["dcrjy"]), p(m["dcrjz"]), p(m["d1"]), d.nxt, d.nyt, d.nzt, DH,
            DT, NGSL + 2, NGSL + 1 + d.nxt, 0, 1)
        box = (slice(NGSL + 2, NGSL + 2 + d.nxt), slice(NGSL + 2, NGSL + 2 + d.nyt),
               slice(ALIGN + 3, ALIGN + d.nzt - 6))
        out = {}
        for n in ("u1", "v1", "w1"):
            a, b = ft[n][box].astype(np.float64), fc[n][box].astype(np.float64)
            assert np.abs(b).max() > 0, n
            out[n] = np.abs(a - b).max() / np.abs(b).max()
        return out

    # density varying along y and z only: the two updates agree
    mx = dict(m)
    mx["d1"] = np.ascontiguousarray(np.broadcast_to(m["d1"][m["d1"].shape[0] // 2],
                                                    m["d1"].shape))
    for n, e in vel_increments(mx).items():
        assert e < 1e-5, (n, e)
    # density varying along x: v and w differ, because the curvilinear update
    # averages density toward i-1 (optimized_velocity.cu:747-751) where
    # dvelcx averages toward i+1 (kernel.cu:212-216); u averages along j and
    # k in both and still agrees
    e = vel_increments(m)
    assert e["u1"] < 1e-5, e
    assert e["v1"] > 1e-4 and e["w1"] > 1e-4, e


def _dm_pair_workload(nt):
    from awp_amd import configs
    return configs.dm_plastic(nx=16, ny=16, nz=12, nz_fine=32, DH=300.0, DT=0.005, nt=nt,
                              nsub=4, ND=4, M0=1.0e16, f0=5.0, src_depth=3, src_grid=1)


def _x_uniform_density(media):
    """Density of each grid made constant along x (its value in the middle
    plane).  The curvilinear velocity update averages density toward i-1 for
    v and w (optimized_velocity.cu:747-751) where dvelcx averages toward i+1
    (kernel.cu:212-216), so a flat curvilinear grid equals the Cartesian one
    only where density does not vary along x."""
    for m in media:
        d1 = m.d1
        d1[...] = d1[d1.shape[0] // 2][None, :, :]
    return media


def test_flat_topography_over_dm_equals_cartesian_dm(oracle_lib):
    """A flat curvilinear grid 0 over a 3x coarser Cartesian grid
    (pmcl3d.c:1661-1694, 1843: the curvilinear updates on grid 0, dvelc2,
    dstrqc2, intp3d and swap3 as on a Cartesian stack) advances like the
    Cartesian two-grid stack.  From one random state, three steps of the
    pmcl3d order: the coarse grid and the lowest 12 levels of grid 0, away
    from the lateral edges, agree to single-precision rounding.  A grid 0 that
    skipped its align+2 level (dvelc2 or dstrqc2) or coupled on another level
    differs there by O(1)."""
    from oracle.oracle import OracleSolver
    w = _dm_pair_workload(3)
    media = _x_uniform_density(w.medium())
    d0 = media[0].dims
    flat = Topography.flat(d0.nxt, d0.nyt)
    ot = OracleSolver(media, w.sources, None, dtype=np.float32, topography=flat)
    oc = OracleSolver(media, w.sources, None, dtype=np.float32)
    rng = np.random.default_rng(5)
    for p in range(2):
        for n in FIELDS:
            a = rng.standard_normal(ot.levels[p].f[n].shape).astype(np.float32)
            ot.levels[p].f[n][...] = a
            oc.levels[p].f[n][...] = a
    ot.run(3)
    oc.run(3)
    m0, m1 = 14, 5  # lateral margins (cells) on grid 0 and grid 1
    boxes = [(slice(NGSL + 2 + m0, NGSL + 2 + d0.nxt - m0),
              slice(NGSL + 2 + m0, NGSL + 2 + d0.nyt - m0), slice(ALIGN, ALIGN + 12))]
    d1 = media[1].dims
    boxes.append((slice(NGSL + 2 + m1, NGSL + 2 + d1.nxt - m1),
                  slice(NGSL + 2 + m1, NGSL + 2 + d1.nyt - m1), slice(ALIGN, ALIGN + d1.nzt)))
    worst = 0.0
    for p, box in enumerate(boxes):
        for n in FIELDS:
            a = ot.levels[p].f[n][box].astype(np.float64)
            b = oc.levels[p].f[n][box].astype(np.float64)
            e = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
            worst = max(worst, e)
            assert e < 1e-5, (p, n, e)
    print(f"flat topography over DM vs Cartesian DM: worst relative difference {worst:.2e}")


def test_topography_over_dm_planted_mistake_is_caught(oracle_lib):
    """The comparison above fails for a grid 0 that skips dvelc2 on its
    align+2 level: the planted mistake (an oracle subclass whose velocity_dm
    does nothing on grid 0) moves the lowest levels by O(1)."""
    from oracle.oracle import OracleSolver

    class NoDvelc2(OracleSolver):
        def velocity_dm(self, p):
            if p != 0:
                super().velocity_dm(p)

    w = _dm_pair_workload(3)
    media = _x_uniform_density(w.medium())
    d0 = media[0].dims
    flat = Topography.flat(d0.nxt, d0.nyt)
    ot = NoDvelc2(media, w.sources, None, dtype=np.float32, topography=flat)
    oc = OracleSolver(media, w.sources, None, dtype=np.float32)
    rng = np.random.default_rng(5)
    for p in range(2):
        for n in FIELDS:
            a = rng.standard_normal(ot.levels[p].f[n].shape).astype(np.float32)
            ot.levels[p].f[n][...] = a
            oc.levels[p].f[n][...] = a
    ot.run(3)
    oc.run(3)
    box = (slice(NGSL + 16, NGSL + 2 + d0.nxt - 14), slice(NGSL + 16, NGSL + 2 + d0.nyt - 14),
           slice(ALIGN + 2, ALIGN + 3))
    a = ot.levels[0].f["u1"][box].astype(np.float64)
    b = oc.levels[0].f["u1"][box].astype(np.float64)
    assert np.abs(a - b).max() / np.abs(b).max() > 1e-3
