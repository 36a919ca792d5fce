"""Lagrange interpolation oracle (oracle/awp_oracle_interp.cpp) and the
library's host coefficient code, on the CPU.

The oracle is pinned by
- the known answers of the reference's own tests
  (tests/interpolation/test_interpolation.c, tests/interpolation/test_lagrange.c);
- the closed form of the Lagrange basis, evaluated in float64;
- polynomial reproduction in 3-D on non-uniform grids and the moments of the
  source weights (partition of unity, first moments = the source location);
- special cases in which the curvilinear and force sources reduce to the
  Cartesian one exactly.
The library's host code (cuinterp_lagrange_h) must then equal the oracle bit
for bit.
"""
import numpy as np
import pytest

from awp_amd.grid import ALIGN, NGSL, Dims
from awp_amd.interpolation import GridDesc, grid_coordinates, metric_grid

FLTOL = 1e-5  # include/awp/definitions.h:43


@pytest.fixture(scope="module")
def oi(oracle_lib):
    from oracle import interp
    return interp


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


# ---- known answers of the reference's tests ---------------------------------

def test_argnearest_known_answers(oi):
    # test_interpolation.c:53-104
    x = np.arange(4, dtype=np.float32)
    assert oi.argnearest(x, 2.1) == (0, 2)
    assert oi.argnearest(x, 0.5) == (0, 0)
    assert oi.argnearest(x, -1) == (700, 0)
    assert oi.argnearest(x, 5)[0] == 701


def test_argnearest_range_known_answers(oi):
    # test_interpolation.c:106-174 (n = 10, x = 0..9)
    x = np.arange(10, dtype=np.float32)
    for q, lu, want_nearest, want in ((4.3, 0, 4, (4, 5)), (4.3, 1, 4, (3, 6)),
                                      (0.1, 2, 0, (0, 5)), (8.9, 2, 9, (5, 10))):
        err, nearest = oi.argnearest(x, q)
        assert (err, nearest) == (0, want_nearest)
        err, first, last = oi.argnearest_range(lu, lu, nearest, 10, q)
        assert (err, first, last) == (0, *want)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_lagrange1_reproduces_monomials(oi, p):
    # test_interpolation.c:176-221: x^p on x = 0..9 at p+1 queries
    x = np.arange(10, dtype=np.float32)
    q = np.array([0.25, 0.625, 0.75, 0.5][:p + 1], np.float32)
    out = oi.lagrange1(x, x.astype(np.float64) ** p, q, p)
    assert np.all(np.abs(out - q.astype(np.float64) ** p) < FLTOL)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_lagrange1_coef_brackets_and_sums_to_one(oi, p):
    # test_interpolation.c:223-280: queries 1.05 and 0.95 on x = 0..9
    x = np.arange(10, dtype=np.float32)
    for q in (1.05, 0.95):
        err, xloc, l, first = oi.lagrange1_coef(x, q, p)
        assert err == 0
        if p > 0:
            assert xloc[0] <= np.float32(q) <= xloc[p]
        assert abs(float(np.sum(l, dtype=np.float64)) - 1.0) < FLTOL
        assert np.array_equal(xloc, x[first:first + p + 1])


def test_lagrange_basis_known_answers(oi):
    # test_lagrange.c:28-90: nodes 0..3
    x = np.arange(4, dtype=np.float32)
    assert oi.lagrange_basis(x, 0.0)[0] == 0
    assert oi.lagrange_basis(x, -0.1)[0] == 600
    assert oi.lagrange_basis(x, 3.1)[0] == 601
    err, l = oi.lagrange_basis(x, 0.5)
    assert err == 0
    assert abs(float(np.sum(l, dtype=np.float64)) - 1.0) < FLTOL
    assert abs(float(np.sum(l * x.astype(np.float64) ** 3)) - 0.5 ** 3) < FLTOL


def _field_of(dims, grid, x, y, z, fn):
    """A field whose value at coordinate point (i, j, k) is fn(x_i, y_j, z_k)."""
    a = np.zeros(dims.size, np.float32)
    X, Y, Z = np.meshgrid(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64),
                          indexing="ij")
    ii, jj, kk = np.meshgrid(np.arange(len(x)), np.arange(len(y)), np.arange(len(z)),
                             indexing="ij")
    pos = grid.off_z + kk + (grid.off_y + jj) * grid.line + (grid.off_x + ii) * grid.slice
    a[pos.ravel()] = fn(X, Y, Z).ravel()
    return a


def test_lagrange3_known_answers(oi):
    # test_interpolation.c:282-393: degree 3 on a 10^3 grid with h = 1/10, the
    # cubes x^3, y^3, z^3 at four query points, to FLTOL.  The grid is the
    # reference test's: ngsl padding in x and y, staggered in y and z.
    n, h = 10, np.float32(0.1)
    dims = Dims(n, n, n)
    grid = metric_grid(dims)
    x = grid_coordinates(grid.size_x, h, shift=0, padding=NGSL)
    y = grid_coordinates(grid.size_y, h, shift=1, padding=NGSL)
    z = grid_coordinates(grid.size_z, h, shift=1, padding=0)
    qx = np.array([0.0, 0.2, 0.4, 0.9], np.float32)
    qy = np.array([0.0, 0.7, 0.4, 0.7], np.float32)
    qz = np.array([0.0, 0.2, 0.3, 0.8], np.float32)
    err, *coef = oi.lagrange_h(x, y, z, qx, qy, qz, 3)
    for axis, q in ((0, qx), (1, qy), (2, qz)):
        fld = _field_of(dims, grid, x, y, z, lambda X, Y, Z: (X, Y, Z)[axis] ** 3)
        out = np.zeros(4, np.float32)
        oi.interp(out, fld, coef, np.arange(4), grid)
        want = q.astype(np.float64) ** 3
        assert np.all(np.abs(out - want) < FLTOL), (axis, out, want)


# ---- mathematics ------------------------------------------------------------

def _nonuniform(rng, n, h=1.0):
    """Strictly increasing float32 nodes with spacing in [0.5h, 1.5h]."""
    return np.cumsum(rng.uniform(0.5 * h, 1.5 * h, n)).astype(np.float32)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5, 6])
def test_weights_match_closed_form(oi, deg):
    rng = np.random.default_rng(10 + deg)
    for _ in range(200):
        xs = _nonuniform(rng, deg + 1)
        x0 = np.float32(rng.uniform(xs[0], xs[-1]))
        err, l = oi.lagrange_basis(xs, x0)
        assert err == 0
        X = xs.astype(np.float64)
        want = np.array([np.prod([(np.float64(x0) - X[k]) / (X[j] - X[k])
                                  for k in range(deg + 1) if k != j])
                         for j in range(deg + 1)])
        # at a node (within FLTOL) the reference returns 1 there
        near = np.abs(np.float64(x0) - X) < FLTOL
        if near.any():
            continue
        assert np.allclose(l, want, rtol=2e-5, atol=2e-6 * np.abs(want).max()), (xs, x0, l, want)


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5])
def test_3d_polynomial_reproduction_nonuniform(oi, deg):
    """A polynomial of degree <= deg in each direction is reproduced exactly
    (to float32 rounding) by degree-deg interpolation on any grid."""
    rng = np.random.default_rng(20 + deg)
    dims = Dims(12, 10, 14)
    grid = metric_grid(dims)
    x = _nonuniform(rng, grid.size_x, 0.1)
    y = _nonuniform(rng, grid.size_y, 0.1)
    z = _nonuniform(rng, grid.size_z, 0.1)
    nq = 300
    qx = rng.uniform(x[0], x[-1], nq).astype(np.float32)
    qy = rng.uniform(y[0], y[-1], nq).astype(np.float32)
    qz = rng.uniform(z[0], z[-1], nq).astype(np.float32)
    c = rng.standard_normal(3)
    fn = lambda X, Y, Z: (1.0 + c[0] * X ** deg) * (0.5 + c[1] * Y ** deg) * (2 + c[2] * Z ** deg)
    fld = _field_of(dims, grid, x, y, z, fn)
    err, *coef = oi.lagrange_h(x, y, z, qx, qy, qz, deg)
    out = np.zeros(nq, np.float32)
    oi.interp(out, fld, coef, np.arange(nq), grid)
    want = fn(qx.astype(np.float64), qy.astype(np.float64), qz.astype(np.float64))
    scale = max(1.0, float(np.abs(fld).max()))
    assert np.max(np.abs(out - want)) < 2e-5 * scale * (deg + 1) ** 3


def _one_source(oi, deg, q, value=1.0):
    dims = Dims(12, 12, 12)
    grid = metric_grid(dims)
    h = np.float32(0.1)
    x = grid_coordinates(grid.size_x, h, padding=NGSL)
    y = grid_coordinates(grid.size_y, h, padding=NGSL)
    z = grid_coordinates(grid.size_z, h)
    err, *coef = oi.lagrange_h(x, y, z, [q[0]], [q[1]], [q[2]], deg)
    out = np.zeros(dims.size, np.float32)
    oi.source_add("cartesian", out, [value], coef, [0], grid, h, 0.01)
    return dims, grid, (x, y, z), out


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_source_moments(oi, deg):
    """The injected amounts sum to -dt/h^3 v (partition of unity) and, for
    degree >= 1, their first moments put the source at its location."""
    q = (0.337, 0.512, 0.448)
    dims, grid, (x, y, z), out = _one_source(oi, deg, q, value=2.0)
    dth = np.float64(np.float32(0.01) / (np.float32(0.1) ** 3))
    total = -dth * 2.0
    nz_idx = np.nonzero(out)[0]
    assert len(nz_idx) == (deg + 1) ** 3 or deg == 0
    assert abs(out.sum(dtype=np.float64) - total) < 1e-5 * abs(total)
    if deg >= 1:
        p = nz_idx - grid.off_z
        k = p % grid.line
        j = (p // grid.line) % (grid.slice // grid.line) - grid.off_y
        i = p // grid.slice - grid.off_x
        w = out[nz_idx].astype(np.float64) / total
        assert abs(np.sum(w * x[i]) - q[0]) < 1e-5
        assert abs(np.sum(w * y[j]) - q[1]) < 1e-5
        assert abs(np.sum(w * z[k]) - q[2]) < 1e-5


def test_curvilinear_and_force_reduce_to_cartesian(oi):
    rng = np.random.default_rng(5)
    dims = Dims(12, 10, 16)
    grid = metric_grid(dims)
    h = np.float32(0.1)
    x = grid_coordinates(grid.size_x, h, padding=NGSL)
    y = grid_coordinates(grid.size_y, h, padding=NGSL)
    z = grid_coordinates(grid.size_z, h)
    nq = 40
    qx = rng.uniform(0.0, 1.1, nq)
    qy = rng.uniform(0.0, 0.9, nq)
    qz = rng.uniform(0.0, 1.4, nq)
    err, *coef = oi.lagrange_h(x, y, z, qx, qy, qz, 3)
    vals = rng.standard_normal(nq).astype(np.float32)
    base = rng.standard_normal(dims.size).astype(np.float32)
    fsz = (dims.nxt + 20) * (dims.nyt + 84)
    gsz = dims.nzt + 2 * ALIGN

    def run(kind, **kw):
        out = base.copy()
        oi.source_add(kind, out, vals, coef, np.arange(nq), grid, h, 0.005, ny=dims.nyt,
                      nz=dims.nzt, **kw)
        return out

    cart = run("cartesian")
    ones_f, ones_g, ones_r = np.ones(fsz), np.ones(gsz), np.ones(dims.size)
    assert np.array_equal(bits(run("curvilinear", f=ones_f, dg=ones_g)), bits(cart))
    assert np.array_equal(bits(run("force", f=ones_f, dg=ones_g, d1=ones_r, quad_weight=-1.0)),
                          bits(cart))
    # a Jacobian of 2 halves every term (exact in binary floating point)
    half = base.copy()
    oi.source_add("cartesian", half, vals * np.float32(0.5), coef, np.arange(nq), grid, h, 0.005)
    assert np.array_equal(bits(run("curvilinear", f=2 * ones_f, dg=ones_g)), bits(half))
    assert np.array_equal(bits(run("force", f=ones_f, dg=ones_g, d1=2 * ones_r,
                                   quad_weight=-1.0)), bits(half))


def test_coincident_sources_add_in_query_order(oi):
    """Two sources at one point add their terms one after the other."""
    q = (0.337, 0.512, 0.448)
    dims, grid, xyz, once = _one_source(oi, 3, q, value=1.0)
    x, y, z = xyz
    err, *coef = oi.lagrange_h(x, y, z, [q[0]] * 2, [q[1]] * 2, [q[2]] * 2, 3)
    twice = np.zeros(dims.size, np.float32)
    oi.source_add("cartesian", twice, [1.0, 1.0], coef, [0, 1], grid, np.float32(0.1), 0.01)
    assert np.array_equal(bits(twice), bits(once + once))


# ---- the library's host code equals the oracle --------------------------------

@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5, 6])
def test_library_coefficients_match_oracle(oi, deg):
    from awp_amd.interpolation import lagrange_h
    rng = np.random.default_rng(30 + deg)
    grid = GridDesc.plain(2, 2, ALIGN, 104, 104 * 60, 37, 29, 41, 10 ** 7)
    x = _nonuniform(rng, grid.size_x, 0.1) - 1.0
    y = grid_coordinates(grid.size_y, 0.1, shift=1, padding=NGSL)
    z = _nonuniform(rng, grid.size_z, 0.05)
    nq = 2000
    qx = rng.uniform(x[0] - 0.2, x[-1] + 0.2, nq)          # some outside: clamped
    qy = rng.uniform(y[0] - 0.2, y[-1] + 0.2, nq)
    qz = rng.uniform(z[0], z[-1], nq)
    qx[:50] = x[rng.integers(0, len(x), 50)]                # on nodes
    qy[50:70] = 0.5 * (y[5:25] + y[6:26])                   # midpoints: ties
    want = oi.lagrange_h(x, y, z, qx, qy, qz, deg)
    got = lagrange_h(x, y, z, grid, qx, qy, qz, deg)
    assert got[0] == want[0]
    for g, w in zip(got[1:4], want[1:4]):
        assert np.array_equal(bits(g), bits(w))
    for g, w in zip(got[4:], want[4:]):
        assert np.array_equal(g, w)


def test_init_rejects_an_empty_query_set():
    from awp_amd.interpolation import Interpolation
    grid = metric_grid(Dims(8, 8, 8))
    x = np.arange(grid.size_x, dtype=np.float32)
    y = np.arange(grid.size_y, dtype=np.float32)
    z = np.arange(grid.size_z, dtype=np.float32)
    with pytest.raises(ValueError, match="ERR_NON_POSITIVE"):
        Interpolation(x, y, z, grid, [], [], [], degree=3)
