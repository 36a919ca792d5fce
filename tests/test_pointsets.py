"""Input-file point sets on a grid stack (awp_amd.pointsets), on the host:
grid coordinates, block assignment and the discontinuous-mesh mapping, with
the known answers of the reference's own tests
(tests/grid/test_grid_3d.c: grid_fill; tests/topography/sources/
test_sources_dm.c on its fixture source_dm.txt)."""
import shutil
from pathlib import Path

import numpy as np
import pytest

from awp_amd import pointsets as P
from awp_amd.grid import NGSL
from awp_amd.inputs import read_input

FIX = Path(__file__).resolve().parent / "golden" / "fixtures"
FLTOL = 1e-5


@pytest.mark.parametrize("rank", [0, 1, 3])
def test_grid_fill_known_answers(rank):
    """test_grid_fill (tests/grid/test_grid_3d.c:54-141), n = 101, h = 1,
    as rank `rank` of a partition."""
    n = 101
    cases = [  # shift, boundary1, boundary2, (x[0], x[1], x[n-1]) without rank*n
        (0, 0, 0, (0.0, 1.0, n - 1)), (0, 1, 0, (0.0, 1.0, n - 1)), (0, 0, 1, (0.0, 1.0, None)),
        (1, 0, 0, (-0.5, 0.5, n - 1.5)), (1, 1, 0, (0.0, 0.5, n - 1.5)),
        (1, 0, 1, (-0.5, 0.5, n - 2))]
    for shift, b1, b2, want in cases:
        x = P.grid_fill1(P.Grid1(rank, n, shift, b1, b2, 0, 0, np.float32(1.0)))
        assert abs(x[0] - (want[0] + rank * n)) < FLTOL
        assert abs(x[1] - (want[1] + rank * n)) < FLTOL
        last = 0.0 if want[2] is None else want[2] + rank * n
        assert abs(x[n - 1] - last) < FLTOL, (shift, b1, b2)


def test_grid_init_offsets_and_sizes():
    """grid_init's padding widens the x/y size and moves the first inner point
    (grid_3d.c:26-75); grids_init gives stress grids ngsl/2 of padding and a
    closed top on a topography block."""
    g = P.grid_init((20, 30, 40), (0, 1, 1), (0, 0, 0), (0, 0, 0), (0, 0, 1), NGSL // 2, 50.0)
    assert (g.size.x, g.size.y, g.size.z) == (28, 38, 40)
    assert (g.offset1.x, g.offset1.y, g.offset1.z) == (6, 6, 32)
    assert (g.mem.x, g.mem.y, g.mem.z) == (40, 50, 104)
    assert g.line == 104 and g.slice == 50 * 104
    grids = P.grids_init(20, 30, 40, 1, 50.0)
    assert grids[P.XX].padding.x == NGSL // 2 and grids[P.X].padding.x == 0
    assert all(grids[t].boundary2.z == 1 for t in grids)
    assert (grids[P.YZ].shift.x, grids[P.YZ].shift.y, grids[P.YZ].shift.z) == (0, 0, 0)


@pytest.fixture
def dm_input(tmp_path, monkeypatch):
    shutil.copy(FIX / "source_dm.txt", tmp_path / "source_dm.txt")
    monkeypatch.chdir(tmp_path)
    return read_input("source_dm.txt")


def _dm_stack():
    """test_sources_dm's three blocks (test_sources_dm.c:62-79): 99^2 x 11 at
    h = 1 over 33^2 x 11 at h = 3 over 11^2 x 11 at h = 9."""
    nx = ny = nz = 11
    return [P.grids_init(9 * nx, 9 * ny, nz, 0, 1.0), P.grids_init(3 * nx, 3 * ny, nz, 0, 3.0),
            P.grids_init(nx, ny, nz, 0, 9.0)]


def test_sources_dm_known_answer(dm_input):
    """The fixture's three sources at z = -2, -11, -34 fall in blocks 0, 1, 2;
    the Mxx stencils of the first two start at x index 4 and 1 past the ngsl
    ghost cells (test_sources_dm.c:125-130).  The third source's y offset
    through two coarsenings (-1 h0 - h1/2 + h0/2 - h1 - h2/2 + h1/2 = -8)
    puts it at y = -6.5, below the lower y bound of the top block's grid
    (-5), so dist_indices drops it on the one rank: the reference test's
    assertion for block 2 sits inside the loop over that block's sources and
    does not run either."""
    stack = _dm_stack()
    assert list(P.find_grid_number(dm_input, stack, np.arange(3))) == [0, 1, 2]
    S = P.PointSet(dm_input, "xx", P.XX, stack, upload=False)
    assert S.lengths == [1, 1, 0]
    assert abs(float(S.yq[1][0]) - (1.5 - 2.0)) < 1e-6
    want_ix = [4, 1]
    for j in range(2):
        err, lx, ly, lz, ix, iy, iz = S.host_coefficients(j)
        assert int(ix[0]) - NGSL == want_ix[j], (j, int(ix[0]) - NGSL)
        assert lx.shape == (1, 1) and abs(float(lx[0, 0]) - 1.0) < 1e-6


def test_block_assignment_overlap_and_surface():
    """Blocks stack downward from z = 0 without gaps: block 0 holds
    (-(nz-1) h0, 0] = (-10, 0]; block 1's top sits OVERLAP (7) cells of the
    block above higher than that block's bottom, so it reaches down to
    -10 + 7 - 30 = -33 and takes (-33, -10]; block 2 takes (-102, -33].
    Surface coordinates and z > 0 stay in block 0; a point below the stack
    is an error."""
    from awp_amd.inputs import Input
    stack = _dm_stack()
    z = np.array([0.5, -1.0, -9.9, -10.0, -3.5, -3.0, -50.0, -80.0], np.float32)
    typ = np.array([0, 0, 0, 0, 0, 0, 0, 1], np.int32)
    inp = Input(length=len(z), type=typ, x=np.zeros(len(z), np.float32),
                y=np.zeros(len(z), np.float32), z=z)
    gn = P.find_grid_number(inp, stack, np.arange(len(z)))
    assert list(gn) == [0, 0, 0, 1, 0, 0, 2, 0]
    with pytest.raises(ValueError):
        inp.z[1] = -102.5
        P.find_grid_number(inp, stack, np.arange(len(z)))


def test_topography_depth_mapping_inverts_the_metric():
    """On a topography block a volume point at depth z below the datum maps
    to the parameter coordinate (H + z) / f(x, y), H = (nz - 2) h, with f the
    metric interpolated at the point (source.c:296-330).  For f linear in x
    and y the degree-3 interpolant is exact, so the mapping must equal the
    closed form at the point's internal (x, y) (x shifted by -h/2); a shifted
    metric index would miss it by the slope times h.  Surface points go to
    the top node level H."""
    from awp_amd.inputs import Input
    nx, ny, nz, h = 30, 26, 24, 50.0
    stack = [P.grids_init(nx, ny, nz, 1, h)]
    cols = ny + 4 + 2 * NGSL + 2 * 32
    f_host = np.zeros((nx + 20, cols), np.float32)

    def F(x, y):
        return 1.0 + 2.0e-4 * x - 1.5e-4 * y

    # metric f at topographic (i, j), stored at row i + 10, column j + 42, sits
    # at the node-grid point x = h (i - 1/2), y = h j (grid_node shift (1, 0, 0))
    i = np.arange(-10, nx + 10)
    j = np.arange(-42, ny + 42)
    f_host[:, :] = F(h * (i[:, None] - 0.5), h * j[None, :]).astype(np.float32)
    z = np.array([-120.0, -300.0, -733.0, 0.0], np.float32)
    x = np.array([400.0, 730.0, 1010.0, 600.0], np.float32)
    y = np.array([310.0, 555.0, 900.0, 450.0], np.float32)
    typ = np.array([0, 0, 0, 1], np.int32)
    inp = Input(length=4, degree=3, type=typ, x=x, y=y, z=z)
    S = P.PointSet(inp, "xx", P.XX, stack, f_host, upload=False)
    H = h * (nz - 2)
    xq = (x.astype(np.float64) - 0.5 * h)
    want = (H + z[:3].astype(np.float64)) / F(xq[:3], y[:3].astype(np.float64))
    np.testing.assert_allclose(S.zq[0][:3], want, rtol=2e-6)
    assert S.zq[0][3] == np.float32(H)
    # the same points without topography map to z - lower = z + (nz - 1) h
    S0 = P.PointSet(inp, "xx", P.XX, [P.grids_init(nx, ny, nz, 0, h)], None, upload=False)
    np.testing.assert_allclose(S0.zq[0][:3], z[:3] + h * (nz - 1), rtol=1e-7)
    assert S0.zq[0][3] == np.float32(h * (nz - 1))
