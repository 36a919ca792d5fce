"""GPU parity of the discontinuous-mesh launchers (dvelc2_H, dstrqc2_H,
intp3d_H, swap_H) and of two-grid runs, through the C ABI, against the FP32
oracle on the same inputs.  Bar: bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

from awp_amd.grid import Dims
from tests._fields import WAVE, first_mismatch, random_state

pytestmark = pytest.mark.gpu

PACK9 = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")
SWAP_ORDER = ("xx", "yy", "zz", "xy", "xz", "yz", "u1", "v1", "w1")
# (coarse, fine): coarse ranges nxt+5 / nyt+5 not multiples of the 4 x 16
# swap tile, several tiles, nzt at the minimum the coupling allows
PAIRS = [(Dims(5, 4, 10), Dims(15, 12, 12)),
         (Dims(23, 37, 20), Dims(69, 111, 16)),
         (Dims(48, 16, 4), Dims(144, 48, 9))]


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awp_amd import _lib
    return torch, _lib.load(), _lib


def _set_grids(lib, fine: Dims, coarse: Dims, DHf=100.0, DT=0.005):
    F, I = C.c_float, C.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 2)(DHf, 3 * DHf), F(DT), (I * 2)(fine.nxt, coarse.nxt),
                            (I * 2)(fine.nyt, coarse.nyt), (I * 2)(fine.nzt, coarse.nzt), 2, 0.0,
                            0.0, z9, z9)


def _dev(torch, d):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}


def _p(t):
    return t.data_ptr()


def _np(a):
    return a.ctypes.data


def _stream(torch):
    return torch.cuda.current_stream().cuda_stream


def _compare(torch, dev, host, names, tag):
    torch.cuda.synchronize()
    for n in names:
        mm = first_mismatch(dev[n].cpu().numpy(), host[n])
        assert mm is None, f"{tag}: field {n} differs: {mm}"


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"{p[1].nxt}x{p[1].nyt}x{p[1].nzt}")
def test_dvelc2_bit_exact(gpu, oracle_lib, pair):
    torch, lib, _lib = gpu
    coarse, fine = pair
    DH, DT = 100.0, 0.005
    _set_grids(lib, fine, coarse, DH, DT)
    st, med = random_state(fine, seed=31 + fine.nxt)
    dev = _dev(torch, {**st, **med})
    lib.dvelc2_H(*(_p(dev[n]) for n in PACK9), _p(dev["dcrjx"]), _p(dev["dcrjy"]),
                 _p(dev["dcrjz"]), _p(dev["d1"]), fine.nxt, fine.nyt, _stream(torch), 0)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    oracle_lib.load().awp_oracle_dvelc2_f32(
        *(_np(host[n]) for n in PACK9), _np(med["dcrjx"]), _np(med["dcrjy"]), _np(med["dcrjz"]),
        _np(med["d1"]), fine.nxt, fine.nyt, fine.nzt, DH, DT)
    _compare(torch, dev, host, WAVE, "dvelc2")


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"{p[1].nxt}x{p[1].nyt}x{p[1].nzt}")
@pytest.mark.parametrize("sub", [None, (12, 30, 11, 13)])
def test_dstrqc2_bit_exact(gpu, oracle_lib, pair, sub):
    torch, lib, _lib = gpu
    coarse, fine = pair
    DH, DT = 100.0, 0.005
    _set_grids(lib, fine, coarse, DH, DT)
    st, med = random_state(fine, seed=41 + fine.nyt)
    (s_i, e_i), (s_j, e_j) = fine.i_range, fine.j_range
    if sub is not None:
        s_i, e_i, s_j, e_j = sub[0], min(sub[1], e_i), sub[2], min(sub[3], e_j)
    dev = _dev(torch, {**st, **med})
    names = ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6", "u1", "v1",
             "w1")
    lib.dstrqc2_H(*(_p(dev[n]) for n in names), _p(dev["lam"]), _p(dev["mu"]), _p(dev["qp"]),
                  _p(dev["qs"]), _p(dev["dcrjx"]), _p(dev["dcrjy"]), _p(dev["dcrjz"]), fine.nxt,
                  fine.nyt, _stream(torch), _p(dev["coeff"]), _p(dev["vx1"]), _p(dev["vx2"]),
                  _p(dev["ww"]), _p(dev["wwo"]), s_i, e_i, s_j, e_j, 0)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    oracle_lib.load().awp_oracle_dstrqc2_f32(
        *(_np(host[n]) for n in names), _np(med["lam"]), _np(med["mu"]), _np(med["qp"]),
        _np(med["qs"]), _np(med["dcrjx"]), _np(med["dcrjy"]), _np(med["dcrjz"]),
        _np(med["coeff"]), _np(med["vx1"]), _np(med["vx2"]), _np(med["ww"]), _np(med["wwo"]),
        fine.nxt, fine.nyt, fine.nzt, DH, DT, s_i, e_i, s_j, e_j)
    _compare(torch, dev, host, WAVE, "dstrqc2")


def _two_states(coarse, fine, seed):
    lo, _ = random_state(coarse, seed=seed)
    hi, _ = random_state(fine, seed=seed + 1)
    return lo, hi


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"{p[0].nxt}x{p[0].nyt}x{p[0].nzt}")
def test_intp3d_bit_exact(gpu, oracle_lib, pair):
    torch, lib, _lib = gpu
    coarse, fine = pair
    _set_grids(lib, fine, coarse)
    lo, hi = _two_states(coarse, fine, 51 + coarse.nxt)
    dlo, dhi = _dev(torch, lo), _dev(torch, hi)
    lib.intp3d_H(*(_p(dlo[n]) for n in PACK9), *(_p(dhi[n]) for n in PACK9), coarse.nxt,
                 coarse.nyt, 0, _stream(torch), 1)
    _lib.check()
    hlo = {k: v.copy() for k, v in lo.items()}
    hhi = {k: v.copy() for k, v in hi.items()}
    pk = lambda d: (C.c_void_p * 9)(*[_np(d[n]) for n in PACK9])
    oracle_lib.load().awp_oracle_intp3d_f32(pk(hlo), pk(hhi), coarse.nxt, coarse.nyt,
                                            coarse.nzt, fine.nxt, fine.nyt, fine.nzt)
    _compare(torch, dhi, hhi, PACK9, "intp3d fine")
    _compare(torch, dlo, hlo, PACK9, "intp3d coarse")


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"{p[0].nxt}x{p[0].nyt}x{p[0].nzt}")
def test_swap3_bit_exact(gpu, oracle_lib, pair):
    torch, lib, _lib = gpu
    coarse, fine = pair
    _set_grids(lib, fine, coarse)
    lo, hi = _two_states(coarse, fine, 61 + coarse.nyt)
    dlo, dhi = _dev(torch, lo), _dev(torch, hi)
    lib.swap_H(*(_p(dlo[n]) for n in SWAP_ORDER), *(_p(dhi[n]) for n in SWAP_ORDER),
               coarse.nxt, coarse.nyt, None, None, None, None, 0, _stream(torch), 0)
    _lib.check()
    hlo = {k: v.copy() for k, v in lo.items()}
    hhi = {k: v.copy() for k, v in hi.items()}
    pk = lambda d: (C.c_void_p * 9)(*[_np(d[n]) for n in PACK9])
    oracle_lib.load().awp_oracle_swap3_f32(pk(hlo), pk(hhi), coarse.nxt, coarse.nyt,
                                           coarse.nzt, fine.nxt, fine.nyt, fine.nzt)
    _compare(torch, dlo, hlo, PACK9, "swap3 coarse")
    _compare(torch, dhi, hhi, PACK9, "swap3 fine")


def test_dm_launchers_reject_bad_grids(gpu):
    """A grid pair that is not 3x refined, or a size disagreeing with the
    table, launches nothing and sets the error channel."""
    torch, lib, _lib = gpu
    F, I = C.c_float, C.c_int
    z9 = (F * 9)()
    fine, coarse = Dims(30, 30, 12), Dims(8, 10, 10)  # 30 != 3*8
    lib.SetDeviceConstValue((F * 2)(100.0, 300.0), F(0.005), (I * 2)(fine.nxt, coarse.nxt),
                            (I * 2)(fine.nyt, coarse.nyt), (I * 2)(fine.nzt, coarse.nzt), 2, 0.0,
                            0.0, z9, z9)
    lo, hi = _two_states(coarse, fine, 3)
    dlo, dhi = _dev(torch, lo), _dev(torch, hi)
    _lib.load().awp_clear_error()
    lib.intp3d_H(*(_p(dlo[n]) for n in PACK9), *(_p(dhi[n]) for n in PACK9), coarse.nxt,
                 coarse.nyt, 0, _stream(torch), 1)
    assert lib.awp_last_error() != 0
    _lib.load().awp_clear_error()
    lib.swap_H(*(_p(dlo[n]) for n in SWAP_ORDER), *(_p(dhi[n]) for n in SWAP_ORDER),
               coarse.nxt + 1, coarse.nyt, None, None, None, None, 0, _stream(torch), 0)
    assert lib.awp_last_error() != 0
    _lib.load().awp_clear_error()
    torch.cuda.synchronize()
    for n in PACK9:
        assert np.array_equal(dlo[n].cpu().numpy(), lo[n])


def _dm_workload(src_grid, nt, M0):
    from awp_amd import configs
    return configs.dm_plastic(nx=16, ny=16, nz=12, nz_fine=18, DH=300.0, DT=0.005, nt=nt,
                              nsub=8, ND=4, M0=M0, f0=5.0, src_depth=6, src_grid=src_grid)


def test_dm_solver_anelastic_matches_oracle(gpu, oracle_lib):
    """Two grids, NVE = 1, source in the coarse grid: bit-exact traces and
    wavefields of both grids against the FP32 oracle; FP64 oracle within
    1e-4 relative L2."""
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    nt = 120
    w = _dm_workload(1, nt, 1.0e16)
    media = w.medium()
    g = Solver(media, w.sources, w.receivers, nt_record=nt)
    g.run(nt, check_every=20)
    o32 = OracleSolver(media, w.sources, w.receivers, nt_record=nt, dtype=np.float32)
    o32.run(nt)
    tg = g.traces()
    assert np.abs(o32.traces()).max() > 0
    assert np.array_equal(tg, o32.traces())
    for p in range(2):
        for n in ("u1", "v1", "w1", "xx", "zz", "xy", "yz", "r1", "r4"):
            assert np.array_equal(g.field(n, p), o32.levels[p].f[n]), (p, n)
    o64 = OracleSolver(media, w.sources, w.receivers, nt_record=nt, dtype=np.float64)
    o64.run(nt)
    err = float(np.linalg.norm(tg - o64.traces()) / np.linalg.norm(o64.traces()))
    print(f"dm anelastic: receiver rel-L2 vs FP64 oracle = {err:.3e}")
    assert err < 1e-4


@pytest.mark.parametrize("texture", ["arrays", "auto", "table"])
def test_dm_solver_plastic_matches_oracle(gpu, oracle_lib, texture):
    """Two grids, NVE = 3 (region split on both grids, yield-factor strips,
    plastic correction), source in the fine grid: bit-exact against the FP32
    oracle, with yielding on the fine grid.  texture "auto": the fine grid's
    stress updates (dstrqc_H's and dstrqc2_H's) evaluate init_texture's parity
    rule; the coarse grid, whose top levels carry the fine grid's texture
    (pmcl3d.c:1049-1058), reads ww and the 8-entry table.  "table": ww and the
    table on both grids (awp_dstrqc_tex, awp_dstrqc2_tex)."""
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    nt = 50
    w = _dm_workload(0, nt, 1.0e17)
    media = w.medium()
    pm = w.params["plastic_fn"](media)
    g = Solver(media, w.sources, w.receivers, nt_record=nt, plastic=pm, texture=texture)
    if texture == "auto":
        assert [L.texmode for L in g.levels] == ["rule", "table"]
    if texture == "table":
        assert [L.texmode for L in g.levels] == ["table", "table"]
    g.run(nt, check_every=10)
    o32 = OracleSolver(media, w.sources, w.receivers, nt_record=nt, dtype=np.float32,
                       plastic=pm)
    o32.run(nt)
    nyield = int((o32.levels[0].f["yldfac"] < 1).sum())
    print(f"dm plastic: {nyield} yielding cells on the fine grid")
    assert nyield > 0
    assert np.array_equal(g.traces(), o32.traces())
    for p in range(2):
        for n in ("u1", "w1", "xx", "xz", "yldfac", "neta"):
            assert np.array_equal(g.field(n, p), o32.levels[p].f[n]), (p, n)
