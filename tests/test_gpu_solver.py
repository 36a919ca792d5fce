"""End-to-end parity: the one-GPU driver (pmcl3d step order through the C ABI)
against the oracle driver on the benchmark workloads at reduced size.

FP32 oracle: bit-exact traces and final wavefield.
FP64 oracle: receiver-trace relative L2 error within 1e-4 (FP32 rounding
accumulated over the run; the measured value is printed).
"""
import numpy as np
import pytest

from awp_amd import configs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


CASES = {
    "elastic_32": lambda: configs.elastic_halfspace(n=32, nt=60),
    "anelastic_48x40x32": lambda: configs.anelastic(nx=48, ny=40, nz=32, nt=50, DH=200.0,
                                                    DT=0.01),
}


@pytest.mark.parametrize("case", list(CASES))
def test_solver_matches_oracle(torch_gpu, oracle_lib, case):
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    w = CASES[case]()
    med = w.medium()
    g = Solver(med, w.sources, w.receivers, nt_record=w.nt)
    g.run(w.nt, check_every=10)
    tg = g.traces()
    o32 = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float32)
    o32.run(w.nt)
    assert np.array_equal(tg, o32.traces()), rel_l2(tg, o32.traces())
    for n in ("u1", "v1", "w1", "xx", "zz", "xz", "r1", "r5"):
        assert np.array_equal(g.field(n), o32.f[n]), n
    o64 = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float64)
    o64.run(w.nt)
    err = rel_l2(tg.astype(np.float64), o64.traces())
    print(f"{case}: receiver rel-L2 vs FP64 oracle = {err:.3e}")
    assert np.abs(o64.traces()).max() > 0
    assert err < 1e-4


def test_plastic_solver_matches_oracle(torch_gpu, oracle_lib):
    """NVE = 3 step order (region split, source before the inner stress box,
    overlapping yield-factor strips) through the C ABI, against the oracle."""
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    w = configs.plastic(nx=40, ny=44, nz=32, nt=60, nsub=12, M0=1.0e18, f0=5.0)
    med = w.medium()
    pm = w.params["plastic_fn"](med)
    g = Solver(med, w.sources, w.receivers, nt_record=w.nt, plastic=pm)
    g.run(w.nt, check_every=10)
    o32 = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float32, plastic=pm)
    o32.run(w.nt)
    nyield = int((o32.f["yldfac"] < 1).sum())
    print(f"plastic: {nyield} yielding cells, max neta {o32.f['neta'].max():.3e}")
    assert nyield > 0
    assert np.array_equal(g.traces(), o32.traces())
    for n in ("u1", "w1", "xx", "xz", "yz", "yldfac", "neta"):
        assert np.array_equal(g.field(n), o32.f[n]), n


@pytest.mark.parametrize("texture", ["table", "rule"])
def test_texture_table_matches_arrays(torch_gpu, oracle_lib, texture):
    """awp_dstrqc_tex (vx1, vx2, wwo from an 8-entry table keyed by the
    mechanism, ww read or from init_texture's rule) against dstrqc_H and the
    oracle on odd and even grid sizes."""
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    for nx, ny, nz in ((47, 41, 36), (40, 38, 32)):
        w = configs.anelastic(nx=nx, ny=ny, nz=nz, nt=40, DH=200.0, DT=0.01)
        print(f"texture {texture}: {w.name}", flush=True)
        med = w.medium()
        g = Solver(med, w.sources, w.receivers, nt_record=w.nt, texture=texture)
        g.run(w.nt, check_every=10)
        a = Solver(med, w.sources, w.receivers, nt_record=w.nt)
        a.run(w.nt)
        o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float32)
        o.run(w.nt)
        assert np.abs(g.traces()).max() > 0
        assert np.array_equal(g.traces(), o.traces())
        for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4",
                  "r5", "r6"):
            assert np.array_equal(g.field(n), a.field(n)), n
            assert np.array_equal(g.field(n), o.f[n]), n


STATE15 = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5",
           "r6")


@pytest.mark.parametrize("texture", ["table", "rule"])
@pytest.mark.parametrize("shape", [(47, 41, 36), (40, 38, 100), (37, 29, 70), (26, 24, 128)])
def test_fused_step_matches_separate_launchers(torch_gpu, oracle_lib, texture, shape):
    """awp_step (velocity, free-surface ghosts and stress in one sweep, state
    A -> state B) against dvelcx_H_opt + dstrqc_H in place and the oracle:
    first 2 steps from a random state in every cell of every array (ghost
    cells included), then a 40-step run from rest with the source."""
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    nx, ny, nz = shape
    w = configs.anelastic(nx=nx, ny=ny, nz=nz, nt=40, DH=200.0, DT=0.01)
    print(f"fused {texture}: {w.name}", flush=True)
    med = w.medium()
    # random state
    a = Solver(med, w.sources, w.receivers, nt_record=3)
    g = Solver(med, w.sources, w.receivers, nt_record=3, texture=texture, fused=True)
    rng = np.random.default_rng(nx * 1000 + nz)
    for n in STATE15:
        x = torch_gpu.from_numpy(rng.standard_normal(a.dims.shape, dtype=np.float32))
        a.f[n].copy_(x)
        g.f[n].copy_(x)
    a.start(); g.start()
    for _ in range(2):
        a.step(); g.step()
    torch_gpu.cuda.synchronize()
    for n in STATE15:
        assert np.array_equal(g.field(n), a.field(n)), n
    assert np.array_equal(g.traces(), a.traces())
    del a, g
    # from rest, 40 steps, against the oracle
    g = Solver(med, w.sources, w.receivers, nt_record=w.nt, texture=texture, fused=True)
    g.run(w.nt, check_every=10)
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float32)
    o.run(w.nt)
    # the source has injected (deep sources need not reach the receivers yet)
    assert np.abs(o.f["xx"]).max() > 0
    assert np.array_equal(g.traces(), o.traces())
    for n in STATE15:
        assert np.array_equal(g.field(n), o.f[n]), n


@pytest.mark.parametrize("kw", [32, 64])
def test_k_run_width_matches_oracle(torch_gpu, oracle_lib, kw):
    """The Cartesian kernels with a k run of 64 levels of one j row per wave
    and of 32 levels of each of two rows (awp_set_tuning("kw"), which the
    idle-lane rule otherwise picks by the k extent): an NVE = 3 run on a grid
    of 70 levels, odd in i and j, with the texture rule and with the per-cell
    texture arrays, bit-exact against the FP32 oracle either way."""
    from awp_amd import _lib
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    lib = _lib.load()
    w = configs.plastic(nx=37, ny=29, nz=70, nt=40, nsub=12, M0=1.0e18, f0=5.0)
    med = w.medium()
    pm = w.params["plastic_fn"](med)
    o32 = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float32, plastic=pm)
    o32.run(w.nt)
    assert int((o32.f["yldfac"] < 1).sum()) > 0
    assert lib.awp_set_tuning(b"kw", kw) == 0
    try:
        for texture in ("rule", "arrays"):
            g = Solver(med, w.sources, w.receivers, nt_record=w.nt, plastic=pm, texture=texture)
            g.run(w.nt, check_every=10)
            torch_gpu.cuda.synchronize()
            assert np.array_equal(g.traces(), o32.traces()), texture
            for n in ("u1", "v1", "w1", "xx", "yy", "xz", "yz", "r1", "r4", "yldfac", "neta"):
                assert np.array_equal(g.field(n), o32.f[n]), (texture, n)
            del g
    finally:
        assert lib.awp_set_tuning(b"kw", 0) == 0
