"""Curvilinear topography on the GPU, against the oracle (bit for bit):
the metric builders, one call of each launcher on random fields over a hilly
surface, whole runs (FP32 and 16-bit storage) and the launchers' error
paths."""
import numpy as np
import pytest

from awp_amd import configs
from awp_amd.topography import F_NAMES, G_NAMES, Topography

NX, NY, NZ = 40, 36, 32


def _workload(nt=40):
    w = configs.topography_hills(nx=NX, ny=NY, nz=NZ, nt=nt, DH=100.0, DT=0.005, nhills=4,
                                 hmax=600.0, wmin=700.0, wmax=1800.0, nrec=8)
    return w


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awp_amd import _lib
    return torch, _lib


def test_neighbour_ranks_are_rejected():
    """The halo-strip launchers have nothing to do without neighbours and
    refuse a neighbour (no GPU call is made)."""
    import ctypes as C
    from awp_amd import _lib
    from awp_amd.refstructs import Topo
    lib = _lib.load()
    T = Topo()
    T.use = 1
    T.x_rank_l = T.x_rank_r = T.y_rank_f = T.y_rank_b = -1
    lib.awp_clear_error()
    for fn in (lib.topo_velocity_front_H, lib.topo_velocity_back_H, lib.topo_stress_left_H,
               lib.topo_stress_right_H):
        fn(C.byref(T))
    assert lib.awp_last_error() == 0
    T.y_rank_f = 3
    lib.topo_velocity_front_H(C.byref(T))
    assert lib.awp_last_error() == 1
    lib.awp_clear_error()
    T.use = 1
    lib.topo_stress_interior_H(C.byref(T))  # no constants, no arrays
    assert lib.awp_last_error() != 0
    lib.awp_clear_error()


@pytest.mark.gpu
def test_gpu_metrics_match_oracle(gpu, oracle_lib):
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    w = _workload()
    med = w.medium()
    topo = w.params["topography"]
    g = Solver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    got = g.topo.metrics()
    for n in F_NAMES:
        assert np.array_equal(got[n].view(np.uint32), o.topo.mf[n].view(np.uint32)), n
    for n in G_NAMES:
        assert np.array_equal(got[n].view(np.uint32), o.topo.mg[n].view(np.uint32)), n


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["velocity", "stress"])
def test_gpu_launcher_one_call_bit_exact(gpu, oracle_lib, which):
    """Random O(1) fields (memory variables too) over the hilly surface."""
    torch, _lib = gpu
    from awp_amd.solver import FIELDS, Solver
    from oracle.oracle import OracleSolver
    w = _workload()
    med = w.medium()
    topo = w.params["topography"]
    g = Solver(med, None, None, topography=topo)
    o = OracleSolver(med, None, None, topography=topo)
    rng = np.random.default_rng(21)
    for n in FIELDS:
        a = rng.standard_normal(o.f[n].shape).astype(np.float32)
        o.f[n][...] = a
        g.f[n].copy_(torch.from_numpy(a))
    if which == "velocity":
        g.velocity(0)
        o.velocity(0)
    else:
        d = g.dims
        g.stress_region(0, *d.i_range, *d.j_range)
        o.stress_region(0, *d.i_range, *d.j_range)
    _lib.check()
    for n in FIELDS:
        a, b = g.field(n), o.f[n]
        bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
        assert len(bad) == 0 or np.array_equal(a[tuple(bad.T)], b[tuple(bad.T)]), \
            (n, len(bad), tuple(bad[0]), a[tuple(bad[0])], b[tuple(bad[0])])


@pytest.mark.gpu
@pytest.mark.parametrize("split", [1, 0])
def test_gpu_topography_velocity_forms_bit_exact(gpu, oracle_lib, split):
    """The curvilinear velocity update as one launch per component (the
    default, awp_set_tuning("topo_vel_split", 1)) and as one launch for all
    three: 40 steps bit-exact against the oracle."""
    torch, _lib = gpu
    from awp_amd.solver import FIELDS, Solver
    from oracle.oracle import OracleSolver
    lib = _lib.load()
    w = _workload()
    med = w.medium()
    topo = w.params["topography"]
    assert lib.awp_set_tuning(b"topo_vel_split", split) == 0
    try:
        g = Solver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
        g.run(w.nt, check_every=10)
        torch.cuda.synchronize()
    finally:
        assert lib.awp_set_tuning(b"topo_vel_split", 1) == 0
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    o.run(w.nt)
    assert np.abs(o.traces()).max() > 0
    assert np.array_equal(g.traces(), o.traces())
    for n in FIELDS:
        assert np.array_equal(g.field(n), o.f[n]), n


@pytest.mark.gpu
@pytest.mark.parametrize("split", [1, 2, 0])
def test_gpu_topography_rate_groups_bit_exact(gpu, oracle_lib, split):
    """The strain-rate pass of the two-pass stress update as two launches
    (normal rates, then shear sums: awp_set_tuning("topo_str_split", 1)), as
    three (normal rates, the xy sum, the xz and yz sums: 2, the default) and
    as one: 40 steps bit-exact against the oracle, FP32 and BF16 storage."""
    torch, _lib = gpu
    from awp_amd.solver import FIELDS, Solver, Solver16
    from awp_amd.storage import FIELDS16
    from oracle.oracle import OracleSolver, OracleSolver16
    lib = _lib.load()
    w = _workload()
    med = w.medium()
    topo = w.params["topography"]
    default = 2
    assert lib.awp_set_tuning(b"topo_str_split", split) == 0
    try:
        g = Solver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo, texture="rule")
        g.run(w.nt, check_every=10)
        g16 = Solver16(med, w.sources, w.receivers, nt_record=w.nt, storage="bf16",
                       topography=topo)
        g16.run(w.nt, check_every=10)
        torch.cuda.synchronize()
    finally:
        assert lib.awp_set_tuning(b"topo_str_split", default) == 0
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    o.run(w.nt)
    assert np.abs(o.traces()).max() > 0
    assert np.array_equal(g.traces(), o.traces())
    for n in FIELDS:
        assert np.array_equal(g.field(n), o.f[n]), n
    o16 = OracleSolver16(med, w.sources, w.receivers, nt_record=w.nt, storage="bf16",
                         topography=topo)
    o16.run(w.nt)
    assert g16.scales == o16.scales and g16.rescales == o16.rescales
    assert np.array_equal(g16.traces(), o16.traces())
    for n in FIELDS16:
        assert np.array_equal(g16.raw(n), o16.levels[0].q[n]), n


@pytest.mark.gpu
@pytest.mark.parametrize("two_pass,texture", [(True, "arrays"), (False, "arrays"),
                                              (True, "rule"), (True, "table")],
                         ids=["two_pass", "one_pass", "two_pass_rule", "two_pass_table"])
def test_gpu_topography_run_bit_exact(gpu, oracle_lib, two_pass, texture):
    """40 steps of the pmcl3d order with the curvilinear updates on grid 0
    (source, free-surface pass, receivers): wavefield and traces, with the
    stress update in one kernel and in two passes (strain rates, then the
    material and memory-variable update), the second pass reading the
    per-cell relaxation texture or the 8-entry table with ww read or
    evaluated from init_texture's parity rule
    (awp_topo_stress_interior_tex_H)."""
    torch, _lib = gpu
    from awp_amd.solver import FIELDS, Solver
    from oracle.oracle import OracleSolver
    w = _workload()
    med = w.medium()
    topo = w.params["topography"]
    g = Solver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo, texture=texture)
    assert g.levels[0].texmode == texture
    if not two_pass:
        g.topo.rates = None
    g.run(w.nt, check_every=10)
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    o.run(w.nt)
    assert np.array_equal(g.traces(), o.traces())
    assert np.abs(o.traces()).max() > 0
    for n in FIELDS:
        assert np.array_equal(g.field(n), o.f[n]), n


@pytest.mark.gpu
def test_gpu_topography_over_dm_bit_exact(gpu, oracle_lib):
    """Curvilinear grid 0 (hills) over a 3x coarser Cartesian grid
    (pmcl3d.c:1661-1694, 1843): the curvilinear updates on grid 0, dvelc2_H
    and dstrqc2_H on its align+2 level, intp3d_H and swap_H between the two.
    120 steps of NVE = 1 with the source in the coarse grid: surface traces
    and both grids' wavefields bit-exact against the FP32 oracle."""
    torch, _lib = gpu
    from awp_amd import configs
    from awp_amd.solver import FIELDS, Solver
    from oracle.oracle import OracleSolver
    nt = 120
    w = configs.dm_plastic(nx=16, ny=16, nz=12, nz_fine=32, DH=300.0, DT=0.005, nt=nt,
                           nsub=4, ND=4, M0=1.0e16, f0=5.0, src_depth=3, src_grid=1)
    media = w.medium()
    d0 = media[0].dims
    topo = Topography.gaussian_hills(d0.nxt, d0.nyt, media[0].DH, n=3, seed=5, hmax=400.0,
                                     wmin=600.0, wmax=1500.0)
    g = Solver(media, w.sources, w.receivers, nt_record=nt, topography=topo)
    g.run(nt, check_every=20)
    o = OracleSolver(media, w.sources, w.receivers, nt_record=nt, dtype=np.float32,
                     topography=topo)
    o.run(nt)
    assert np.abs(o.traces()).max() > 0
    assert np.array_equal(g.traces(), o.traces())
    for p in range(2):
        for n in FIELDS:
            assert np.array_equal(g.field(n, p), o.levels[p].f[n]), (p, n)


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["bf16", "fp16"])
def test_gpu_topography_16bit_bit_exact(gpu, oracle_lib, storage):
    from awp_amd.solver import Solver16
    from awp_amd.storage import FIELDS16
    from oracle.oracle import OracleSolver16
    w = _workload(nt=30)
    med = w.medium()
    topo = w.params["topography"]
    g = Solver16(med, w.sources, w.receivers, nt_record=w.nt, storage=storage, topography=topo)
    g.run(w.nt, check_every=10)
    o = OracleSolver16(med, w.sources, w.receivers, nt_record=w.nt, storage=storage,
                       topography=topo)
    o.run(w.nt)
    assert g.scales == o.scales and g.rescales == o.rescales
    assert np.array_equal(g.traces(), o.traces())
    for n in FIELDS16:
        assert np.array_equal(g.raw(n), o.levels[0].q[n]), n


@pytest.mark.gpu
@pytest.mark.parametrize("storage,tol", [("f32", 5e-5), ("bf16", 0.1), ("fp16", 2e-2)])
def test_gpu_topography_workload_against_fp64(gpu, storage, tol):
    """The full topography workload (384x384x192, six hills, 1500 steps) on
    the GPU against the FP64 oracle's receiver traces
    (golden/topography_384x384x192_fp64_traces.npz, written by
    golden/fp64_topography.py).  Stated tolerances on the relative L2 error
    over all receivers, components and samples: FP32 5e-5 (measured 1.4e-5),
    BF16 0.1 (7.2e-2), FP16 2e-2 (8.8e-3)."""
    from pathlib import Path
    from awp_amd.solver import Solver, Solver16
    torch, _lib = gpu
    ref = np.load(Path(__file__).parent / "golden" / "topography_384x384x192_fp64_traces.npz")
    ref = ref["traces"]
    nt = ref.shape[0] - 1
    w = configs.topography_hills(nx=384, ny=384, nz=192, nt=nt)
    med = w.medium()
    topo = w.params["topography"]
    if storage == "f32":
        s = Solver(med, w.sources, w.receivers, nt_record=nt + 1, topography=topo)
    else:
        s = Solver16(med, w.sources, w.receivers, nt_record=nt + 1, storage=storage,
                     topography=topo)
    del med
    s.run(nt)
    got = s.traces().astype(np.float64)
    del s
    torch.cuda.empty_cache()
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"{storage}: receiver rel-L2 vs FP64 oracle {err:.3e}")
    assert err < tol


@pytest.mark.gpu
@pytest.mark.parametrize("nseg", [8, 2, 1])
def test_gpu_topography_stress_chunks_bit_exact(gpu, oracle_lib, nseg):
    """The curvilinear stress kernel's k chunks on a 200-level column:
    awp_set_tuning("topo_nseg") 8 (the default: one chunk of four waves
    sharing an LDS column, 56 lanes below level 0), 2 (an interior chunk of
    two waves and the top chunk, overlapping by six levels) and 1 (one-wave
    segments of 58 outputs).  24 steps with the two-pass (three rate groups)
    and one-pass updates, bit-exact against the oracle."""
    torch, _lib = gpu
    from awp_amd.solver import FIELDS, Solver
    from oracle.oracle import OracleSolver
    lib = _lib.load()
    w = configs.topography_hills(nx=24, ny=20, nz=200, nt=24, DH=100.0, DT=0.005, nhills=3,
                                 hmax=500.0, wmin=600.0, wmax=1500.0, nrec=6)
    med = w.medium()
    topo = w.params["topography"]
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    o.run(w.nt)
    assert lib.awp_set_tuning(b"topo_nseg", nseg) == 0
    try:
        for two_pass in (True, False):
            g = Solver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
            if not two_pass:
                g.topo.rates = None
            g.run(w.nt, check_every=8)
            torch.cuda.synchronize()
            assert np.array_equal(g.traces(), o.traces()), two_pass
            for n in FIELDS:
                assert np.array_equal(g.field(n), o.f[n]), (two_pass, n)
    finally:
        assert lib.awp_set_tuning(b"topo_nseg", 8) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("nseg,wsync", [(1, 1), (8, 1), (1, 0)])
def test_gpu_topography_wave_sync_bit_exact(gpu, oracle_lib, nseg, wsync):
    """The curvilinear kernels with wave-scope LDS ordering where a wave
    reads only its own LDS (awp_set_tuning("topo_wsync", 1), the default:
    the velocity kernel, and the strain-rate kernels in one-wave segments;
    chunks of several waves keep the workgroup barrier) and with the
    workgroup barrier everywhere (0): 24 steps on a 200-level column,
    bit-exact against the oracle."""
    torch, _lib = gpu
    from awp_amd.solver import FIELDS, Solver
    from oracle.oracle import OracleSolver
    lib = _lib.load()
    w = configs.topography_hills(nx=24, ny=20, nz=200, nt=24, DH=100.0, DT=0.005, nhills=3,
                                 hmax=500.0, wmin=600.0, wmax=1500.0, nrec=6)
    med = w.medium()
    topo = w.params["topography"]
    o = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
    o.run(w.nt)
    assert lib.awp_set_tuning(b"topo_wsync", wsync) == 0
    assert lib.awp_set_tuning(b"topo_nseg", nseg) == 0
    try:
        g = Solver(med, w.sources, w.receivers, nt_record=w.nt, topography=topo)
        g.run(w.nt, check_every=8)
        torch.cuda.synchronize()
    finally:
        assert lib.awp_set_tuning(b"topo_wsync", 1) == 0
        assert lib.awp_set_tuning(b"topo_nseg", 8) == 0
    assert np.array_equal(g.traces(), o.traces())
    for n in FIELDS:
        assert np.array_equal(g.field(n), o.f[n]), n
