"""Input-file point sets on the GPU (awp_amd.pointsets.PointIO) against the
oracle driving the same input files (oracle/pointio.py).

Two runs:
- a discontinuous mesh (48x48x18 at h = 100 m over 16x16x12 at h = 300 m),
  moment-tensor sources read from files in both blocks, velocity receivers
  in both blocks and at the surface (two writes per file), SGT output;
- a curvilinear topography grid (six hills), sources with the curvilinear
  Jacobians, surface forces, velocity receivers and SGT output.
Every wavefield array after the run is bit for bit the oracle's (the file
sources inject the same values at the same cells), and every receiver, SGT
and material file holds the oracle's samples of its fields at the steps
the file covers, in the reference's file layout.
"""
import numpy as np
import pytest

from awp_amd import configs
from awp_amd.inputs import Input, Version, read_input, receiver_step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _write_input(name, file, pts, steps, stride, degree, gpu, cpu, num_writes):
    inp = Input(version=Version(2, 1, 0), file=file, length=len(pts), steps=steps, stride=stride,
                degree=degree, gpu_buffer_size=gpu, cpu_buffer_size=cpu,
                num_writes=num_writes)
    pts = np.asarray(pts, dtype=np.float64)
    inp.type = pts[:, 0].astype(np.int32)
    inp.x, inp.y, inp.z = (pts[:, k].astype(np.float32) for k in (1, 2, 3))
    assert inp.write(name) == 0
    return read_input(name)


def _write_series(inp, suffixes, rng, scale):
    for s in suffixes:
        t = np.arange(inp.steps)
        amp = rng.uniform(0.5, 1.5, (inp.length, 1))
        data = (scale * amp * np.sin(2 * np.pi * (t[None, :] + 3.0) / 17.0)
                * np.exp(-((t[None, :] - 10.0) / 8.0) ** 2)).astype(np.float32)
        data.tofile(f"{inp.file}_{s}")


def _check_files(files, base, inp, recorded, local_index, nt):
    """Each file of base: receiver r's num_writes blocks of gpu*cpu samples,
    the samples of host-full k at steps (k*rows + row)*stride; blocks of
    host-fulls the run did not complete hold zeros."""
    rows = inp.gpu_buffer_size * inp.cpu_buffer_size
    unit = rows * inp.num_writes * inp.stride
    mine = sorted(f for f in files if f.split("/")[-1].startswith(base + "_"))
    assert mine, base
    for f in mine:
        last = int(f.rsplit("_", 1)[1])
        q = last // unit - 1
        got = np.fromfile(f, dtype=np.float32).reshape(inp.length, inp.num_writes, rows)
        for w in range(inp.num_writes):
            k = q * inp.num_writes + w
            if ((k + 1) * rows - 1) * inp.stride > nt:
                assert not got[:, w, :].any()
                continue
            for r in range(rows):
                step = (k * rows + r) * inp.stride
                assert receiver_step(inp.gpu_buffer_size, inp.cpu_buffer_size, inp.num_writes,
                                     inp.stride, step) == last
                want = recorded[step]
                for i, g in enumerate(local_index):
                    assert got[g, w, r] == want[i], (f, g, w, r)
    return len(mine)


def _compare_fields(g, o, names):
    for p in range(len(o.levels)):
        for n in names:
            a = g.field(n, p)
            assert np.array_equal(a, o.levels[p].f[n]), (p, n, int((a != o.levels[p].f[n]).sum()))


FIELDS = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r4")


def test_pointsets_discontinuous_mesh(torch_gpu, oracle_lib, tmp_path, monkeypatch):
    from awp_amd.pointsets import grids_init
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    from oracle.pointio import OraclePointIO
    monkeypatch.chdir(tmp_path)
    nt = 36
    print(f"case point sets on a 2-block mesh, {nt} steps", flush=True)
    rng = np.random.default_rng(31)
    src = _write_input("src.txt", "src", [[0, 2150.0, 2050.0, -600.0],
                                          [0, 2350.0, 2600.0, -2500.0],
                                          [0, 2600.0, 2300.0, -900.0]],
                       steps=48, stride=1, degree=3, gpu=2, cpu=3, num_writes=1)
    _write_series(src, ("xx", "yy", "zz", "xy", "xz", "yz"), rng, 1.0e13)
    rec = _write_input("rec.txt", "rec", [[1, 1500.0, 1700.0, 0.0], [0, 2000.0, 3000.0, -400.0],
                                          [0, 3100.0, 1900.0, -2100.0],
                                          [0, 2400.0, 2500.0, -1300.0]],
                       steps=nt, stride=2, degree=3, gpu=2, cpu=3, num_writes=2)
    sgt = _write_input("sgt.txt", "sgt", [[0, 2250.0, 2250.0, -750.0],
                                          [0, 2700.0, 2100.0, -2400.0]],
                       steps=nt, stride=1, degree=1, gpu=2, cpu=2, num_writes=1)
    w = configs.dm_plastic(nx=16, ny=16, nz=12, nz_fine=18, DH=300.0, DT=0.005, nt=nt, nsub=8,
                           ND=4, M0=1.0e16, f0=5.0, src_depth=6, src_grid=1)
    media = w.medium()
    opts = dict(sources="src.txt", receivers="rec.txt", sgt="sgt.txt")
    g = Solver(media, None, None, pointio=dict(nt=nt, directory=str(tmp_path), **opts))
    g.run(nt, check_every=12)
    stack = [grids_init(m.dims.nxt, m.dims.nyt, m.dims.nzt, 0, m.DH) for m in media]
    o = OracleSolver(media, None, None, dtype=np.float32)
    o.pointio = OraclePointIO(stack, **opts)
    o.pointio_steps = nt
    o.run(nt)
    assert np.abs(o.levels[0].f["xx"]).max() > 0 and np.abs(o.levels[1].f["xx"]).max() > 0
    _compare_fields(g, o, FIELDS)
    P = g.pointio
    assert P.sources.m[0].lengths == [2, 1]
    files = P.receivers.files + P.sgt.files
    nfiles = 0
    for name, rset in zip(("x", "y", "z"), P.receivers.r):
        nfiles += _check_files(files, f"rec_{name}", rec, o.pointio.recorded[name],
                               rset.indices, nt)
    for (name, _), rset in zip(P.sgt.COMPONENTS, P.sgt.g):
        nfiles += _check_files(files, f"sgt_{name}", sgt, o.pointio.recorded[name], rset.indices, nt)
    for m in ("d1", "lami", "mui"):
        got = np.fromfile(str(tmp_path / f"sgt_{m}"), dtype=np.float32)
        want = np.zeros(sgt.length, np.float32)
        want[P.sgt.mat.indices] = o.pointio.material[m]
        assert np.array_equal(got, want), m
    print(f"point sets: {nfiles} receiver/SGT files checked", flush=True)
    g.pointio.close()


def test_pointsets_topography(torch_gpu, oracle_lib, tmp_path, monkeypatch):
    from awp_amd.pointsets import grids_init
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    from oracle.pointio import OraclePointIO
    monkeypatch.chdir(tmp_path)
    nt = 30
    print(f"case point sets on a topography grid, {nt} steps", flush=True)
    rng = np.random.default_rng(32)
    w = configs.topography_hills(nx=40, ny=36, nz=32, nt=nt)
    topo = w.params["topography"]
    h = w.DH
    L = 20 * h
    src = _write_input("src.txt", "src", [[0, 0.45 * 40 * h, 0.5 * 36 * h, -0.3 * L],
                                          [0, 0.55 * 40 * h, 0.45 * 36 * h, -0.5 * L]],
                       steps=36, stride=1, degree=3, gpu=2, cpu=3, num_writes=1)
    _write_series(src, ("xx", "yy", "zz", "xy", "xz", "yz"), rng, 1.0e14)
    frc = _write_input("frc.txt", "frc", [[1, 0.5 * 40 * h, 0.5 * 36 * h, 0.0],
                                          [1, 0.4 * 40 * h, 0.6 * 36 * h, 0.0]],
                       steps=36, stride=1, degree=3, gpu=3, cpu=2, num_writes=1)
    _write_series(frc, ("fx", "fy", "fz"), rng, 1.0e9)
    rec = _write_input("rec.txt", "rec", [[1, 0.3 * 40 * h, 0.3 * 36 * h, 0.0],
                                          [0, 0.6 * 40 * h, 0.55 * 36 * h, -0.2 * L],
                                          [1, 0.7 * 40 * h, 0.4 * 36 * h, 0.0]],
                       steps=36, stride=1, degree=3, gpu=2, cpu=2, num_writes=3)
    sgt = _write_input("sgt.txt", "sgt", [[0, 0.5 * 40 * h, 0.5 * 36 * h, -0.25 * L]],
                       steps=36, stride=2, degree=3, gpu=1, cpu=3, num_writes=1)
    med = w.medium()
    opts = dict(sources="src.txt", receivers="rec.txt", forces="frc.txt", sgt="sgt.txt")
    g = Solver(med, None, None, topography=topo,
               pointio=dict(nt=nt, directory=str(tmp_path), **opts))
    g.run(nt, check_every=10)
    o = OracleSolver(med, None, None, dtype=np.float32, topography=topo)
    stack = [grids_init(med.dims.nxt, med.dims.nyt, med.dims.nzt, 1, med.DH)]
    o.pointio = OraclePointIO(stack, f_host=o.topo.mf["f"], **opts)
    o.pointio_steps = nt
    o.run(nt)
    assert np.abs(o.f["u1"]).max() > 0
    _compare_fields(g, o, FIELDS)
    P = g.pointio
    files = P.receivers.files + P.sgt.files
    for name, rset in zip(("x", "y", "z"), P.receivers.r):
        _check_files(files, f"rec_{name}", rec, o.pointio.recorded[name], rset.indices, nt)
    for (name, _), rset in zip(P.sgt.COMPONENTS, P.sgt.g):
        _check_files(files, f"sgt_{name}", sgt, o.pointio.recorded[name], rset.indices, nt)
    g.pointio.close()
