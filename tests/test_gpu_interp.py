"""Lagrange-interpolated receivers and sources on the GPU (cuinterp_interp_H,
cusource_add_{cartesian,curvilinear,force}_H through the C ABI) against the
oracle, bit for bit: receivers over degrees 0-5 with queries near the grid
edges, outside it (clamped) and on nodes; sources with overlapping and
coincident stencils, whose terms every run adds in query order; the
launchers' bounds checks."""
import numpy as np
import pytest

from awp_amd.grid import ALIGN, NGSL, Dims
from awp_amd.interpolation import Interpolation, grid_coordinates, metric_grid

pytestmark = pytest.mark.gpu

H = np.float32(0.1)


@pytest.fixture(scope="module")
def env(oracle_lib):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awp_amd import _lib
    from oracle import interp
    _lib.load()
    return torch, _lib, interp


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


def _grid(dims):
    grid = metric_grid(dims)
    x = grid_coordinates(grid.size_x, H, padding=NGSL)
    y = grid_coordinates(grid.size_y, H, shift=1, padding=NGSL)
    z = grid_coordinates(grid.size_z, H, shift=1)
    return grid, x, y, z


def _receiver_queries(rng, x, y, z, nq):
    qx = rng.uniform(x[0] - 0.3, x[-1] + 0.3, nq)
    qy = rng.uniform(y[0], y[-1], nq)
    qz = rng.uniform(z[0] - 0.1, z[-1], nq)
    qx[:20] = x[rng.integers(0, len(x), 20)]
    qy[20:40] = 0.5 * (y[3:23] + y[4:24])
    qz[40:60] = z[-1]
    return qx, qy, qz


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 5])
def test_gpu_receivers_match_oracle(env, deg):
    torch, _lib, oi = env
    rng = np.random.default_rng(100 + deg)
    dims = Dims(40, 36, 48)
    grid, x, y, z = _grid(dims)
    nq = 3000
    qx, qy, qz = _receiver_queries(rng, x, y, z, nq)
    ridx = rng.permutation(nq + 5)[:nq].astype(np.int32)   # scattered output slots
    field = rng.standard_normal(dims.size).astype(np.float32)
    print(f"receivers deg={deg} nq={nq} dims={dims}", flush=True)
    I = Interpolation(x, y, z, grid, qx, qy, qz, ridx, degree=deg)
    coef = I.coefficients()
    want_coef = oi.lagrange_h(x, y, z, qx, qy, qz, deg)[1:]
    for g, w in zip(coef, want_coef):
        assert np.array_equal(np.asarray(g).view(np.int32), np.asarray(w).view(np.int32))
    out = torch.full((nq + 5,), -7.0, device="cuda")
    I.interp(out, torch.from_numpy(field).cuda())
    torch.cuda.synchronize()
    want = np.full(nq + 5, -7.0, np.float32)
    oi.interp(want, field, coef, ridx, grid)
    assert np.array_equal(bits(out.cpu().numpy()), bits(want))
    I.close()


def _jacobians(rng, dims):
    f = rng.uniform(0.5, 1.5, (dims.nxt + 20) * (dims.nyt + 84)).astype(np.float32)
    dg = rng.uniform(0.5, 1.5, dims.nzt + 2 * ALIGN).astype(np.float32)
    rho = rng.uniform(1.5, 3.0, dims.size).astype(np.float32)
    return f, dg, rho


@pytest.mark.parametrize("deg", [1, 3])
@pytest.mark.parametrize("kind", ["cartesian", "curvilinear", "force"])
def test_gpu_sources_match_oracle(env, kind, deg):
    """Clustered sources (most stencils overlap, some coincide): the GPU adds
    each grid point's terms in query order, as the oracle's sequential loop."""
    torch, _lib, oi = env
    rng = np.random.default_rng(200 + deg + 10 * len(kind))
    dims = Dims(32, 28, 40)
    grid, x, y, z = _grid(dims)
    nq = 600
    qx = rng.normal(1.2, 0.15, nq).clip(0.0, 2.9)
    qy = rng.normal(1.0, 0.15, nq).clip(0.0, 2.5)
    qz = rng.normal(2.0, 0.2, nq).clip(0.0, 3.8)
    qx[100:110], qy[100:110], qz[100:110] = qx[0], qy[0], qz[0]   # coincident
    lidx = rng.integers(0, 50, nq).astype(np.int32)               # shared inputs
    vals = rng.standard_normal(50).astype(np.float32)
    base = rng.standard_normal(dims.size).astype(np.float32)
    f, dg, rho = _jacobians(rng, dims)
    print(f"sources kind={kind} deg={deg} nq={nq}", flush=True)
    I = Interpolation(x, y, z, grid, qx, qy, qz, lidx, degree=deg)
    assert I.num_cells < nq * (deg + 1) ** 3                      # stencils overlap
    coef = I.coefficients()
    dt, qw = 0.004, 0.25
    kw = {}
    if kind != "cartesian":
        kw = dict(f=f, dg=dg)
    if kind == "force":
        kw.update(d1=rho, quad_weight=qw)
    want = base.copy()
    oi.source_add(kind, want, vals, coef, lidx, grid, H, dt, ny=dims.nyt, nz=dims.nzt, **kw)

    def run():
        out = torch.from_numpy(base).cuda()
        v = torch.from_numpy(vals).cuda()
        if kind == "cartesian":
            I.add_cartesian(out, v, H, dt)
        elif kind == "curvilinear":
            I.add_curvilinear(out, v, H, dt, torch.from_numpy(f).cuda(), dims.nyt,
                              torch.from_numpy(dg).cuda())
        else:
            I.add_force(out, v, torch.from_numpy(rho).cuda(), H, dt, qw,
                        torch.from_numpy(f).cuda(), dims.nxt, dims.nyt, dims.nzt,
                        torch.from_numpy(dg).cuda())
        torch.cuda.synchronize()
        return out.cpu().numpy()

    first, second = run(), run()
    assert np.array_equal(bits(first), bits(want))
    assert np.array_equal(bits(second), bits(first))
    I.close()


def test_gpu_many_receivers_reproduce_a_polynomial(env):
    """20000 receivers, degree 3, on a 160 x 160 x 96 field holding a cubic:
    the interpolant reproduces it, and equals the oracle bit for bit."""
    torch, _lib, oi = env
    rng = np.random.default_rng(7)
    dims = Dims(160, 160, 96)
    grid, x, y, z = _grid(dims)
    nq = 20000
    qx = rng.uniform(x[0], x[-1], nq)
    qy = rng.uniform(y[0], y[-1], nq)
    qz = rng.uniform(z[0], z[-1], nq)
    a = np.zeros(dims.size, np.float32)
    X, Y, Z = (v.astype(np.float64) for v in (x, y, z))
    ii, jj, kk = np.meshgrid(np.arange(len(x)), np.arange(len(y)), np.arange(len(z)),
                             indexing="ij")
    pos = grid.off_z + kk + (grid.off_y + jj) * grid.line + (grid.off_x + ii) * grid.slice
    fn = lambda X, Y, Z: 0.01 * X ** 3 - 0.2 * Y ** 2 * Z + 0.05 * Z ** 3 + 1.0  # noqa: E731
    a[pos.ravel()] = fn(X[ii], Y[jj], Z[kk]).ravel()
    print(f"receivers nq={nq} dims={dims}", flush=True)
    I = Interpolation(x, y, z, grid, qx, qy, qz, degree=3)
    out = torch.zeros(nq, device="cuda")
    I.interp(out, torch.from_numpy(a).cuda())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = np.zeros(nq, np.float32)
    oi.interp(want, a, I.coefficients(), np.arange(nq), grid)
    assert np.array_equal(bits(got), bits(want))
    exact = fn(*(np.float32(v).astype(np.float64) for v in (qx, qy, qz)))
    assert np.max(np.abs(got - exact)) < 2e-4 * max(1.0, float(np.abs(a).max()))
    I.close()


def test_gpu_force_rejects_stencils_outside_the_density_array(env):
    """The force kernel reads density at the stencil index + ngsl + 2 (the
    reference's layout); a source whose stencil would leave the density array
    is refused before any launch and the field is left alone."""
    torch, _lib, oi = env
    rng = np.random.default_rng(9)
    dims = Dims(24, 24, 32)
    grid, x, y, z = _grid(dims)
    f, dg, rho = _jacobians(rng, dims)
    I = Interpolation(x, y, z, grid, [x[-3]], [1.0], [1.0], degree=3)
    out = torch.zeros(dims.size, device="cuda")
    with pytest.raises(_lib.AwpError, match="leave the metric/density arrays"):
        I.add_force(out, torch.ones(1, device="cuda"), torch.from_numpy(rho).cuda(), H, 0.01,
                    1.0, torch.from_numpy(f).cuda(), dims.nxt, dims.nyt, dims.nzt,
                    torch.from_numpy(dg).cuda())
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
    I.close()


def test_gpu_htod_rejects_indices_outside_the_field(env):
    """cuinterp_htod checks every stencil point against the field before it
    uploads a map (host indices edited after init)."""
    import ctypes as C
    torch, _lib, oi = env
    dims = Dims(16, 16, 16)
    grid, x, y, z = _grid(dims)
    I = Interpolation(x, y, z, grid, [0.5], [0.5], [0.5], degree=1)
    I.I.ix[0] = grid.size_x + 40
    assert _lib.load().cuinterp_htod(C.byref(I.I)) == 400
    I.close()


def test_gpu_receiver_file_drives_sampling(env, tmp_path, monkeypatch):
    """The reference's receiver fixture (tests/fixtures/receiver.txt: four
    surface-level receivers, degree 3, buffers of 2 x 3 records) read by the
    v2 reader, mapped to query coordinates, sampled every step into the
    device buffer and flushed to the host buffer: every record equals the
    oracle's interpolation of that step's field."""
    import shutil
    from pathlib import Path
    from awp_amd import inputs
    torch, _lib, oi = env
    fix = Path(__file__).resolve().parent / "golden" / "fixtures" / "receiver.txt"
    shutil.copy(fix, tmp_path / "receiver.txt")
    monkeypatch.chdir(tmp_path)
    inp = inputs.read_input("receiver.txt")
    dims = Dims(16, 16, 24)
    grid, x, y, z = _grid(dims)
    qx, qy, qz = inputs.cartesian_queries(inp, H, dims.nzt)
    I = Interpolation(x, y, z, grid, qx, qy, qz, degree=inp.degree)
    buf = inputs.StepBuffer(inp.length, inp.gpu_buffer_size, inp.cpu_buffer_size, inp.stride)
    rng = np.random.default_rng(4)
    coef = I.coefficients()
    want = np.zeros_like(buf.h_buffer)
    nsteps = inp.gpu_buffer_size * inp.cpu_buffer_size
    for step in range(nsteps):
        field = rng.standard_normal(dims.size).astype(np.float32)
        if buf.is_device_ready(step):
            I.interp(buf.device_slot(step), torch.from_numpy(field).cuda())
            rec = np.zeros(inp.length, np.float32)
            oi.interp(rec, field, coef, np.arange(inp.length), grid)
            o = buf.host_offset(step) + buf.device_offset(step)
            want[o:o + inp.length] = rec
        torch.cuda.synchronize()
        buf.copy_to_host(step)
    assert buf.is_host_full(nsteps - 1)
    assert np.array_equal(bits(buf.h_buffer), bits(want))
    I.close()


def test_gpu_receiver_files(env, tmp_path, monkeypatch):
    """receiver_write's file output for the reference's receiver fixture: six
    samples per host buffer (2 x 3), files recv_0006 and recv_0012, each
    receiver's six samples contiguous, equal to the oracle's interpolation of
    each step's field."""
    import shutil
    from pathlib import Path
    from awp_amd import inputs
    torch, _lib, oi = env
    fix = Path(__file__).resolve().parent / "golden" / "fixtures" / "receiver.txt"
    shutil.copy(fix, tmp_path / "receiver.txt")
    monkeypatch.chdir(tmp_path)
    inp = inputs.read_input("receiver.txt")
    dims = Dims(16, 16, 24)
    grid, x, y, z = _grid(dims)
    I = Interpolation(x, y, z, grid, *inputs.cartesian_queries(inp, H, dims.nzt),
                      degree=inp.degree)
    w = inputs.ReceiverWriter(I, inp, "recv", inp.steps, directory=str(tmp_path))
    rng = np.random.default_rng(6)
    coef = I.coefficients()
    want = []
    for step in range(12):
        field = rng.standard_normal(dims.size).astype(np.float32)
        w.write(step, torch.from_numpy(field).cuda())
        rec = np.zeros(inp.length, np.float32)
        oi.interp(rec, field, coef, np.arange(inp.length), grid)
        want.append(rec)
    torch.cuda.synchronize()
    assert [Path(f).name for f in w.files] == ["recv_0006", "recv_0012"]
    for n, f in enumerate(w.files):
        got = np.fromfile(f, dtype=np.float32).reshape(inp.length, 6)
        exp = np.stack(want[6 * n:6 * n + 6], 1)
        assert np.array_equal(bits(got), bits(exp)), f
    I.close()


def test_gpu_receivers_at_the_benchmark_size(env):
    """200,000 degree-3 receivers on the 512 x 512 x 256 layout of the
    benchmark, on a field holding a cubic: bit for bit against the oracle,
    and the cubic reproduced to float32 rounding."""
    torch, _lib, oi = env
    rng = np.random.default_rng(21)
    dims = Dims(512, 512, 256)
    grid, x, y, z = _grid(dims)
    nq = 200000
    qx = rng.uniform(x[0], x[-1], nq)
    qy = rng.uniform(y[0], y[-1], nq)
    qz = rng.uniform(z[0], z[-1], nq)
    print(f"receivers nq={nq} dims={dims}", flush=True)
    # the field on the device: c(x, y, z) at every coordinate point, zero elsewhere
    X = torch.from_numpy(x.astype(np.float64)).cuda()
    Y = torch.from_numpy(y.astype(np.float64)).cuda()
    Z = torch.from_numpy(z.astype(np.float64)).cuda()
    cub = (0.01 * X[:, None, None] ** 3 - 0.2 * Y[None, :, None] ** 2 * Z[None, None, :]
           + 0.05 * Z[None, None, :] ** 3 + 1.0).float()
    a = torch.zeros(dims.shape, device="cuda")
    a[2:2 + len(x), 2:2 + len(y), ALIGN:ALIGN + len(z)] = cub
    del cub
    a = a.reshape(-1)
    I = Interpolation(x, y, z, grid, qx, qy, qz, degree=3)
    out = torch.zeros(nq, device="cuda")
    I.interp(out, a)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    host = a.cpu().numpy()
    want = np.zeros(nq, np.float32)
    oi.interp(want, host, I.coefficients(), np.arange(nq), grid)
    assert np.array_equal(bits(got), bits(want))
    q = [np.float32(v).astype(np.float64) for v in (qx, qy, qz)]
    exact = 0.01 * q[0] ** 3 - 0.2 * q[1] ** 2 * q[2] + 0.05 * q[2] ** 3 + 1.0
    assert np.max(np.abs(got - exact)) < 2e-6 * float(np.abs(host).max()) * 64
    I.close()


def test_gpu_sources_at_the_benchmark_size(env):
    """4,000 degree-3 point sources one cell apart on a fault plane of the
    512 x 512 x 256 layout (256,000 terms on about 79,000 grid points): bit
    for bit against the oracle's sequential loop, and the injected total is
    -dt/h^3 times the sum of the source values (partition of unity)."""
    torch, _lib, oi = env
    rng = np.random.default_rng(22)
    dims = Dims(512, 512, 256)
    grid, x, y, z = _grid(dims)
    nq = 4000
    u, v = np.meshgrid(np.linspace(0.3, 0.7, 64), np.linspace(0.4, 0.7, 63), indexing="ij")
    qx = np.full(nq, 0.5 * 512 * H) + rng.uniform(-0.5, 0.5, nq) * H
    qy = u.ravel()[:nq] * 512 * H
    qz = v.ravel()[:nq] * 252 * H
    vals = rng.standard_normal(nq).astype(np.float32)
    print(f"sources nq={nq} dims={dims}", flush=True)
    I = Interpolation(x, y, z, grid, qx, qy, qz, degree=3)
    assert I.num_cells < nq * 64
    dt = 0.004
    out = torch.zeros(dims.size, device="cuda")
    I.add_cartesian(out, torch.from_numpy(vals).cuda(), H, dt)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = np.zeros(dims.size, np.float32)
    oi.source_add("cartesian", want, vals, I.coefficients(), np.arange(nq), grid, H, dt)
    assert np.array_equal(bits(got), bits(want))
    dth = float(np.float32(dt) / (H * H * H))
    total = -dth * float(vals.astype(np.float64).sum())
    assert abs(float(got.astype(np.float64).sum()) - total) < 1e-4 * dth * np.abs(vals).sum()
    I.close()
