"""Kinematic finite-fault sources (addkinsrc_H with SetDeviceFilterParameters)
and prescribed-velocity sources (frcvel_H) on the GPU against the oracle.

frcvel is plain single-precision arithmetic: bit for bit.  addkinsrc calls
sinf, cosf, expf and powf, whose device and glibc versions differ in the last
bits, so its stresses, moments and filter states are compared to a relative
1e-5 of each array's largest value (a few ulp carried through 80 steps)."""
import ctypes as C

import numpy as np
import pytest

from awp_amd.grid import Dims
from tests._fields import first_mismatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(oracle_lib):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awp_amd import _lib
    from oracle import kin
    return torch, _lib, _lib.load(), kin


def _set_grid(lib, dims, DH, DT):
    F, I = C.c_float, C.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 1)(DH), F(DT), (I * 1)(dims.nxt), (I * 1)(dims.nyt),
                            (I * 1)(dims.nzt), 1, 0.0, 0.0, z9, z9)


def _cells(rng, dims, n):
    """n distinct source cells (1-based x, y; z counted from the surface)."""
    flat = rng.choice(dims.nxt * dims.nyt * dims.nzt, n, replace=False)
    x, r = np.divmod(flat, dims.nyt * dims.nzt)
    y, z = np.divmod(r, dims.nzt)
    return np.ascontiguousarray(np.stack([x + 1, y + 1, z + 1], 1).astype(np.int32))


@pytest.mark.parametrize("order", [0, 4])
def test_gpu_addkinsrc_matches_oracle(env, order):
    from scipy import signal
    torch, _lib, lib, ok = env
    dims = Dims(40, 36, 48)
    DH, DT = 50.0, 0.002
    _set_grid(lib, dims, DH, DT)
    rng = np.random.default_rng(11 + order)
    n = 300
    psrc = _cells(rng, dims, n)
    cols = [np.zeros(2 * n, np.float32) for _ in range(6)]
    cols[0][0::2] = rng.integers(1, 3, n)           # Brune or Liu
    cols[1][0::2] = rng.uniform(0.5, 2.0, n)        # slip
    cols[2][0::2] = rng.uniform(0.0, 0.08, n)       # rupture time
    cols[3][0::2] = rng.uniform(0.02, 0.1, n)       # rise time
    cols[4][0::2] = rng.uniform(1e4, 4e4, n)        # area
    cols[0][1::2] = rng.uniform(0, 360, n)          # strike
    cols[1][1::2] = rng.uniform(10, 90, n)          # dip
    cols[2][1::2] = rng.uniform(-180, 180, n)       # rake
    mu = rng.uniform(1e10, 3e10, dims.size).astype(np.float32)
    if order:
        b, a = signal.butter(order, 0.3)
        bd, ad = np.ascontiguousarray(b), np.ascontiguousarray(a)
        lib.SetDeviceFilterParameters(order, bd.ctypes.data, ad.ctypes.data)
    else:
        b = a = None
        lib.SetDeviceFilterParameters(0, None, None)
    _lib.check()
    nf = order + 1
    host = [np.zeros(dims.size, np.float32) for _ in range(6)]
    mom = np.zeros(n, np.float32)
    filt = np.zeros(n * nf, np.float64)
    dev = [torch.zeros(dims.size, device="cuda") for _ in range(6)]
    dmom = torch.zeros(n, device="cuda")
    dfilt = torch.zeros(n * nf, dtype=torch.float64, device="cuda")
    dpsrc = torch.from_numpy(psrc.reshape(-1).copy()).cuda()
    dcols = [torch.from_numpy(c).cuda() for c in cols]
    dmu = torch.from_numpy(mu).cuda()
    st = torch.cuda.current_stream().cuda_stream
    print(f"addkinsrc n={n} order={order}", flush=True)
    for step in range(1, 81):
        lib.addkinsrc_H(step, 3, dpsrc.data_ptr(), n, st, dmu.data_ptr(),
                        *(c.data_ptr() for c in dcols), *(f.data_ptr() for f in dev),
                        dmom.data_ptr(), dfilt.data_ptr(), 0)
        ok.addkinsrc(step, psrc, mu, cols[:5], host, mom, filt if order else None, b, a,
                     np.float32(DT), np.float32(DH), dims)
    torch.cuda.synchronize()
    _lib.check()
    for name, g, w in [("mom", dmom, mom), ("filter", dfilt, filt)] + \
            [(f"field{c}", dev[c], host[c]) for c in range(6)]:
        g = g.cpu().numpy()
        scale = max(float(np.abs(w).max()), 1e-30)
        err = float(np.abs(g.astype(np.float64) - w).max()) / scale
        assert err < 1e-5, (name, err)
    assert np.count_nonzero(mom) > n // 2
    lib.SetDeviceFilterParameters(0, None, None)


def test_gpu_frcvel_matches_oracle(env):
    torch, _lib, lib, ok = env
    dims = Dims(24, 20, 32)
    DT = 0.004
    _set_grid(lib, dims, 50.0, DT)
    rng = np.random.default_rng(5)
    n, read_step, tskp = 64, 6, 3
    psrc = _cells(rng, dims, n)
    cols = [rng.standard_normal((read_step + 1) * n).astype(np.float32) for _ in range(3)]
    u = [rng.standard_normal(dims.size).astype(np.float32) for _ in range(3)]
    du = [torch.from_numpy(x.copy()).cuda() for x in u]
    dcols = [torch.from_numpy(c.copy()).cuda() for c in cols]
    dpsrc = torch.from_numpy(psrc.reshape(-1).copy()).cuda()
    xmin, xmax = 14, 30          # a band of x indices (array index x + 9)
    st = torch.cuda.current_stream().cuda_stream
    print(f"frcvel n={n} read_step={read_step} tskp={tskp}", flush=True)
    for step in range(1, read_step * tskp + 5):
        lib.frcvel_H(step, read_step, 3, dpsrc.data_ptr(), n, tskp, st,
                     *(c.data_ptr() for c in dcols), 0, 0, 0, *(x.data_ptr() for x in du),
                     xmin, xmax, 0)
        ok.frcvel(step, read_step, psrc, tskp, cols, u, xmin, xmax, np.float32(DT), dims)
    torch.cuda.synchronize()
    _lib.check()
    for name, g, w in [("u1", du[0], u[0]), ("v1", du[1], u[1]), ("w1", du[2], u[2]),
                       ("axx", dcols[0], cols[0]), ("ayy", dcols[1], cols[1]),
                       ("azz", dcols[2], cols[2])]:
        assert first_mismatch(g.cpu().numpy(), w) is None, name


def test_filter_order_is_checked(env):
    torch, _lib, lib, ok = env
    b = np.ones(25)
    lib.awp_clear_error()
    lib.SetDeviceFilterParameters(20, b.ctypes.data, b.ctypes.data)
    assert lib.awp_last_error() == 1
    lib.awp_clear_error()


def test_gpu_kinematic_run_matches_oracle(env):
    """A whole run with a kinematic fault in place of the point sources
    (pmcl3d's IFAULT = 5 order): 40 subfaults, Brune and Liu, a 2nd-order
    filter, 60 steps of the anelastic step.  Traces and wavefield agree with
    the oracle's run to a relative 1e-5 (the source time functions call the
    device's transcendental functions)."""
    from scipy import signal
    from awp_amd import configs
    from awp_amd.kinematic import KinematicFault
    from awp_amd.solver import Solver
    from oracle.oracle import OracleSolver
    torch, _lib, lib, ok = env
    w = configs.anelastic(nx=48, ny=40, nz=32, nt=60, DH=200.0, DT=0.01)
    med = w.medium()
    rng = np.random.default_rng(12)
    n = 40
    cells = _cells(rng, med.dims, n)
    cells[:, 2] = 1 + cells[:, 2] % 20
    _, keep = np.unique(cells, axis=0, return_index=True)
    cells = cells[np.sort(keep)]
    n = len(cells)
    b, a = signal.butter(2, 0.4)
    fault = KinematicFault(cells=cells, stf_type=1 + np.arange(n) % 2,
                           slip=rng.uniform(0.5, 2.0, n), ruptime=rng.uniform(0.0, 0.2, n),
                           risetime=rng.uniform(0.1, 0.3, n), area=np.full(n, 200.0 ** 2),
                           strike=rng.uniform(0, 360, n), dip=rng.uniform(30, 90, n),
                           rake=rng.uniform(-180, 180, n), b=b, a=a)
    print(f"kinematic run n={n}", flush=True)
    g = Solver(med, None, w.receivers, nt_record=w.nt, kinematic=fault)
    g.run(w.nt, check_every=10)
    tg = g.traces()
    o = OracleSolver(med, None, w.receivers, nt_record=w.nt, dtype=np.float32, kinematic=fault)
    o.run(w.nt)
    to = o.traces()
    assert np.abs(to).max() > 0
    err = float(np.linalg.norm(tg - to) / np.linalg.norm(to))
    print(f"kinematic run: trace rel-L2 vs oracle {err:.2e}")
    assert err < 1e-5
    for name in ("u1", "w1", "xx", "xy"):
        got, want = g.field(name), o.f[name]
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max(), name
    lib.SetDeviceFilterParameters(0, None, None)


def test_gpu_addkinsrc_at_the_benchmark_size(env):
    """80,000 subfaults on one vertical plane of the 512 x 512 x 256 layout
    with a 4th-order filter, five steps after every rupture time: stresses,
    moments and filter states within 1e-5 of the oracle."""
    from scipy import signal
    torch, _lib, lib, ok = env
    dims = Dims(512, 512, 256)
    DH, DT = 50.0, 0.0025
    _set_grid(lib, dims, DH, DT)
    rng = np.random.default_rng(31)
    flat = rng.choice(460 * 174, 80000, replace=False)
    psrc = np.ascontiguousarray(np.stack([np.full(80000, 256), 27 + flat // 174, 1 + flat % 174],
                                         1).astype(np.int32))
    n = len(psrc)
    cols = [np.zeros(2 * n, np.float32) for _ in range(6)]
    cols[0][0::2] = 1 + np.arange(n) % 2
    cols[1][0::2] = rng.uniform(0.5, 3.0, n)
    cols[2][0::2] = rng.uniform(0.0, 0.01, n)
    cols[3][0::2] = rng.uniform(0.05, 0.2, n)
    cols[4][0::2] = DH * DH
    cols[0][1::2], cols[1][1::2], cols[2][1::2] = 30.0, 90.0, rng.uniform(150, 210, n)
    b, a = signal.butter(4, 0.2)
    lib.SetDeviceFilterParameters(4, np.ascontiguousarray(b).ctypes.data,
                                  np.ascontiguousarray(a).ctypes.data)
    mu = np.full(dims.size, 3.0e10, np.float32)
    host = [np.zeros(dims.size, np.float32) for _ in range(6)]
    mom = np.zeros(n, np.float32)
    filt = np.zeros(n * 5, np.float64)
    dev = [torch.zeros(dims.size, device="cuda") for _ in range(6)]
    dmom = torch.zeros(n, device="cuda")
    dfilt = torch.zeros(n * 5, dtype=torch.float64, device="cuda")
    dpsrc = torch.from_numpy(psrc.reshape(-1).copy()).cuda()
    dcols = [torch.from_numpy(c).cuda() for c in cols]
    dmu = torch.from_numpy(mu).cuda()
    st = torch.cuda.current_stream().cuda_stream
    print(f"addkinsrc n={n} dims={dims}", flush=True)
    for step in range(1, 10):
        lib.addkinsrc_H(step, 3, dpsrc.data_ptr(), n, st, dmu.data_ptr(),
                        *(c.data_ptr() for c in dcols), *(f.data_ptr() for f in dev),
                        dmom.data_ptr(), dfilt.data_ptr(), 0)
        ok.addkinsrc(step, psrc, mu, cols[:5], host, mom, filt, b, a, np.float32(DT),
                     np.float32(DH), dims)
    torch.cuda.synchronize()
    _lib.check()
    for name, g, w in [("mom", dmom, mom), ("filter", dfilt, filt)] + \
            [(f"field{c}", dev[c], host[c]) for c in range(6)]:
        g = g.cpu().numpy()
        err = float(np.abs(g.astype(np.float64) - w).max()) / max(float(np.abs(w).max()), 1e-30)
        assert err < 1e-5, (name, err)
    assert np.count_nonzero(mom) == n
    lib.SetDeviceFilterParameters(0, None, None)
