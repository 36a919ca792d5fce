"""GPU parity: every launcher of libawp_hip.so, called through the C ABI on
seeded random states, against the FP32 oracle on the same inputs.

Bar: bit-exact.  Both sides evaluate the reference's expression trees in the
same order without contraction, with single-precision denormals flushed (the
reference's -use_fast_math mode), so any difference is a bug.
"""
import ctypes as C

import numpy as np
import pytest

from awp_amd.grid import Dims
from tests._fields import WAVE, first_mismatch, random_state

pytestmark = pytest.mark.gpu

# odd, non-multiple-of-64 sizes exercise every tile edge
SIZES = [Dims(24, 20, 40), Dims(37, 29, 70), Dims(16, 66, 130)]


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awp_amd import _lib
    lib = _lib.load()
    return torch, lib, _lib


def _set_grid(lib, dims: Dims, DH=50.0, DT=0.0025):
    F, I = C.c_float, C.c_int
    z9 = (F * 9)()
    lib.SetDeviceConstValue((F * 1)(DH), F(DT), (I * 1)(dims.nxt), (I * 1)(dims.nyt),
                            (I * 1)(dims.nzt), 1, 0.0, 0.0, z9, z9)


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
        got = dev[n].cpu().numpy()
        mm = first_mismatch(got, host[n])
        assert mm is None, f"{tag}: field {n} differs: {mm}"


@pytest.mark.parametrize("dims", SIZES, ids=lambda d: f"{d.nxt}x{d.nyt}x{d.nzt}")
@pytest.mark.parametrize("bj,ichunk,remap", [(4, 0, 1), (1, 3, 0), (8, 7, 1), (2, 1, 1), (16, 2, 1)])
def test_dvelcx_bit_exact(gpu, oracle_lib, dims, bj, ichunk, remap):
    torch, lib, _lib = gpu
    _lib.set_tuning("vel_bj", bj)
    _lib.set_tuning("vel_ichunk", ichunk)
    _lib.set_tuning("xcd_remap", remap)
    st, med = random_state(dims, seed=11 + dims.nzt)
    DH, DT = 50.0, 0.0025
    _set_grid(lib, dims, DH, DT)
    dev = _dev(torch, {**st, **med})
    s_i, e_i = dims.i_range
    names = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz")
    lib.dvelcx_H_opt(*(_p(dev[n]) for n in names), _p(dev["dcrjx"]), _p(dev["dcrjy"]),
                     _p(dev["dcrjz"]), _p(dev["d1"]), dims.nyt, dims.nzt, _stream(torch), s_i,
                     e_i, 0, 1)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    oracle_lib.load().awp_oracle_dvelcx_f32(
        *(_np(host[n]) for n in names), _np(med["dcrjx"]), _np(med["dcrjy"]),
        _np(med["dcrjz"]), _np(med["d1"]), dims.nxt, dims.nyt, dims.nzt, DH, DT, s_i, e_i, 0, 1)
    _compare(torch, dev, host, WAVE, "dvelcx")
    _lib.set_tuning("vel_bj", 4)
    _lib.set_tuning("vel_ichunk", 0)
    _lib.set_tuning("xcd_remap", 1)


def _run_dstrqc(torch, lib, _lib, oracle_lib, dims, variant, NPC=0, sub=None, d_i_grids=(0, 1)):
    st, med = random_state(dims, seed=5 + dims.nxt)
    DH, DT = 50.0, 0.0025
    _set_grid(lib, dims, DH, DT)
    dev = _dev(torch, {**st, **med})
    s_i, e_i = dims.i_range
    s_j, e_j = dims.j_range
    if sub is not None:
        s_i, e_i, s_j, e_j = sub
    names = ("xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6", "u1", "v1",
             "w1")
    fn = lib.dstrqc_H if variant == "dstrqc" else lib.dstrqc_H_new
    fn(*(_p(dev[n]) for n in names), _p(dev["lam"]), _p(dev["mu"]), _p(dev["qp"]),
       _p(dev["coeff"]), _p(dev["qs"]), _p(dev["dcrjx"]), _p(dev["dcrjy"]), _p(dev["dcrjz"]),
       dims.nyt, dims.nzt, _stream(torch), _p(dev["lam_mu"]), _p(dev["vx1"]), _p(dev["vx2"]),
       _p(dev["ww"]), _p(dev["wwo"]), dims.nxt, NPC, 0, 0, s_i, e_i, s_j, e_j, 0)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    oname = "awp_oracle_dstrqc_f32" if variant == "dstrqc" else "awp_oracle_dstrqc_new_f32"
    getattr(oracle_lib.load(), oname)(
        *(_np(host[n]) for n in names), _np(med["lam"]), _np(med["mu"]), _np(med["qp"]),
        _np(med["coeff"]), _np(med["qs"]), _np(med["dcrjx"]), _np(med["dcrjy"]),
        _np(med["dcrjz"]), _np(med["lam_mu"]), _np(med["vx1"]), _np(med["vx2"]),
        _np(med["ww"]), _np(med["wwo"]), dims.nxt, dims.nyt, dims.nzt, DH, DT, dims.nxt, NPC, 0,
        0, s_i, e_i, s_j, e_j, 0)
    _compare(torch, dev, host, WAVE, variant)


@pytest.mark.parametrize("dims", SIZES, ids=lambda d: f"{d.nxt}x{d.nyt}x{d.nzt}")
@pytest.mark.parametrize("variant", ["dstrqc", "dstrqc_new"])
def test_dstrqc_bit_exact(gpu, oracle_lib, dims, variant):
    torch, lib, _lib = gpu
    _run_dstrqc(torch, lib, _lib, oracle_lib, dims, variant)


@pytest.mark.parametrize("bj,ichunk,remap", [(1, 5, 0), (8, 2, 1), (2, 0, 1), (4, 3, 1), (16, 1, 0)])
def test_dstrqc_launch_shapes(gpu, oracle_lib, bj, ichunk, remap):
    torch, lib, _lib = gpu
    _lib.set_tuning("str_bj", bj)
    _lib.set_tuning("str_ichunk", ichunk)
    _lib.set_tuning("xcd_remap", remap)
    try:
        _run_dstrqc(torch, lib, _lib, oracle_lib, Dims(21, 19, 67), "dstrqc")
    finally:
        _lib.set_tuning("str_bj", 16)
        _lib.set_tuning("str_ichunk", 0)
        _lib.set_tuning("xcd_remap", 1)


def test_dstrqc_strips_and_periodic(gpu, oracle_lib):
    """pmcl3d's strip launches (sub-ranges of i and j) and NPC=2."""
    torch, lib, _lib = gpu
    d = Dims(20, 18, 36)
    _run_dstrqc(torch, lib, _lib, oracle_lib, d, "dstrqc", sub=(12, 27, 10, 27))
    _run_dstrqc(torch, lib, _lib, oracle_lib, d, "dstrqc_new", NPC=2, sub=(10, 11, 11, 20))


def test_fstr_and_addsrc_bit_exact(gpu, oracle_lib):
    torch, lib, _lib = gpu
    dims = Dims(22, 18, 40)
    st, med = random_state(dims, seed=3)
    DH, DT = 75.0, 0.004
    _set_grid(lib, dims, DH, DT)
    rng = np.random.default_rng(9)
    nsrc, nst = 7, 12
    psrc = np.stack([rng.integers(1, dims.nxt + 1, nsrc), rng.integers(1, dims.nyt + 1, nsrc),
                     rng.integers(1, dims.nzt + 1, nsrc)], 1).astype(np.int32)
    psrc[:, 2] = dims.nzt + 1 - psrc[:, 2]
    # distinct cells (duplicates race in the reference)
    _, keep = np.unique(psrc, axis=0, return_index=True)
    psrc = np.ascontiguousarray(psrc[np.sort(keep)])
    nsrc = len(psrc)
    cols = [(1e15 * rng.standard_normal(nsrc * nst)).astype(np.float32) for _ in range(6)]
    dev = _dev(torch, st)
    dsrc = torch.from_numpy(psrc.reshape(-1).copy()).cuda()
    dcols = [torch.from_numpy(c).cuda() for c in cols]
    s_i, e_i = dims.i_range
    s_j, e_j = dims.j_range
    order = ("xx", "yy", "zz", "xy", "yz", "xz")
    for step in (1, 5, nst):
        lib.addsrc_H(step, nst, 3, _p(dsrc), nsrc, _stream(torch), *(_p(c) for c in dcols),
                     *(_p(dev[n]) for n in order), 0)
    lib.fstr_H(_p(dev["zz"]), _p(dev["xz"]), _p(dev["yz"]), _stream(torch), s_i, e_i, s_j, e_j)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    o = oracle_lib.load()
    for step in (1, 5, nst):
        o.awp_oracle_addsrc_f32(step, nst, 3, _np(psrc), nsrc, *(_np(c) for c in cols),
                                *(_np(host[n]) for n in order), dims.nxt, dims.nyt, dims.nzt,
                                DH, DT)
    o.awp_oracle_fstr_f32(_np(host["zz"]), _np(host["xz"]), _np(host["yz"]), dims.nxt,
                          dims.nyt, dims.nzt, s_i, e_i, s_j, e_j)
    _compare(torch, dev, host, WAVE, "addsrc+fstr")


def test_velbuffer_and_receivers(gpu):
    torch, lib, _lib = gpu
    dims = Dims(20, 24, 36)
    st, _ = random_state(dims, seed=8)
    _set_grid(lib, dims)
    dev = _dev(torch, st)
    nbgx, nedx, nskpx = 2, 17, 3
    nbgy, nedy, nskpy = 1, 23, 2
    nbgz, nedz, nskpz = 0, 6, 2
    rnx = (nedx - nbgx) // nskpx + 1
    rny = (nedy - nbgy) // nskpy + 1
    rnz = (nedz - nbgz) // nskpz + 1
    bufs = [torch.zeros(rnx * rny * rnz, device="cuda") for _ in range(4)]
    lib.velbuffer_H(_p(dev["u1"]), _p(dev["v1"]), _p(dev["w1"]), _p(dev["u1"]),
                    *(_p(b) for b in bufs), 1, nbgx, nedx, nskpx, nbgy, nedy, nskpy, nbgz, nedz,
                    nskpz, rnx, rny, rnz, _stream(torch), 0, None, 0)
    _lib.check()
    torch.cuda.synchronize()
    got = [b.cpu().numpy().reshape(rnz, rny, rnx) for b in bufs[:3]]
    for kk in range(rnz):
        for jj in range(rny):
            for ii in range(rnx):
                i = 10 + nbgx + ii * nskpx
                j = 10 + nbgy + jj * nskpy
                ko = dims.nzt + 31 - (nbgz + kk * nskpz)
                assert got[0][kk, jj, ii] == st["u1"][i, j, ko]
                assert got[1][kk, jj, ii] == st["v1"][i, j, ko]
                assert got[2][kk, jj, ii] == st["w1"][i, j, ko - 1]
    idx = np.array([[10, 10, dims.ktop], [29, 33, dims.ktop], [15, 20, 40]], dtype=np.int32)
    didx = torch.from_numpy(idx.reshape(-1).copy()).cuda()
    out = torch.zeros((2, 3, 3), device="cuda")
    lib.awp_receivers_gather(_p(dev["u1"]), _p(dev["v1"]), _p(dev["w1"]), _p(didx), 3, _p(out),
                             1, _stream(torch), 0)
    _lib.check()
    o = out.cpu().numpy()[1]
    for n, (i, j, k) in enumerate(idx):
        assert o[n, 0] == st["u1"][i, j, k]
        assert o[n, 1] == st["v1"][i, j, k]
        assert o[n, 2] == st["w1"][i, j, k - 1]


def _plastic_state(dims, seed):
    from awp_amd.plasticity import lithostatic_sigma2
    st, med = random_state(dims, seed=seed)
    rng = np.random.default_rng(seed + 100)
    sh = dims.shape
    med["sigma2"] = lithostatic_sigma2(dims, 75.0, med["d1"].astype(np.float64)).astype(np.float32)
    med["cohes"] = rng.uniform(0.0, 3.0e6, sh).astype(np.float32)
    med["phi"] = rng.uniform(15.0, 40.0, sh).astype(np.float32)
    st["neta"] = rng.uniform(0.0, 1e-3, sh).astype(np.float32)
    y = rng.uniform(0.3, 1.0, sh)
    y[rng.uniform(size=sh) < 0.3] = 1.0
    st["yldfac"] = y.astype(np.float32)
    return st, med


@pytest.mark.parametrize("Tv", [0.0, 0.01])
def test_drprecpc_calc_bit_exact(gpu, oracle_lib, Tv):
    from awp_amd.plasticity import PlasticParams
    torch, lib, _lib = gpu
    dims = Dims(19, 22, 70)
    st, med = _plastic_state(dims, 21)
    pp = PlasticParams(fmajor=1.2, fminor=0.8, strike=(30.0, 0.0, 120.0))
    Rz, RzT = pp.rotation()
    F, I = C.c_float, C.c_int
    lib.SetDeviceConstValue((F * 1)(75.0), F(0.004), (I * 1)(dims.nxt), (I * 1)(dims.nyt),
                            (I * 1)(dims.nzt), 1, pp.fmajor, pp.fminor, (F * 9)(*Rz.tolist()),
                            (F * 9)(*RzT.tolist()))
    dev = _dev(torch, {**st, **med})
    xs, xe, ys, ye = 10, dims.nxt + 9, 11, dims.nyt + 8
    args = [_p(dev[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "mu", "d1", "sigma2",
                                   "yldfac", "cohes", "phi", "neta")]
    if Tv > 0:
        lib.awp_drprecpc_calc_tv(*args, dims.nzt, xs, xe, ys, ye, _stream(torch), 0, Tv)
    else:
        lib.drprecpc_calc_H_opt(*args, dims.nzt, xs, xe, ys, ye, _stream(torch), 0)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    pl = np.ascontiguousarray(pp.packed())
    oracle_lib.load().awp_oracle_drprecpc_calc_f32(
        *(_np(host[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")), _np(med["mu"]),
        _np(med["d1"]), _np(med["sigma2"]), _np(host["yldfac"]), _np(med["cohes"]),
        _np(med["phi"]), _np(host["neta"]), dims.nxt, dims.nyt, dims.nzt, 75.0, 0.004, xs, xe,
        ys, ye, _np(pl), Tv)
    assert (host["yldfac"] < 1).sum() > 100, "too few yielding cells for a meaningful test"
    _compare(torch, dev, host, ("yldfac", "neta"), "drprecpc_calc")


def test_drprecpc_app_bit_exact(gpu, oracle_lib):
    from awp_amd.plasticity import PlasticParams
    torch, lib, _lib = gpu
    dims = Dims(21, 18, 66)
    st, med = _plastic_state(dims, 33)
    pp = PlasticParams(fmajor=1.3, fminor=0.75, strike=(15.0, 0.0, 105.0))
    Rz, RzT = pp.rotation()
    F, I = C.c_float, C.c_int
    lib.SetDeviceConstValue((F * 1)(75.0), F(0.004), (I * 1)(dims.nxt), (I * 1)(dims.nyt),
                            (I * 1)(dims.nzt), 1, pp.fmajor, pp.fminor, (F * 9)(*Rz.tolist()),
                            (F * 9)(*RzT.tolist()))
    dev = _dev(torch, {**st, **med})
    xs, xe, ys, ye = 10, dims.nxt + 9, 10, dims.nyt + 9
    lib.drprecpc_app_H(*(_p(dev[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz", "mu",
                                               "sigma2", "yldfac")), dims.nzt, xs, xe, ys, ye,
                       _stream(torch), 0)
    _lib.check()
    host = {k: v.copy() for k, v in st.items()}
    pl = np.ascontiguousarray(pp.packed())
    oracle_lib.load().awp_oracle_drprecpc_app_f32(
        *(_np(host[n]) for n in ("xx", "yy", "zz", "xy", "xz", "yz")), _np(med["sigma2"]),
        _np(host["yldfac"]), dims.nxt, dims.nyt, dims.nzt, 75.0, xs, xe, ys, ye, _np(pl))
    _compare(torch, dev, host, ("xx", "yy", "zz", "xy", "xz", "yz"), "drprecpc_app")
