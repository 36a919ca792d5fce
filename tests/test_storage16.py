"""16-bit wavefield storage (bfloat16 / binary16 with per-field power-of-two
scales, binary32 arithmetic).

CPU: the oracle's codecs against numpy's binary16 conversion (round to
nearest even, subnormals, overflow) and against the definition of
round-to-nearest-even bfloat16; the scale policy; the oracle's 16-bit run
against the FP64 oracle (measured: bf16 1.1e-2 receiver rel-L2, fp16 0.26 --
a single power-of-two scale per field leaves binary16 ~24 octaves of normal
range, and the far field sits ~10^8 below the point-source cell, so most
far-field values are subnormal).

GPU: every 16-bit launcher bit-exact against the oracle's 16-bit run; the
GPU codecs against numpy / the definition on every binary16 bit pattern and
on random and tie-point binary32 values.
"""
import ctypes as C
import math

import numpy as np
import pytest

from awp_amd import configs
from awp_amd.storage import (FIELDS16, default_check_every, initial_scales, pow2_scale,
                             rescale_shift)


def _enc(o, x, fmt, inv=1.0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.empty(x.shape, dtype=np.uint16)
    o.load().awp_oracle_s16_encode(x.ctypes.data, q.ctypes.data, x.size, fmt, inv)
    return q


def _dec(o, q, fmt, s=1.0):
    q = np.ascontiguousarray(q, dtype=np.uint16)
    x = np.empty(q.shape, dtype=np.float32)
    o.load().awp_oracle_s16_decode(q.ctypes.data, x.ctypes.data, q.size, fmt, s)
    return x


def _test_floats(rng, n=200000):
    """Random binary32 values over the binary16 range and beyond, plus exact
    binary16 / bfloat16 tie points."""
    e = rng.uniform(-30, 18, n)
    x = (rng.choice([-1, 1], n) * np.exp2(e)).astype(np.float32)
    h = np.arange(0, 0x7c00, 37, dtype=np.uint16).view(np.float16).astype(np.float32)
    h_next = np.nextafter(h.astype(np.float16), np.float16(np.inf)).astype(np.float32)
    ties16 = ((h.astype(np.float64) + h_next) / 2).astype(np.float32)
    b = (rng.integers(0, 0x7f80, 5000).astype(np.uint32) << 16)
    ties_bf = (b | 0x8000).view(np.float32)
    # every bfloat16 pattern of the normal range with the low halves at and
    # around the tie point, and the largest values (which round to infinity)
    allb = np.arange(0x0080, 0x7f80, dtype=np.uint32) << 16
    near_bf = (allb[:, None] | np.uint32([0x7fff, 0x8000, 0x8001, 0xffff])[None, :])
    near_bf = near_bf.ravel().view(np.float32)
    return np.concatenate([x, h, ties16, -ties16, ties_bf, -ties_bf, near_bf, -near_bf,
                           np.uint32([0x7f7f7fff, 0x7f7f8000, 0x7f7fffff]).view(np.float32),
                           np.float32([0.0, -0.0, 65504, 65519.99, 65520, 1e10, 6e-8, 3e-8])])


def _bf16_reference(x):
    """Nearest bfloat16 by value, ties to an even last bit (the definition)."""
    b = x.view(np.uint32)
    lo = (b & 0xFFFF0000).astype(np.uint32)
    hi = lo + 0x10000
    xl = lo.view(np.float32).astype(np.float64)
    xv = x.astype(np.float64)
    # above the largest bfloat16 the next candidate is 2^128 (infinity)
    xh = np.where((hi & 0x7FFFFFFF) == 0x7F800000, np.copysign(2.0 ** 128, xv),
                  hi.view(np.float32).astype(np.float64))
    dl, dh = np.abs(xv - xl), np.abs(xh - xv)
    pick_hi = (dh < dl) | ((dh == dl) & (((lo >> 16) & 1) == 1))
    return np.where(pick_hi, hi >> 16, lo >> 16).astype(np.uint16)


def test_oracle_codecs(oracle_lib):
    rng = np.random.default_rng(70)
    x = _test_floats(rng)
    finite = np.isfinite(x) & (np.abs(x) >= np.float32(2.0 ** -126))
    x = x[finite]
    q16 = _enc(oracle_lib, x, 2)
    with np.errstate(over="ignore"):
        want16 = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(q16, want16)
    allh = np.arange(65536, dtype=np.uint16)
    fin = np.isfinite(allh.view(np.float16))
    d16 = _dec(oracle_lib, allh, 2)
    assert np.array_equal(d16[fin], allh.view(np.float16)[fin].astype(np.float32))
    qb = _enc(oracle_lib, x, 1)
    assert np.array_equal(qb, _bf16_reference(x))
    db = _dec(oracle_lib, qb, 1)
    assert np.array_equal(db.view(np.uint32), qb.astype(np.uint32) << 16)


def test_scale_policy():
    assert pow2_scale(1000.0, 0) == 512.0           # 1000 / 512 in [1, 2)
    assert pow2_scale(1000.0, 10) == 0.5
    assert rescale_shift("fp16", 0.0) == 0
    assert rescale_shift("fp16", 2.0 ** 10.5) == 0
    assert rescale_shift("fp16", 2.0 ** 13.2) == -3
    assert rescale_shift("bf16", 2.0 ** -5) == 5
    with pytest.raises(Exception):
        rescale_shift("fp16", math.inf)
    w = configs.anelastic(nx=24, ny=24, nz=24, nt=20, DH=200.0, DT=0.01)
    sc = initial_scales("fp16", w.medium(), w.sources)
    assert set(sc) == set(FIELDS16)
    assert all(math.frexp(v)[0] == 0.5 for v in sc.values())
    # bfloat16: unit scales and no checks by default, the binary16 policy on request
    assert set(initial_scales("bf16", w.medium(), w.sources).values()) == {1.0}
    scaled = initial_scales("bf16", w.medium(), w.sources, scaling="scaled")
    assert all(math.frexp(v)[0] == 0.5 for v in scaled.values()) and set(scaled.values()) != {1.0}
    assert default_check_every("bf16") == 0 and default_check_every("fp16") == 8
    assert default_check_every("bf16", "scaled") == 8


def _small():
    return configs.anelastic(nx=48, ny=40, nz=32, nt=50, DH=200.0, DT=0.01)


def test_oracle_16bit_accuracy(oracle_lib):
    from oracle.oracle import OracleSolver, OracleSolver16
    w = _small()
    med = w.medium()
    o64 = OracleSolver(med, w.sources, w.receivers, nt_record=w.nt, dtype=np.float64)
    o64.run(w.nt)
    ref = o64.traces()
    err = {}
    for st, scaling in (("bf16", "auto"), ("bf16", "scaled"), ("fp16", "auto")):
        o = OracleSolver16(med, w.sources, w.receivers, nt_record=w.nt, storage=st,
                           scaling=scaling)
        o.run(w.nt)
        err[st, scaling] = float(np.linalg.norm(o.traces() - ref) / np.linalg.norm(ref))
        if st == "bf16" and scaling == "auto":
            assert o.rescales == 0 and set(o.scales.values()) == {1.0}
        else:
            assert o.rescales > 0
    print(f"16-bit storage, receiver rel-L2 vs FP64 oracle: {err}")
    assert err["bf16", "auto"] < 0.03
    assert err["bf16", "scaled"] < 0.03
    assert err["fp16", "auto"] < 0.5


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awp_amd import _lib
    return torch, _lib.load(), _lib


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [1, 2], ids=["bf16", "fp16"])
def test_gpu_codecs(gpu, oracle_lib, fmt):
    torch, lib, _lib = gpu
    rng = np.random.default_rng(71)
    x = _test_floats(rng)
    x = x[np.isfinite(x) & (np.abs(x) >= np.float32(2.0 ** -126))]
    xd = torch.from_numpy(x).cuda()
    qd = torch.empty(x.shape, dtype=torch.int16, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    lib.awp_s16_encode(xd.data_ptr(), qd.data_ptr(), x.size, fmt, 1.0, st)
    _lib.check()
    got = qd.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, _enc(oracle_lib, x, fmt))
    allh = np.arange(65536, dtype=np.uint16)
    qa = torch.from_numpy(allh.view(np.int16).copy()).cuda()
    out = torch.empty(65536, dtype=torch.float32, device="cuda")
    lib.awp_s16_decode(qa.data_ptr(), out.data_ptr(), 65536, fmt, 1.0, st)
    _lib.check()
    d = out.cpu().numpy()
    want = _dec(oracle_lib, allh, fmt)
    fin = np.isfinite(want)
    assert np.array_equal(d[fin].view(np.uint32), want[fin].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("storage,scaling,texture", [("bf16", "auto", "arrays"),
                                                     ("bf16", "scaled", "arrays"),
                                                     ("fp16", "auto", "arrays"),
                                                     ("bf16", "auto", "rule"),
                                                     ("fp16", "auto", "table")])
def test_gpu_16bit_run_bit_exact(gpu, oracle_lib, storage, scaling, texture):
    """A 50-step anelastic run with 16-bit storage (velocity, stress with FS
    ghosts, source, fstr, receivers; with scaling, rescaling every 8 steps):
    stored values and traces bit-exact against the oracle's 16-bit run.
    bfloat16 "auto" keeps unit scales and no checks, "scaled" follows the
    binary16 policy.  texture "rule" / "table": the stress update with the
    8-entry relaxation-texture table (awp_dstrqc_tex_s16), ww from
    init_texture's parity rule or read."""
    from awp_amd.solver import Solver16
    from oracle.oracle import OracleSolver16
    w = _small()
    med = w.medium()
    g = Solver16(med, w.sources, w.receivers, nt_record=w.nt, storage=storage, scaling=scaling,
                 texture=texture)
    assert g.levels[0].texmode == texture
    g.run(w.nt, check_every=10)
    o = OracleSolver16(med, w.sources, w.receivers, nt_record=w.nt, storage=storage,
                       scaling=scaling)
    o.run(w.nt)
    assert g.scales == o.scales and g.rescales == o.rescales
    if storage == "bf16" and scaling == "auto":
        assert g.rescales == 0 and set(g.scales.values()) == {1.0}
    else:
        assert g.rescales > 0
    assert np.array_equal(g.traces(), o.traces())
    for n in FIELDS16:
        a, b = g.raw(n), o.levels[0].q[n]
        bad = np.argwhere(a != b)
        assert len(bad) == 0, (n, len(bad), tuple(bad[0]))


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [1, 2], ids=["bf16", "fp16"])
def test_gpu_absmax_and_rescale(gpu, oracle_lib, fmt):
    """awp_s16_absmax and awp_s16_rescale against the host codecs: ragged
    lengths, 16-byte aligned and unaligned starts (vector and scalar paths),
    a NaN anywhere."""
    torch, lib, _lib = gpu
    rng = np.random.default_rng(5)
    st = torch.cuda.current_stream().cuda_stream
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    for n, off in ((1, 0), (7, 0), (8, 0), (1000003, 0), (1000003, 1), (4096, 3)):
        m = n + off
        x = (rng.standard_normal(m) * 10.0 ** rng.uniform(-3, 3, m)).astype(np.float32)
        q = _enc(oracle_lib, x, fmt)
        qd = torch.from_numpy(q.view(np.int16).copy()).cuda()
        lib.awp_s16_absmax(qd.data_ptr() + 2 * off, n, fmt, out.data_ptr(), st)
        _lib.check()
        assert out.item() == float(np.abs(_dec(oracle_lib, q[off:], fmt)).max()), (n, off)
        lib.awp_s16_rescale(qd.data_ptr() + 2 * off, n, fmt, -3, st)
        _lib.check()
        want = q.copy()
        want[off:] = _enc(oracle_lib, _dec(oracle_lib, q[off:], fmt) * np.float32(2.0 ** -3), fmt)
        assert np.array_equal(qd.cpu().numpy().view(np.uint16), want), (n, off)
        if n > 5:
            qn = q.copy()
            qn[off + n - 2] = 0x7FC0 if fmt == 1 else 0x7E00
            qd = torch.from_numpy(qn.view(np.int16).copy()).cuda()
            lib.awp_s16_absmax(qd.data_ptr() + 2 * off, n, fmt, out.data_ptr(), st)
            _lib.check()
            assert math.isnan(out.item()), (n, off)
