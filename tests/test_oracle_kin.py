"""Kinematic finite-fault and prescribed-velocity sources: the oracle
(oracle/awp_oracle_kin.cpp) pinned by what the mathematics fixes.

- compmt: a double-couple moment tensor is traceless, with eigenvalues
  -1, 0, 1, for every strike, dip and rake.
- brune and liu: unit area (Brune: integral of w^2 t exp(-w t) = 1; Liu:
  the normalisation CN makes the piecewise integral 1) and Brune's peak at
  t = 1/w.
- lfilter: scipy.signal.lfilter (direct form II transposed) on a Butterworth
  filter.
- addkinsrc: after the slip the accumulated moment reaches slip * area / mu,
  and the stress increments are proportional to the moment tensor.
- frcvel: linear interpolation between samples, the holding force stored at
  the last prescribed step and applied afterwards.
No reference tests or fixtures cover these kernels.
"""
import numpy as np
import pytest

from awp_amd.grid import Dims


@pytest.fixture(scope="module")
def ok(oracle_lib):
    from oracle import kin
    return kin


def _sym(m):
    xx, yy, zz, xz, yz, xy = (float(v) for v in m)
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def test_moment_tensor_is_a_unit_double_couple(ok):
    rng = np.random.default_rng(1)
    for _ in range(500):
        s, d, r = rng.uniform(0, 2 * np.pi), rng.uniform(0, np.pi / 2), rng.uniform(-np.pi, np.pi)
        M = _sym(ok.compmt(s, d, r))
        assert abs(np.trace(M)) < 1e-6
        assert np.allclose(np.linalg.eigvalsh(M), [-1.0, 0.0, 1.0], atol=3e-6)


def test_brune_unit_area_and_peak(ok):
    for rise in (0.5, 1.0, 2.0):
        f = np.float32(1.0 / rise)
        w = 2 * np.pi * float(f)
        dt = 1e-3 / w
        t = np.arange(1, int(60 / (w * dt))) * dt
        s = np.array([ok.brune(f, v) for v in t], np.float64)
        assert abs(s.sum() * dt - 1.0) < 2e-3
        assert abs(t[np.argmax(s)] - 1.0 / w) < 2 * dt
    assert ok.brune(1.0, 0.0) == 0.0 and ok.brune(1.0, -1.0) == 0.0


def test_liu_unit_area(ok):
    for tau in (0.4, 1.0, 3.0):
        dt = tau * 1e-5
        t = (np.arange(int(tau / dt)) + 0.5) * dt
        s = np.array([ok.liu(tau, v) for v in t], np.float64)
        assert abs(s.sum() * dt - 1.0) < 1e-4
        assert ok.liu(tau, tau) == 0.0 and ok.liu(tau, 2 * tau) == 0.0
        assert min(s) >= 0.0


def test_lfilter_matches_scipy(ok):
    from scipy import signal
    rng = np.random.default_rng(2)
    for order in (1, 2, 4):
        b, a = signal.butter(order, 0.2)
        x = rng.standard_normal(300)
        assert np.allclose(ok.lfilter(b, a, x), signal.lfilter(b, a, x), rtol=1e-12, atol=1e-13)


def _kin_setup(dims, n, stf_types, rng):
    psrc = np.stack([rng.choice(np.arange(1, dims.nxt + 1), n, replace=False),
                     rng.integers(1, dims.nyt + 1, n), rng.integers(1, dims.nzt + 1, n)],
                    1).astype(np.int32)
    cols = [np.zeros(2 * n, np.float32) for _ in range(5)]
    cols[0][0::2] = stf_types
    cols[1][0::2] = rng.uniform(0.5, 2.0, n)        # slip
    cols[2][0::2] = rng.uniform(0.0, 0.05, n)       # rupture time
    cols[3][0::2] = rng.uniform(0.1, 0.3, n)        # rise time
    cols[4][0::2] = rng.uniform(1e4, 4e4, n)        # area
    cols[0][1::2] = rng.uniform(0, 360, n)          # strike
    cols[1][1::2] = rng.uniform(10, 90, n)          # dip
    cols[2][1::2] = rng.uniform(-180, 180, n)       # rake
    return psrc, cols


@pytest.mark.parametrize("stf", [1, 2])
def test_addkinsrc_moment_and_tensor(ok, stf):
    dims = Dims(12, 10, 16)
    rng = np.random.default_rng(3 + stf)
    n = 5
    psrc, cols = _kin_setup(dims, n, stf, rng)
    mu = rng.uniform(1e10, 3e10, dims.size).astype(np.float32)
    fields = [np.zeros(dims.size, np.float32) for _ in range(6)]
    mom = np.zeros(n, np.float32)
    DT, DH = np.float32(5e-4), np.float32(50.0)
    for step in range(1, 3000):
        ok.addkinsrc(step, psrc, mu, cols, fields, mom, None, None, None, DT, DH, dims)
    idx = psrc[:, 0] + 9
    idy = psrc[:, 1] + 9
    idz = psrc[:, 2] + 31
    pos = idx * dims.slice + idy * dims.yline + idz
    m0 = cols[1][0::2].astype(np.float64) * cols[4][0::2] / mu[pos]
    assert np.allclose(mom, m0, rtol=2e-3)
    # stress increments: -DT/DH^3 * moment * tensor, order xx, yy, zz, xy, yz, xz
    order = (0, 1, 2, 5, 4, 3)
    for j in range(n):
        M = ok.compmt(np.float32(cols[0][2 * j + 1] / 180 * np.pi),
                      np.float32(cols[1][2 * j + 1] / 180 * np.pi),
                      np.float32(cols[2][2 * j + 1] / 180 * np.pi))
        got = np.array([fields[c][pos[j]] for c in range(6)], np.float64)
        want = -float(mom[j]) / float(DH) ** 3 * M[list(order)]
        assert np.allclose(got, want, rtol=2e-3, atol=2e-3 * np.abs(want).max())


def test_frcvel_interpolates_then_holds(ok):
    dims = Dims(8, 8, 8)
    psrc = np.array([[3, 4, 5]], np.int32)
    pos = (3 + 9) * dims.slice + (4 + 9) * dims.yline + (5 + 31)
    read_step, tskp = 3, 4
    axx = np.array([0.0, 4.0, 8.0, 12.0], np.float32)   # rows 0..READ_STEP
    ayy = 2 * axx
    azz = -axx
    u = [np.zeros(dims.size, np.float32) for _ in range(3)]
    DT = np.float32(0.5)
    seen = []
    for step in range(1, read_step * tskp + 4):
        ok.frcvel(step, read_step, psrc, tskp, (axx, ayy, azz), u, -1, -1, DT, dims)
        seen.append(float(u[0][pos]))
    # steps 1..12 sample (step / tskp) * 4 = step, linearly
    assert seen[:read_step * tskp - 1] == [float(s) for s in range(1, read_step * tskp)]
    assert seen[read_step * tskp - 1] == 12.0
    # at step 12 the velocity sample row moves to row 0 and row 3 becomes the
    # holding force (12 - u_before) / DT = (12 - 11) / 0.5 = 2
    assert axx[0] == 12.0 and axx[3] == 2.0
    assert seen[read_step * tskp:] == [12.0 + 1.0, 14.0, 15.0]
    # v: samples 0, 8, 16, 24, force (24 - 22) / 0.5 = 4, three steps of 4 * 0.5
    assert float(u[1][pos]) == 24.0 + 3 * 2.0
    # w: samples 0, -4, -8, -12, force (-12 + 11) / 0.5 = -2
    assert float(u[2][pos]) == -12.0 - 3 * 1.0
