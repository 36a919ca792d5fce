"""Seeded random states on the padded grid for kernel-level parity tests.

Every array, ghosts included, is filled, so that an index error anywhere in
a stencil changes the result."""
import numpy as np

from awp_amd.grid import Dims
from awp_amd.media import q_model

WAVE = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz", "r1", "r2", "r3", "r4", "r5", "r6")


def random_state(dims: Dims, seed: int, qmax: float = 0.02):
    rng = np.random.default_rng(seed)
    sh = dims.shape
    st = {}
    for n in ("u1", "v1", "w1"):
        st[n] = rng.standard_normal(sh).astype(np.float32)
    for n in ("xx", "yy", "zz", "xy", "xz", "yz"):
        st[n] = (1e6 * rng.standard_normal(sh)).astype(np.float32)
    for n in ("r1", "r2", "r3", "r4", "r5", "r6"):
        st[n] = (1e4 * rng.standard_normal(sh)).astype(np.float32)
    med = {}
    med["d1"] = rng.uniform(1800.0, 3000.0, sh).astype(np.float32)
    med["mu"] = (1.0 / (med["d1"] * rng.uniform(1000.0, 4000.0, sh) ** 2)).astype(np.float32)
    med["lam"] = (1.0 / (med["d1"] * rng.uniform(1000.0, 9000.0, sh) ** 2)).astype(np.float32)
    # Q^-1 straddling the 1/(2q) <= 200 branch point q = 0.0025
    med["qp"] = rng.uniform(0.0, qmax, sh).astype(np.float32)
    med["qs"] = rng.uniform(0.0, qmax, sh).astype(np.float32)
    med["vx1"] = rng.uniform(0.0, 0.5, sh).astype(np.float32)
    med["vx2"] = rng.uniform(0.0, 1.0, sh).astype(np.float32)
    med["wwo"] = rng.uniform(0.05, 14.0, sh).astype(np.float32)
    med["ww"] = rng.integers(1, 9, sh).astype(np.int32)
    med["coeff"] = q_model(0.0).coeff.astype(np.float32)
    med["lam_mu"] = rng.uniform(0.1, 0.5, (sh[0], sh[1])).astype(np.float32)
    med["dcrjx"] = rng.uniform(0.9, 1.0, sh[0]).astype(np.float32)
    med["dcrjy"] = rng.uniform(0.9, 1.0, sh[1]).astype(np.float32)
    med["dcrjz"] = rng.uniform(0.9, 1.0, sh[2]).astype(np.float32)
    return st, med


def first_mismatch(a: np.ndarray, b: np.ndarray):
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if len(bad) == 0:
        return None
    idx = tuple(bad[0])
    return dict(count=len(bad), index=idx, got=float(a[idx]), want=float(b[idx]))
