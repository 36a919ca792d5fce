"""Per-field power-of-two scales for 16-bit wavefield storage.

Every field f (u1 v1 w1 xx yy zz xy xz yz r1..r6) is stored as q = x / s_f
rounded to bfloat16 or binary16, s_f a power of two.  The policy:

  * initial scales from the sources: the stress bound is the largest
    accumulated moment over a cell volume, sum_t |M'(t)| DT / DH^3; the
    velocity bound is that stress over the smallest impedance rho*vs; the
    memory variables share the stress bound.  Each scale puts its bound at
    2^target in stored units (binary16: 2^12, two octaves below the largest
    finite value 65504; bfloat16: 2^0, its range is binary32's);
  * every check_every steps, the largest stored magnitude of each field is
    measured and the field is rescaled by an exact power of two when its
    exponent drifts more than two octaves from 2^target (binary16 target
    2^10), so the values keep the format's full precision and range;
  * a non-finite largest magnitude (binary16 overflow) is an error.
bfloat16 keeps binary32's exponent range, so a power-of-two scale changes none
of its stored significands (only where binary32's flush of subnormals sets
in).  Its default policy ("auto") is therefore unit scales and no drift check;
scaling="scaled" applies the binary16 policy to bfloat16 too.
Both the GPU driver and the oracle's 16-bit driver apply this same policy to
bit-identical states, so they take the same decisions.
"""
from __future__ import annotations

import math

import numpy as np

FIELDS16 = ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
            "r1", "r2", "r3", "r4", "r5", "r6")
FORMATS = {"bf16": 1, "fp16": 2}
INIT_TARGET = {"bf16": 0, "fp16": 12}
TARGET = {"bf16": 0, "fp16": 10}
DRIFT = 2


class StorageOverflow(RuntimeError):
    pass


def pow2_scale(bound: float, target: int) -> float:
    """Power of two s with bound / s in [2^target, 2^(target+1))."""
    if not (bound > 0.0) or not math.isfinite(bound):
        return 1.0
    e = math.floor(math.log2(bound)) - target
    return math.ldexp(1.0, max(min(e, 119), -120))


SCALINGS = ("auto", "scaled")


def uses_scaling(fmt: str, scaling: str = "auto") -> bool:
    if scaling not in SCALINGS:
        raise ValueError(f"scaling must be one of {SCALINGS}")
    return not (fmt == "bf16" and scaling == "auto")


def default_check_every(fmt: str, scaling: str = "auto") -> int:
    """Steps between scale checks: 8, or 0 (never) for unit-scale bfloat16."""
    return 8 if uses_scaling(fmt, scaling) else 0


def initial_scales(fmt: str, medium, sources, scaling: str = "auto") -> dict:
    """Per-field scales before the first step (see the module docstring)."""
    scales = {f: 1.0 for f in FIELDS16}
    if not uses_scaling(fmt, scaling) or sources is None or sources.npsrc == 0:
        return scales
    DH, DT = float(medium.DH), float(medium.DT)
    acc = np.abs(sources.rates).sum(axis=1).max() * DT  # largest accumulated moment [N m]
    sig = acc / DH ** 3
    d = medium.dims
    ii, jj, kk = d.interior()
    rho = medium.d1[ii, jj, kk]
    vs = np.sqrt(1.0 / (medium.mu[ii, jj, kk] * rho))
    vel = sig / float((rho * vs).min())
    t = INIT_TARGET[fmt]
    for f in FIELDS16:
        b = vel if f in ("u1", "v1", "w1") else sig
        scales[f] = pow2_scale(b, t)
    return scales


def rescale_shift(fmt: str, m: float) -> int:
    """Exponent shift for a field whose largest stored magnitude is m (0 when
    it stays, or when the field is zero)."""
    if m == 0.0:
        return 0
    if not math.isfinite(m):
        raise StorageOverflow(f"{fmt} storage overflowed (largest stored magnitude {m})")
    e = math.floor(math.log2(m))
    t = TARGET[fmt]
    if abs(e - t) <= DRIFT:
        return 0
    return t - e


class ScaledStorage:
    """Mixin: scale bookkeeping for a driver with _absmax_all() -> {field:
    largest stored magnitude at unit scale} and _rescale(name, shift)."""

    storage: str
    scales: dict
    check_every: int
    rescales: int = 0

    def scale_array(self):
        return np.array([self.scales[f] for f in FIELDS16], dtype=np.float32)

    def check_scales(self) -> list:
        changed = []
        ms = self._absmax_all()
        for f in FIELDS16:
            m = float(ms[f])
            k = rescale_shift(self.storage, m)
            if k:
                self._rescale(f, k)
                self.scales[f] = self.scales[f] * math.ldexp(1.0, -k)
                changed.append((f, k))
        self.rescales += len(changed)
        return changed
