"""Synthetic workloads: media, sources and receivers for the benchmark
configurations of the time-stepping path.

  elastic_halfspace_64   64^3, h=100 m, dt=5 ms, vp 6000, vs 3464, rho 2700,
                         elastic (Q^-1 = 0), isotropic moment-tensor Ricker
                         source (M0 = 1e15 N m, f0 = 2 Hz) at the centre,
                         200 steps, 8 seeded surface receivers.
  anelastic_512          512x512x256, h=50 m, dt=2.5 ms, 1000 steps, layered
                         vs 1500->4500 m/s, vp = 1.73 vs, Gardner rho, 5 %
                         von Karman perturbation (a = 500 m, Hurst 0.5),
                         Qs = 50 vs[km/s], Qp = 2 Qs, double-couple Ricker
                         source f0 = 1 Hz.
  topography_384         384x384x192, h=100 m, dt=5 ms, 1500 steps, the same
                         medium under a free surface of 6 seeded Gaussian
                         hills (height <= 1.5 km, widths 2-8 km).
Every random draw is seeded.  Sizes can be scaled down (``scale``) for parity
tests; the physics parameters stay the same.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import Dims
from .media import Medium, build_medium


def ricker(t: np.ndarray, f0: float, t0: float | None = None) -> np.ndarray:
    """Ricker wavelet with unit peak, centred at t0 (default 1.2/f0)."""
    if t0 is None:
        t0 = 1.2 / f0
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


@dataclass
class PointSources:
    """Point moment-rate sources in the reference's input convention
    (src/awp/source.c:160-200): cells are 1-based (x, y) and z counted 1 at
    the surface; rates are (npsrc, nst, 6) columns xx, yy, zz, xz, yz, xy."""
    cells: np.ndarray          # (npsrc, 3) int, model (x, y, z) 1-based
    rates: np.ndarray          # (npsrc, nst, 6) float64 [N m / s]

    @property
    def npsrc(self) -> int:
        return int(self.cells.shape[0])

    @property
    def nst(self) -> int:
        return int(self.rates.shape[1])

    def psrc(self, dims: Dims) -> np.ndarray:
        """tpsrc as pmcl3d stores it: (x, y, nzt+1-z) (source.c:179)."""
        p = self.cells.astype(np.int32).copy()
        p[:, 2] = dims.nzt + 1 - self.cells[:, 2]
        return p

    def columns(self) -> list[np.ndarray]:
        """axx, ayy, azz, axz, ayz, axy as flat [j*READ_STEP + n] arrays."""
        return [np.ascontiguousarray(self.rates[:, :, c]).reshape(-1) for c in range(6)]


@dataclass
class Receivers:
    """Surface receivers: model cells (x, y) 1-based; sampled with the
    velbuffer rule (u, v at the surface k, w one cell below)."""
    xy: np.ndarray  # (n, 2)

    def array_index(self, dims: Dims) -> np.ndarray:
        n = self.xy.shape[0]
        idx = np.empty((n, 3), dtype=np.int32)
        idx[:, 0] = self.xy[:, 0] + 1 + 8
        idx[:, 1] = self.xy[:, 1] + 1 + 8
        idx[:, 2] = dims.ktop
        return idx


@dataclass
class Workload:
    name: str
    dims: Dims
    DH: float
    DT: float
    nt: int
    medium_fn: object = None
    sources: PointSources | None = None
    receivers: Receivers | None = None
    params: dict = field(default_factory=dict)

    def medium(self) -> Medium:
        return self.medium_fn()


def seeded_surface_receivers(dims: Dims, n: int, seed: int, margin: int) -> Receivers:
    rng = np.random.default_rng(seed)
    lo, hx, hy = 1 + margin, dims.nxt - margin, dims.nyt - margin
    if hx < lo or hy < lo:
        lo, hx, hy = 1, dims.nxt, dims.nyt
    xs = rng.integers(lo, hx + 1, size=n)
    ys = rng.integers(lo, hy + 1, size=n)
    return Receivers(np.stack([xs, ys], axis=1))


def elastic_halfspace(n: int = 64, nt: int = 200, seed: int = 1234, nrec: int = 8,
                      sponge: bool = True) -> Workload:
    dims = Dims(n, n, n)
    DH, DT = 100.0, 0.005
    vp, vs, rho = 6000.0, 3464.0, 2700.0
    f0, M0 = 2.0, 1.0e15
    t = np.arange(nt) * DT
    rate = M0 * ricker(t, f0)
    rates = np.zeros((1, nt, 6))
    rates[0, :, 0] = rates[0, :, 1] = rates[0, :, 2] = rate
    c = n // 2
    src = PointSources(np.array([[c, c, c]]), rates)
    rec = seeded_surface_receivers(dims, nrec, seed, margin=min(20, n // 4))

    def mk():
        return build_medium(dims, DH, DT, vp, vs, rho, Qp=0.0, Qs=0.0, sponge=sponge)

    return Workload("elastic_halfspace_%d" % n, dims, DH, DT, nt, mk, src, rec,
                    dict(vp=vp, vs=vs, rho=rho, f0=f0, M0=M0))


def layered_vs(depth: np.ndarray, zmax: float, nlayers: int = 6, vmin: float = 1500.0,
               vmax: float = 4500.0) -> np.ndarray:
    """Piecewise-constant vs rising from vmin at the surface to vmax in
    nlayers equal-thickness layers."""
    layer = np.minimum((depth / zmax * nlayers).astype(int), nlayers - 1)
    return vmin + (vmax - vmin) * layer / (nlayers - 1)


def von_karman(shape: tuple[int, int, int], h: float, a: float, nu: float, std: float,
               seed: int) -> np.ndarray:
    """Seeded 3-D von Karman random field (PSD ~ (1 + k^2 a^2)^-(nu+3/2)) with
    the given standard deviation, float32."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal(shape, dtype=np.float32)
    F = np.fft.rfftn(white)
    kx = np.fft.fftfreq(shape[0], d=h) * 2 * np.pi
    ky = np.fft.fftfreq(shape[1], d=h) * 2 * np.pi
    kz = np.fft.rfftfreq(shape[2], d=h) * 2 * np.pi
    k2 = (kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = (1.0 + k2 * a * a) ** (-(nu + 1.5) / 2.0)
    F *= amp.astype(np.float32)
    del amp, k2
    f = np.fft.irfftn(F, s=shape, axes=(0, 1, 2)).astype(np.float32)
    f -= f.mean()
    f *= std / f.std()
    return f


def layered_random_model(dims: Dims, DH: float, seed: int = 7, std: float = 0.05,
                         corr: float = 500.0, nu: float = 0.5):
    """vp, vs, rho, Qp, Qs of the layered + von Karman medium (model order x, y,
    z-down).  Returns float64 arrays."""
    z = (np.arange(dims.nzt) + 0.5) * DH
    vs1d = layered_vs(z, dims.nzt * DH)
    pert = von_karman((dims.nxt, dims.nyt, dims.nzt), DH, corr, nu, std, seed)
    vs = vs1d[None, None, :] * (1.0 + pert.astype(np.float64))
    del pert
    vp = 1.73 * vs
    rho = 310.0 * vp ** 0.25
    Qs = 50.0 * vs / 1000.0
    Qp = 2.0 * Qs
    return vp, vs, rho, Qp, Qs


def anelastic(nx: int = 512, ny: int = 512, nz: int = 256, nt: int = 1000, DH: float = 50.0,
              DT: float = 0.0025, seed: int = 7, nrec: int = 8) -> Workload:
    dims = Dims(nx, ny, nz)
    f0, M0 = 1.0, 1.0e15
    t = np.arange(nt) * DT
    rates = np.zeros((1, nt, 6))
    rates[0, :, 5] = M0 * ricker(t, f0)  # xy: vertical strike-slip double couple
    src = PointSources(np.array([[nx // 2, ny // 2, nz // 2]]), rates)
    rec = seeded_surface_receivers(dims, nrec, seed + 1, margin=min(20, nx // 4))

    def mk():
        vp, vs, rho, Qp, Qs = layered_random_model(dims, DH, seed)
        return build_medium(dims, DH, DT, vp, vs, rho, Qp, Qs)

    return Workload("anelastic_%dx%dx%d" % (nx, ny, nz), dims, DH, DT, nt, mk, src, rec,
                    dict(f0=f0, M0=M0, seed=seed))


def topography_hills(nx: int = 384, ny: int = 384, nz: int = 192, nt: int = 1500,
                     DH: float = 100.0, DT: float = 0.005, seed: int = 7, nhills: int = 6,
                     hmax: float = 1500.0, wmin: float = 2000.0, wmax: float = 8000.0,
                     nrec: int = 16) -> Workload:
    """Curvilinear topography: the anelastic workload's layered + random
    medium with Q and double-couple Ricker source on a grid whose free surface
    is a seeded sum of nhills Gaussian hills (height <= hmax, widths in
    [wmin, wmax]); surface receivers.  The topography rides in
    params["topography"]."""
    from .topography import Topography
    w = anelastic(nx=nx, ny=ny, nz=nz, nt=nt, DH=DH, DT=DT, seed=seed, nrec=nrec)
    w.name = "topography_%dx%dx%d" % (nx, ny, nz)
    w.params["topography"] = Topography.gaussian_hills(nx, ny, DH, n=nhills, seed=seed + 3,
                                                       hmax=hmax, wmin=wmin, wmax=wmax)
    w.params.update(hmax=hmax, wmin=wmin, wmax=wmax, nhills=nhills)
    return w


def line_source(dims: Dims, nsub: int, nt: int, DT: float, f0: float, M0: float, seed: int,
                depth: int | None = None, x0: int | None = None) -> PointSources:
    """A line of nsub double-couple (xy) point subfaults along x at one depth,
    each with a seeded amplitude share of M0 and a Ricker moment rate."""
    rng = np.random.default_rng(seed)
    nsub = min(nsub, max(dims.nxt - 2, 1))
    x0 = x0 if x0 is not None else (dims.nxt - nsub) // 2 + 1
    z = depth if depth is not None else max(dims.nzt // 2, 1)
    y = dims.nyt // 2
    cells = np.array([[x0 + n, y, z] for n in range(nsub)])
    amp = rng.uniform(0.5, 1.5, nsub)
    amp *= M0 / amp.sum()
    t = np.arange(nt) * DT
    rates = np.zeros((nsub, nt, 6))
    for n in range(nsub):
        rates[n, :, 5] = amp[n] * ricker(t, f0)
    return PointSources(cells, rates)


def plastic(nx: int = 64, ny: int = 64, nz: int = 48, nt: int = 100, DH: float = 75.0,
            DT: float = 0.004, seed: int = 11, nsub: int = 200, M0: float = 1.0e19,
            cohesion: float = 5.0e6, friction: float = 0.75, Tv: float = 0.01,
            nrec: int = 8, f0: float = 1.0) -> Workload:
    """Layered + random medium with Q and Drucker-Prager plasticity
    (cohesion 5 MPa, friction 0.75, viscoplastic relaxation 0.01 s) driven by
    a line of double-couple subfaults."""
    from .plasticity import PlasticParams, constant_medium, protect_sources
    dims = Dims(nx, ny, nz)
    src = line_source(dims, nsub, nt, DT, f0, M0, seed)
    rec = seeded_surface_receivers(dims, nrec, seed + 1, margin=min(20, nx // 4))

    def mk():
        vp, vs, rho, Qp, Qs = layered_random_model(dims, DH, seed)
        return build_medium(dims, DH, DT, vp, vs, rho, Qp, Qs)

    def mk_plastic(med):
        pm = constant_medium(dims, DH, med.d1, cohesion, friction,
                             PlasticParams(fmajor=1.0, fminor=1.0, Tv=Tv))
        protect_sources(pm, dims, src.psrc(dims))
        return pm

    w = Workload("plastic_%dx%dx%d" % (nx, ny, nz), dims, DH, DT, nt, mk, src, rec,
                 dict(cohesion=cohesion, friction=friction, Tv=Tv, M0=M0))
    w.params["plastic_fn"] = mk_plastic
    return w


def layered_random_lattice(nx: int, ny: int, zmax: float, h: float, seed: int,
                           std: float = 0.05, corr: float = 500.0, nu: float = 0.5):
    """A seeded von Karman perturbation on a lattice of spacing h covering
    nx*h x ny*h x zmax, and the model function (x, y, depth) -> vp, vs, rho,
    Qp, Qs of the layered + random medium sampled from it (trilinear), for
    meshes whose grids have different spacings."""
    from .dm import sample_lattice
    nz = int(np.ceil(zmax / h))
    F = von_karman((nx, ny, nz), h, corr, nu, std, seed)

    def model(x, y, z):
        pert = sample_lattice(F, h, x, y, z).astype(np.float64)
        vs1d = layered_vs(z, zmax)
        vs = vs1d[None, None, :] * (1.0 + pert)
        vp = 1.73 * vs
        rho = 310.0 * vp ** 0.25
        Qs = 50.0 * vs / 1000.0
        return vp, vs, rho, 2.0 * Qs, Qs

    return model


def dm_plastic(nx: int = 256, ny: int = 256, nz: int = 192, nz_fine: int = 96,
               DH: float = 75.0, DT: float = 0.0025, nt: int = 3000, seed: int = 13,
               nsub: int = 200, M0: float = 4.0e19, cohesion: float = 5.0e6,
               friction: float = 0.75, Tv: float = 0.01, nrec: int = 8, f0: float = 1.0,
               ND: int = 20, src_depth: int | None = None, src_grid: int = 1) -> Workload:
    """Discontinuous mesh + plasticity: a fine top grid (3nx x 3ny x nz_fine,
    spacing DH/3) over the coarse grid (nx x ny x nz, spacing DH), layered +
    random medium with Q, Drucker-Prager plasticity (cohesion, friction,
    viscoplastic relaxation Tv) on both grids, and a line of nsub double-couple
    subfaults along x in grid src_grid (default the coarse grid) at model depth
    src_depth of that grid.  medium() returns the two media (fine first);
    sources is the per-grid list."""
    from .dm import build_stack, depth_offsets, stack_dims
    from .plasticity import PlasticParams, constant_medium, protect_sources
    dims = stack_dims(nx, ny, (nz_fine, nz))
    DHs = [DH / 3.0, DH]
    zmax = depth_offsets(dims, DHs)[1] + nz * DH
    zsrc = src_depth if src_depth is not None else max(dims[src_grid].nzt // 2, 1)
    sources = [None, None]
    sources[src_grid] = line_source(dims[src_grid], nsub, nt, DT, f0, M0, seed, depth=zsrc)
    rec = seeded_surface_receivers(dims[0], nrec, seed + 1, margin=min(20 * 3, dims[0].nxt // 4))

    def mk():
        model = layered_random_lattice(nx, ny, zmax, DH, seed)
        return build_stack(dims, DHs[0], DT, model, ND=ND)

    def mk_plastic(media):
        out = []
        for p, med in enumerate(media):
            pm = constant_medium(med.dims, med.DH, med.d1, cohesion, friction,
                                 PlasticParams(fmajor=1.0, fminor=1.0, Tv=Tv))
            if sources[p] is not None:
                protect_sources(pm, med.dims, sources[p].psrc(med.dims))
            out.append(pm)
        return out

    w = Workload("dm_plastic_%dx%dx%d_%dx%dx%d" % (dims[0].nxt, dims[0].nyt, nz_fine, nx, ny, nz),
                 dims[0], DHs[0], DT, nt, mk, sources, rec,
                 dict(cohesion=cohesion, friction=friction, Tv=Tv, M0=M0, dims=dims, DH=DHs))
    w.params["plastic_fn"] = mk_plastic
    return w


def adjoint_pair(nx: int = 256, ny: int = 256, nz: int = 128, nt: int = 2000, DH: float = 100.0,
                 DT: float = 0.005, nsrc: int = 4, nrec_side: int = 8, seed: int = 17,
                 anomaly: float = 0.10, f0: float = 1.0, M0: float = 1.0e15):
    """The adjoint-gradient workload: a smooth background (vs rising linearly
    with depth 1500 -> 4500 m/s, vp = 1.73 vs, Gardner rho, Qs = 50 vs[km/s],
    Qp = 2 Qs) and the same model with a Gaussian vs anomaly of the given
    relative size at the centre (width a sixth of the box); nsrc seeded
    explosive (isotropic) Ricker sources at mid depth; a nrec_side^2 grid of
    surface receivers.  Returns (background medium, true medium, sources,
    receivers)."""
    dims = Dims(nx, ny, nz)
    rng = np.random.default_rng(seed)
    z = (np.arange(nz) + 0.5) * DH
    vs1 = 1500.0 + 3000.0 * z / (nz * DH)
    x = (np.arange(nx) + 0.5) * DH
    y = (np.arange(ny) + 0.5) * DH
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    cx, cy, cz = nx * DH / 2, ny * DH / 2, nz * DH / 2
    w = min(nx, ny, nz) * DH / 6.0
    gauss = np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2 * w * w))

    def medium(vs):
        vp = 1.73 * vs
        rho = 310.0 * vp ** 0.25
        Qs = 50.0 * vs / 1000.0
        return build_medium(dims, DH, DT, vp, vs, rho, 2.0 * Qs, Qs)

    vs_bg = np.broadcast_to(vs1[None, None, :], (nx, ny, nz))
    bg = medium(vs_bg)
    true = medium(vs_bg * (1.0 + anomaly * gauss))
    t = np.arange(nt) * DT
    cells = np.stack([rng.integers(nx // 4, 3 * nx // 4 + 1, nsrc),
                      rng.integers(ny // 4, 3 * ny // 4 + 1, nsrc),
                      rng.integers(nz // 3, 2 * nz // 3 + 1, nsrc)], axis=1)
    rates = np.zeros((nsrc, nt, 6))
    for n in range(nsrc):
        rates[n, :, 0] = rates[n, :, 1] = rates[n, :, 2] = M0 * ricker(t, f0, t0=1.2 / f0 + 0.05 * n)
    src = PointSources(cells, rates)
    gx = np.linspace(max(1, nx // 10), nx - max(1, nx // 10), nrec_side).round().astype(int)
    gy = np.linspace(max(1, ny // 10), ny - max(1, ny // 10), nrec_side).round().astype(int)
    rec = Receivers(np.array([[a, b] for a in gx for b in gy]))
    return bg, true, src, rec
