"""Synthetic ROMS-style forcing windows and releases for tests and benchmarks.

No NetCDF is involved: fields are generated in memory on the normalised rho
grid of OceanModelAdapter (OceanModelAdapter.hpp:20-34) and stored as FP32,
the reference's forcing storage type.

plain_window():  120(eta) x 100(xi) x 10 sigma levels at 200 m, no land,
                 uniform u=0.2, v=0.1, w=0, AKt=0, h=50 m, hourly records.
full_window():   1024 x 768 x 11 sigma levels at 150 m; u, v = sum of 64
                 random plane-wave modes (wavelengths 5-40 km, rms speed
                 0.3 m/s, seeded phases drifting in time) with vertical shape
                 exp(z/30 m); w from the continuity equation integrated from
                 the bottom; land = Gaussian-smoothed white noise above its
                 70th percentile (mask=0, h=0, u=v=w=0 there, as
                 ROMSAdapter.cpp:244-264 leaves land); h=40 m elsewhere;
                 AKt = 1e-4 m2/s in water; hourly records.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

R_EARTH = 6371000.0


@dataclass
class ForcingWindow:
    """A window of forcing records on a rectilinear lon/lat grid."""
    s_rho: int
    eta: int
    xi: int
    lonRad: np.ndarray
    latRad: np.ndarray
    h: np.ndarray
    mask: np.ndarray
    depthIntervals: np.ndarray       # s_w-1 values, entry kI+s_w-2
    times: np.ndarray                # record times [T] (s)
    record_fn: object = None         # t -> (zeta, u, v, w, akt) FP32
    _cache: dict = field(default_factory=dict)

    @property
    def s_w(self):
        return self.s_rho + 1

    @property
    def T(self):
        return len(self.times)

    def record(self, t: int):
        if t not in self._cache:
            self._cache[t] = self.record_fn(t)
        return self._cache[t]

    def drop(self, t: int):
        self._cache.pop(t, None)

    def stacked(self):
        """All records as [T]-leading arrays (zeta, u, v, w, akt)."""
        recs = [self.record(t) for t in range(self.T)]
        return tuple(np.stack([r[q] for r in recs]) for q in range(5))

    def interval(self, ocean_time_idx: int, direction: int, deltat: float = 3600.0):
        """(start, end, record pair) as Particle.cpp:110-115, 183-184."""
        nxt = ocean_time_idx + direction
        start = float(self.times[ocean_time_idx])
        end = float(self.times[nxt]) if 0 <= nxt < self.T else start + direction * deltat
        if nxt < 0 or nxt >= self.T:
            nxt = ocean_time_idx
        return start, end, ocean_time_idx, nxt


def rectilinear(eta, xi, spacing_m, lat0_deg=40.75, lon0_deg=14.0):
    dlat = spacing_m / R_EARTH
    dlon = spacing_m / (R_EARTH * math.cos(math.radians(lat0_deg)))
    jj, ii = np.meshgrid(np.arange(eta), np.arange(xi), indexing="ij")
    lat = math.radians(lat0_deg) + jj * dlat
    lon = math.radians(lon0_deg) + ii * dlon
    return lon, lat


def depth_to_k(depth_m, h, s_w):
    """deplatlon2kji's vertical inversion (OceanModelAdapter.cpp:425-441) for
    uniform sigma_w(k) = k/(s_w-1)."""
    sw = lambda k: k / (s_w - 1)   # noqa: E731
    k_bottom = -s_w + 2
    if depth_m <= h * abs(sw(0)):
        return 0.0
    if depth_m >= h * abs(sw(k_bottom)):
        return float(k_bottom)
    kI = 0
    while kI - 1 > k_bottom and depth_m >= h * abs(sw(kI - 1)):
        kI -= 1
    upper = h * abs(sw(kI))
    a = abs(h * abs(sw(kI - 1)) - upper)
    return kI - (depth_m - upper) / a


def plain_window(T=25, eta=120, xi=100, s_rho=10, spacing=200.0, u=0.2, v=0.1, h=50.0) -> ForcingWindow:
    lon, lat = rectilinear(eta, xi, spacing)
    s_w = s_rho + 1

    def rec(t):
        return (np.zeros((eta, xi), np.float32), np.full((s_rho, eta, xi), u, np.float32),
                np.full((s_rho, eta, xi), v, np.float32), np.zeros((s_w, eta, xi), np.float32),
                np.zeros((s_w, eta, xi), np.float32))

    return ForcingWindow(s_rho, eta, xi, lon, lat, np.full((eta, xi), h), np.ones((eta, xi)),
                         np.full(s_w - 1, 1.0 / (s_w - 1)), np.arange(T) * 3600.0, rec)


def _gaussian_smooth(a, sigma):
    """Separable Gaussian filter with wrap-around, via FFT (no scipy needed)."""
    ny, nx = a.shape
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    g = np.exp(-2.0 * (math.pi * sigma) ** 2 * (kx ** 2 + ky ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(a) * g))


def full_window(T=25, eta=1024, xi=768, s_rho=11, spacing=150.0, seed=5489, n_modes=64, rms=0.3,
                land_fraction=0.30, h_ocean=40.0, H_decay=30.0, akt=1e-4) -> ForcingWindow:
    rng = np.random.default_rng(seed)
    lon, lat = rectilinear(eta, xi, spacing)
    s_w = s_rho + 1
    # land: smoothed white noise above its (1-land_fraction) quantile
    noise = _gaussian_smooth(rng.standard_normal((eta, xi)), 8.0)
    land = noise > np.quantile(noise, 1.0 - land_fraction)
    mask = np.where(land, 0.0, 1.0)
    h = np.where(land, 0.0, h_ocean)
    # modes: wavelength log-uniform in [5, 40] km, random direction and
    # polarisation, phase drifting with a period of 12-48 h
    lam = np.exp(rng.uniform(math.log(5e3), math.log(40e3), n_modes))
    theta = rng.uniform(0, 2 * math.pi, n_modes)
    kx = 2 * math.pi / lam * np.cos(theta)
    ky = 2 * math.pi / lam * np.sin(theta)
    beta = rng.uniform(0, 2 * math.pi, n_modes)
    phi0 = rng.uniform(0, 2 * math.pi, n_modes)
    omega = 2 * math.pi / rng.uniform(12 * 3600.0, 48 * 3600.0, n_modes) * rng.choice([-1, 1], n_modes)
    amp = math.sqrt(2.0 * rms ** 2 / n_modes)          # E[u^2 + v^2] = sum amp^2 / 2 = rms^2
    x = np.arange(xi) * spacing
    y = np.arange(eta) * spacing
    # vertical structure on rho levels (sigma_rho(k) = (2k-1)/(2 s_rho)) and W levels
    k_rho = np.arange(-s_rho + 1, 1)
    k_w = np.arange(-s_w + 1, 1)
    z_rho = (2 * k_rho - 1) / (2 * s_rho) * h_ocean
    z_w = k_w / (s_w - 1) * h_ocean
    shape_rho = np.exp(z_rho / H_decay)
    # w(z) = -div_h(U) * integral_{-h}^{z} exp(z'/H) dz'
    shape_w = H_decay * (np.exp(z_w / H_decay) - math.exp(-h_ocean / H_decay))
    water = (~land).astype(np.float32)
    cx, sx = np.cos(np.outer(x, kx)), np.sin(np.outer(x, kx))      # [xi, M]

    def rec(t):
        ph = np.outer(y, ky) + phi0 + omega * (t * 3600.0)            # [eta, M]
        cy, sy = np.cos(ph), np.sin(ph)
        au, av = amp * np.cos(beta), amp * np.sin(beta)
        # cos(kx x + ky y + phi) = cy cx - sy sx ; sin(...) = sy cx + cy sx
        u2 = (cy * au) @ cx.T - (sy * au) @ sx.T
        v2 = (cy * av) @ cx.T - (sy * av) @ sx.T
        dcoef = -(au * kx + av * ky)                                   # d/dx, d/dy of cos -> -k sin
        div = (sy * dcoef) @ cx.T + (cy * dcoef) @ sx.T
        u = (shape_rho[:, None, None] * (u2 * water)[None]).astype(np.float32)
        v = (shape_rho[:, None, None] * (v2 * water)[None]).astype(np.float32)
        w = (-shape_w[:, None, None] * (div * water)[None]).astype(np.float32)
        a = np.full((s_w, eta, xi), akt, np.float32) * water[None]
        return np.zeros((eta, xi), np.float32), u, v, w, a.astype(np.float32)

    return ForcingWindow(s_rho, eta, xi, lon, lat, h, mask, np.full(s_w - 1, 1.0 / (s_w - 1)),
                         np.arange(T) * 3600.0, rec)


def point_sources(window: ForcingWindow, n_sources: int, seed: int = 5489, margin: int = 8):
    """Uniform-random release points inside cells whose four corners are water."""
    rng = np.random.default_rng(seed + 1)
    wet = window.mask > 0
    ok = wet[:-1, :-1] & wet[1:, :-1] & wet[:-1, 1:] & wet[1:, 1:]
    ok[:margin, :] = False
    ok[-margin:, :] = False
    ok[:, :margin] = False
    ok[:, -margin:] = False
    cells = np.argwhere(ok)
    pick = cells[rng.choice(len(cells), n_sources, replace=False)]
    off = rng.uniform(0.05, 0.95, (n_sources, 2))
    return pick[:, 0] + off[:, 0], pick[:, 1] + off[:, 1]


def release(window: ForcingWindow, n_particles: int, n_sources: int, depth_m: float = 1.0, time: float = 0.0,
            seed: int = 5489, spread: float = 0.0):
    """A single pulse of n_particles split evenly over n_sources point
    sources (ids 0..n-1 in source order, Source.cpp:25-44 numbering).
    spread > 0 adds seeded N(0, spread^2) offsets to j and i of each particle,
    as randomSources does with 0.25 (SourceEmission.hpp:20-23)."""
    from .types import make_particles
    js, is_ = point_sources(window, n_sources, seed)
    per = np.full(n_sources, n_particles // n_sources)
    per[: n_particles % n_sources] += 1
    p = make_particles(n_particles, time=time)
    h_src = float(window.h[window.mask > 0].mean())
    p["k"] = depth_to_k(depth_m, h_src, window.s_w)
    p["j"] = np.repeat(js, per)
    p["i"] = np.repeat(is_, per)
    if spread > 0:
        rng = np.random.default_rng(seed + 2)
        p["j"] += spread * rng.standard_normal(n_particles)
        p["i"] += spread * rng.standard_normal(n_particles)
    return p


def ensemble(window: ForcingWindow, n_scenarios: int = 256, total: int = 16_000_000, median: float = 40_000,
             lo: int = 10_000, hi: int = 500_000, kh_range=(0.1, 10.0), max_offset_h: float = 6.0,
             dti: float = 30.0, depth_m: float = 1.0, seed: int = 5489, spread: float = 0.25):
    """The batched-ensemble workload: n_scenarios release scenarios on one
    grid and forcing.  Particle counts are lognormal (median `median`,
    sigma 1) clipped to [lo, hi] and rescaled so they sum to about `total`;
    each scenario has its own ocean source cell, Kh log-uniform in kh_range
    (sigma = sqrt(2 Kh dti)) and a release offset uniform in [0,
    max_offset_h] hours after the window start.  Particles of a scenario are
    contiguous, ids 0..N-1, each offset by N(0, spread^2) cells.
    Returns (particles, scenario uint8, sigma per scenario, counts)."""
    from .types import make_particles
    rng = np.random.default_rng(seed + 7)
    counts = np.clip(rng.lognormal(math.log(median), 1.0, n_scenarios), lo, hi)
    counts = np.clip(np.round(counts * total / counts.sum()), lo, hi).astype(np.int64)
    kh = np.exp(rng.uniform(math.log(kh_range[0]), math.log(kh_range[1]), n_scenarios))
    sigma = np.sqrt(2.0 * kh * dti)
    offset = rng.uniform(0.0, max_offset_h * 3600.0, n_scenarios)
    js, is_ = point_sources(window, n_scenarios, seed + 11)
    n = int(counts.sum())
    p = make_particles(n, time=0.0)
    scen = np.repeat(np.arange(n_scenarios, dtype=np.uint8), counts)
    h_src = float(window.h[window.mask > 0].mean())
    p["k"] = depth_to_k(depth_m, h_src, window.s_w)
    p["j"] = np.repeat(js, counts) + spread * rng.standard_normal(n)
    p["i"] = np.repeat(is_, counts) + spread * rng.standard_normal(n)
    p["time"] = float(window.times[0]) + np.repeat(offset, counts)
    return p, scen, sigma, counts
