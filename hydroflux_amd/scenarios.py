"""Synthetic inputs for the rain-on-grid flood workloads.

Construction (seeded, deterministic):

- bed: fractional-Brownian-motion surface by spectral synthesis — complex
  Gaussian coefficients with amplitude |k|^-(H+1), H = 0.8, inverse real FFT
  (Hermitian symmetry by construction), rescaled to `relief` metres
  peak-to-peak; plus a planar slope of 0.01 descending toward +x (the
  transmissive outflow edge) and a Gaussian-section valley `valley_depth`
  metres deep (sigma = `valley_sigma` m) carved along x at mid-domain; plus a
  constant `offset` (e.g. +2000 m for Andean elevations).
- Manning n: nearest of 64 seeded Voronoi sites, each assigned one of the
  land-cover classes {0.03, 0.05, 0.08, 0.12}.
- initial state: dry. Forcing: uniform rain `rain_mm_h`, a point inflow of
  `inflow` m^3/s at the upstream valley cell (row n_rows/2, column 0).
- boundaries: N, S, W walls, E transmissive.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

LAND_COVER_N = (0.03, 0.05, 0.08, 0.12)
SEED_DEM = 0x48464401
SEED_LAND = 0x48464402


@dataclass
class FloodScenario:
    n_rows: int
    n_cols: int
    dx: float
    bed: np.ndarray
    manning: np.ndarray
    land_class: np.ndarray
    rain_rate: float                       # m/s
    sources: list = field(default_factory=list)
    bc_kinds: tuple = (1, 1, 1, 0)         # N, S, W, E
    t_end: float = 3600.0
    max_dt: float = 1.0
    cfl: float = 0.4

    @property
    def n_cells(self) -> int:
        return self.n_rows * self.n_cols

    def dry_state(self) -> np.ndarray:
        return np.zeros((self.n_rows, self.n_cols, 3))


def fbm_surface(n_rows, n_cols, hurst=0.8, relief=50.0, seed=SEED_DEM):
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(n_rows)[:, None]
    kx = np.fft.rfftfreq(n_cols)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf
    amp = k ** (-(hurst + 1.0))
    shape = amp.shape
    coef = amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    z = np.fft.irfft2(coef, s=(n_rows, n_cols))
    z = (z - z.min()) / (z.max() - z.min()) * relief
    return z


def voronoi_classes(n_rows, n_cols, n_sites=64, seed=SEED_LAND):
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    sites = rng.uniform([0, 0], [n_rows, n_cols], size=(n_sites, 2))
    cls = rng.integers(0, len(LAND_COVER_N), size=n_sites)
    tree = cKDTree(sites)
    workers = -1
    out = np.empty((n_rows, n_cols), np.int8)
    rows = np.arange(n_rows) + 0.5
    cols = np.arange(n_cols) + 0.5
    step = max(1, (1 << 22) // n_cols)
    for r0 in range(0, n_rows, step):
        rr = rows[r0:r0 + step]
        pts = np.stack(np.meshgrid(rr, cols, indexing="ij"), -1).reshape(-1, 2)
        _, idx = tree.query(pts, workers=workers)
        out[r0:r0 + step] = cls[idx].reshape(len(rr), n_cols)
    return out


def flood_scenario(n=1024, dx=5.0, offset=0.0, rain_mm_h=50.0, inflow=100.0, valley_depth=8.0,
                   valley_sigma=100.0, seed_dem=SEED_DEM, seed_land=SEED_LAND, n_cols=None,
                   relief=50.0) -> FloodScenario:
    n_rows, n_cols = int(n), int(n_cols or n)
    z = fbm_surface(n_rows, n_cols, relief=relief, seed=seed_dem)
    x = (np.arange(n_cols) + 0.5) * dx
    y = (np.arange(n_rows) + 0.5) * dx
    z = z + 0.01 * (x[-1] - x)[None, :]
    z = z - valley_depth * np.exp(-((y - y[n_rows // 2]) ** 2) / (2.0 * valley_sigma ** 2))[:, None]
    z = np.ascontiguousarray(z + offset)
    land = voronoi_classes(n_rows, n_cols, seed=seed_land)
    man = np.asarray(LAND_COVER_N)[land]
    return FloodScenario(n_rows, n_cols, dx, z, np.ascontiguousarray(man), land,
                         rain_rate=rain_mm_h * 1e-3 / 3600.0,
                         sources=[(n_rows // 2, 0, inflow)] if inflow else [])


def scenario_mix(n=256, seed=0x48464403, min_side=128, max_side=2048, rain_range=(10.0, 100.0),
                 duration_range=None, dx=5.0):
    """The batching workload: `n` scenarios, long side log-uniform in
    [min_side, max_side], aspect ratio U(1, 4) (short side = max(min_side,
    long / aspect)), each with its own seeded DEM and land cover, rain
    U(10, 100) mm/h, an inflow at its upstream valley cell, and (optionally)
    a duration U(duration_range) simulated seconds."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        s = int(round(np.exp(rng.uniform(np.log(min_side), np.log(max_side)))))
        a = rng.uniform(1.0, 4.0)
        short = max(min_side, int(round(s / a)))
        rows, cols = (short, s) if rng.random() < 0.5 else (s, short)
        rain = rng.uniform(*rain_range)
        sc = flood_scenario(n=rows, n_cols=cols, dx=dx, rain_mm_h=rain, inflow=100.0,
                            seed_dem=seed * 1000 + 2 * k, seed_land=seed * 1000 + 2 * k + 1)
        if duration_range is not None:
            sc.t_end = float(rng.uniform(*duration_range))
        out.append(sc)
    return out
