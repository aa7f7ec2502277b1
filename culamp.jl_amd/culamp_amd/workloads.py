"""Seeded synthetic inputs shaped like the benchmark configurations.

* ``grid_histograms``  -- two g x g grid histograms on [0,1]^2 (z = 0): one
  Uniform(0,1), one a mixture of 3-8 isotropic Gaussian bumps; the grid
  coordinates follow the reference's ``--kernel`` layout
  ((k-1) div g, (k-1) mod g, 0) (src/CmdLineUtils.jl:208-213), scaled by
  1/(g-1).  Marginals get the reference's +1e-6 mixing and renormalisation
  (src/CmdLineUtils.jl:187-188).
* ``uniform_grid_pair`` -- two g x g histograms, entries Uniform(0,1)+1e-3.
* ``gmm_clouds``       -- two RGB-like point clouds in [0,1]^3 drawn from
  Gaussian mixtures, uniform weights.
* ``cell_batch``       -- independent cell-similarity instances: sizes
  log-uniform in [lo, hi], d-dimensional rows |N(0,1)| * Bernoulli(density),
  L2-normalised (all-zero rows redrawn), Dirichlet(1) weights.
* ``divergence_clouds`` -- N(0, I_3) source, 4-component Gaussian-mixture
  target, Uniform(0.5, 1.5) weights normalised.
"""
from __future__ import annotations

import numpy as np


def grid_coords(g: int, unit: bool = True) -> np.ndarray:
    k = np.arange(g * g)
    x = np.stack([k // g, k % g, np.zeros(g * g)], axis=1).astype(np.float64)
    return x / (g - 1) if unit else x


def _mix(a: np.ndarray) -> np.ndarray:
    a = a / a.sum()
    a = a + 1e-6
    return a / a.sum()


def bumps(rng: np.random.Generator, g: int) -> np.ndarray:
    k = int(rng.integers(3, 9))
    centres = rng.uniform(0, 1, size=(k, 2))
    sig = rng.uniform(0.03, 0.15, size=k)
    w = rng.dirichlet(np.ones(k))
    xy = grid_coords(g)[:, :2]
    h = np.zeros(g * g)
    for c, s, wi in zip(centres, sig, w):
        h += wi * np.exp(-np.sum((xy - c) ** 2, axis=1) / (2 * s * s))
    return h


def grid_histograms(g: int = 64, seed_a: int = 201, seed_b: int = 202):
    """x, y (g*g, 3) on [0,1]^2 and marginals r (uniform), c (bumps)."""
    x = grid_coords(g)
    r = _mix(np.random.default_rng(seed_a).uniform(0, 1, g * g))
    c = _mix(bumps(np.random.default_rng(seed_b), g))
    return x, x.copy(), r, c


def uniform_grid_pair(g: int = 32, seed_a: int = 101, seed_b: int = 102, unit: bool = True):
    x = grid_coords(g, unit)
    r = _mix(np.random.default_rng(seed_a).uniform(0, 1, g * g) + 1e-3)
    c = _mix(np.random.default_rng(seed_b).uniform(0, 1, g * g) + 1e-3)
    return x, x.copy(), r, c


def gmm_cloud(rng: np.random.Generator, n: int, k: int) -> np.ndarray:
    centres = rng.uniform(0.15, 0.85, size=(k, 3))
    sig = rng.uniform(0.03, 0.15, size=k)
    w = rng.dirichlet(np.ones(k))
    comp = rng.choice(k, size=n, p=w)
    pts = centres[comp] + rng.normal(size=(n, 3)) * sig[comp, None]
    return np.clip(pts, 0.0, 1.0)


def gmm_clouds(n: int = 16384, m: int = 16384, seed_x: int = 301, seed_y: int = 302):
    x = gmm_cloud(np.random.default_rng(seed_x), n, 5)
    y = gmm_cloud(np.random.default_rng(seed_y), m, 8)
    return x, y, np.full(n, 1.0 / n), np.full(m, 1.0 / m)


def sparse_unit_rows(rng: np.random.Generator, n: int, d: int, density: float) -> np.ndarray:
    x = np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < density)
    bad = ~(x.sum(axis=1) > 0)
    while bad.any():
        x[bad] = np.abs(rng.normal(size=(bad.sum(), d))) * (rng.uniform(size=(bad.sum(), d)) < density)
        bad = ~(x.sum(axis=1) > 0)
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def cell_batch(count: int = 256, lo: int = 256, hi: int = 4096, d: int = 128, density: float = 0.3,
               seed: int = 401):
    """List of (x (n, d), y (m, d), r, c) instances."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n, m = (int(np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(2))
        x = sparse_unit_rows(rng, n, d, density)
        y = sparse_unit_rows(rng, m, d, density)
        out.append((x, y, rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m))))
    return out


def divergence_clouds(n: int = 65536, m: int = 65536, seed_x: int = 501, seed_y: int = 502):
    rx, ry = np.random.default_rng(seed_x), np.random.default_rng(seed_y)
    x = rx.normal(size=(n, 3))
    centres = ry.normal(0, 2.0, size=(4, 3))
    sig = ry.uniform(0.2, 0.6, size=4)
    comp = ry.choice(4, size=m, p=ry.dirichlet(np.ones(4)))
    y = centres[comp] + ry.normal(size=(m, 3)) * sig[comp, None]
    a = rx.uniform(0.5, 1.5, n)
    b = ry.uniform(0.5, 1.5, m)
    return x, y, a / a.sum(), b / b.sum()
