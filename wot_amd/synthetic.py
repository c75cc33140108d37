"""Seeded synthetic single-cell time courses in PCA coordinates.

The workloads the solver is measured on are Gaussian mixtures in a d-dimensional
PCA space.  Source and target days share mixture components; the target
means are shifted (one day pair) or drift linearly with time (a time course).
Growth rates are log-normal.  Everything is float64 and fully determined by the
seed, so the GPU path and the CPU oracle see identical inputs.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class DayPair:
    x: np.ndarray          # (I, d) source coordinates, float64
    y: np.ndarray          # (J, d) target coordinates, float64
    growth: np.ndarray     # (I,) growth rate per source cell, float64
    params: dict = field(default_factory=dict)


def _mixture(rng, n, means, sigma):
    k = means.shape[0]
    comp = rng.integers(0, k, size=n)
    return means[comp] + sigma * rng.standard_normal((n, means.shape[1]))


def day_pair(n_source=2000, n_target=2500, dim=30, n_components=5, mean_scale=3.0,
             sigma=1.0, shift=0.5, growth_sigma=0.3, seed=0):
    """One day pair: equal-weight mixture, means ~ N(0, mean_scale^2 I), sigma
    per component; target means = source means + N(0, shift^2 I); growth ~
    LogNormal(0, growth_sigma)."""
    rng = np.random.default_rng(seed)
    mu = mean_scale * rng.standard_normal((n_components, dim))
    mu_t = mu + shift * rng.standard_normal((n_components, dim))
    x = _mixture(rng, n_source, mu, sigma)
    y = _mixture(rng, n_target, mu_t, sigma)
    g = np.exp(growth_sigma * rng.standard_normal(n_source))
    return DayPair(x=x, y=y, growth=g,
                   params=dict(n_source=n_source, n_target=n_target, dim=dim,
                               n_components=n_components, mean_scale=mean_scale,
                               sigma=sigma, shift=shift, growth_sigma=growth_sigma, seed=seed))


def time_course(n_days=39, min_cells=1000, max_cells=20000, dim=30, n_components=20,
                mean_scale=3.0, drift=0.25, sigma=1.0, growth_sigma=0.3, seed=0, counts=None):
    """A time course of ``n_days`` days (``n_days - 1`` consecutive pairs).
    Cells per day ~ U{min_cells..max_cells} (or the given ``counts``);
    component means drift linearly
    (mu_k(t) = mu_k + t * drift * w_k, w_k ~ N(0, I)).  Returns (coords per
    day, growth per day, cell counts)."""
    rng = np.random.default_rng(seed)
    mu0 = mean_scale * rng.standard_normal((n_components, dim))
    w = rng.standard_normal((n_components, dim))
    drawn = rng.integers(min_cells, max_cells + 1, size=n_days)
    counts = np.asarray(counts if counts is not None else drawn)
    n_days = len(counts)
    coords, growth = [], []
    for t in range(n_days):
        mu_t = mu0 + t * drift * w
        coords.append(_mixture(rng, int(counts[t]), mu_t, sigma))
        growth.append(np.exp(growth_sigma * rng.standard_normal(int(counts[t]))))
    return coords, growth, counts
