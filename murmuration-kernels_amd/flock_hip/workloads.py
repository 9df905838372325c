"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

plain   one flock of n birds, uniform cube at 0.64 birds per 4 m cell
mix     independent flocks of log-uniform sizes, each at the same density
stream  a Poisson stream of flock-simulation requests
clumps  Gaussian clumps with aligned headings
"""
from __future__ import annotations

import numpy as np

from . import cube_side_for, gen_clumps, gen_cube, gen_request


def plain(n: int, seed: int = 1) -> np.ndarray:
    return gen_cube(n, cube_side_for(n), seed)


def mix_sizes(flocks: int = 256, lo: int = 1000, hi: int = 200_000, seed: int = 1) -> np.ndarray:
    """Flock sizes round(exp(U(ln lo, ln hi)))."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), flocks))).astype(np.int64)


def mix(flocks: int = 256, lo: int = 1000, hi: int = 200_000, seed: int = 1) -> list[np.ndarray]:
    """Start states of the mixed workload: flock f = gen_request(n_f, seed * 1000 + f)."""
    return [gen_request(int(n), seed * 1000 + f) for f, n in enumerate(mix_sizes(flocks, lo, hi, seed))]


def stream(requests: int = 2000, rate: float = 100.0, lo: int = 2000, hi: int = 500_000, steps_lo: int = 50,
           steps_hi: int = 500, seed: int = 1) -> dict:
    """Poisson arrivals at `rate` requests/s; birds log-uniform [lo, hi];
    steps uniform {steps_lo..steps_hi}; one seed per request."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
    birds = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), requests))).astype(np.uint32)
    steps = rng.integers(steps_lo, steps_hi + 1, requests).astype(np.uint32)
    seeds = (np.arange(requests, dtype=np.uint64) + np.uint64(seed) * np.uint64(1_000_003))
    return {"arrival_s": arrivals, "birds": birds, "steps": steps, "seeds": seeds}


def clumps(n: int = 10_000_000, k: int = 8, sigma: float = 5.0, cube: float = 1000.0, seed: int = 1) -> np.ndarray:
    return gen_clumps(n, k, sigma, cube, seed)
