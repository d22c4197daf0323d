"""Seeded synthetic columns for the benchmark workloads and the tests.

The reference generates test data with cuRAND / rand() seeded by time(NULL)
(util/generator/cuda_array_generator.cpp:15, :110, :146), so its inputs are
not reproducible; these generators are counter-seeded numpy streams.
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def uniform_int(n: int, hi: int = 1024, seed: int = 0) -> np.ndarray:
    """int32 uniform in [0, hi) (the 1M-row FL round-trip column)."""
    return rng(seed).integers(0, hi, n, dtype=np.int32)


def timestamps(n: int, seed: int = 1, start: int = 1_400_000_000, p: float = 0.3) -> np.ndarray:
    """int32 timestamps: start + cumulative geometric(p) increments (>= 1)."""
    inc = rng(seed).geometric(p, n).astype(np.int64)
    inc[0] = 0
    return (start + np.cumsum(inc)).astype(np.int32)


def step_function(n: int, seed: int = 2, mean_run: float = 200.0, hi: int = 1 << 20) -> np.ndarray:
    """int32 step function: run lengths geometric with the given mean, values uniform in [0, hi)."""
    g = rng(seed)
    runs = []
    total = 0
    while total < n:
        k = max(1024, int((n - total) / mean_run * 1.2) + 16)
        lens = g.geometric(1.0 / mean_run, k)
        runs.append(lens)
        total += int(lens.sum())
    lens = np.concatenate(runs)
    ends = np.cumsum(lens)
    m = int(np.searchsorted(ends, n, side="left")) + 1
    lens = lens[:m]
    vals = g.integers(0, hi, m, dtype=np.int32)
    return np.repeat(vals, lens)[:n].astype(np.int32)


def zipf_codes(n: int, seed: int = 3, s: float = 1.2, codes: int = 50) -> np.ndarray:
    """int32 codes 0..codes-1 with P(k) proportional to (k+1)^-s."""
    w = 1.0 / np.arange(1, codes + 1, dtype=np.float64) ** s
    w /= w.sum()
    return rng(seed).choice(codes, size=n, p=w).astype(np.int32)


def normal_with_outliers(n: int, seed: int = 4, mu: float = 1000.0, sigma: float = 50.0,
                         frac: float = 0.01) -> np.ndarray:
    """int32 round(N(mu, sigma)) with `frac` of the rows replaced by uniform [0, 2^31)."""
    g = rng(seed)
    x = np.rint(g.normal(mu, sigma, n)).astype(np.int32)
    k = int(round(n * frac))
    pos = g.choice(n, size=k, replace=False) if k else np.array([], dtype=np.int64)
    x[pos] = g.integers(0, 2**31 - 1, k, dtype=np.int64).astype(np.int32)
    return x


def price_walk_f64(n: int, seed: int = 5, start_cents: int = 10000) -> np.ndarray:
    """float64 random walk from 100.00 with N(0, 0.01) steps, kept at 2 decimals
    (integer cents divided by 100)."""
    steps = np.rint(rng(seed).normal(0.0, 1.0, n)).astype(np.int64)
    steps[0] = 0
    cents = start_cents + np.cumsum(steps)
    return cents.astype(np.float64) / 100.0


def price_walk_f32(n: int, seed: int = 6) -> np.ndarray:
    return price_walk_f64(n, seed).astype(np.float32)
