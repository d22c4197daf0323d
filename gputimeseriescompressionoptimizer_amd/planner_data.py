"""Synthetic int32 blocks for the learned planner: the generator families of
the benchmark columns with randomised parameters, plus a few shapes they do
not cover (uniform bit widths, sorted sequences, near-constant columns,
periodic signals).  A stream draws its family and parameters once per regime
and changes regime every `regime` blocks."""
from __future__ import annotations

from typing import Callable, Iterator, List, Tuple

import numpy as np

FAMILIES = ["timestamps", "step", "zipf", "outliers", "uniform", "sorted", "near_constant", "periodic", "walk"]


def _params(family: str, r: np.random.Generator) -> dict:
    if family == "timestamps":
        return {"p": float(r.uniform(0.02, 0.95))}
    if family == "step":
        return {"mean_run": float(np.exp(r.uniform(np.log(2), np.log(5000)))), "bits": int(r.integers(1, 31))}
    if family == "zipf":
        return {"s": float(r.uniform(1.05, 2.5)), "codes": int(np.exp(r.uniform(np.log(2), np.log(20000))))}
    if family == "outliers":
        return {"sigma": float(np.exp(r.uniform(0, np.log(1e6)))), "frac": float(r.uniform(0, 0.05)),
                "mu": float(r.uniform(0, 1e6))}
    if family == "uniform":
        return {"bits": int(r.integers(1, 32))}
    if family == "sorted":
        return {"step_bits": int(r.integers(0, 12))}
    if family == "near_constant":
        return {"frac": float(r.uniform(0, 0.2)), "bits": int(r.integers(1, 31))}
    if family == "periodic":
        return {"period": float(np.exp(r.uniform(np.log(4), np.log(5000)))), "amp": float(np.exp(r.uniform(0, 20))),
                "noise": float(r.uniform(0, 3))}
    return {"step": int(r.integers(1, 1000))}


def block(family: str, p: dict, n: int, r: np.random.Generator) -> np.ndarray:
    """One int32 block of `n` rows."""
    if family == "timestamps":
        inc = r.geometric(p["p"], n).astype(np.int64)
        x = int(r.integers(1_000_000_000, 1_500_000_000)) + np.cumsum(inc)
    elif family == "step":
        runs = r.geometric(1.0 / p["mean_run"], n // 2 + 2)
        vals = r.integers(0, 1 << p["bits"], runs.size)
        x = np.repeat(vals, runs)[:n]
        if x.size < n:
            x = np.resize(x, n)
    elif family == "zipf":
        k = np.arange(1, p["codes"] + 1, dtype=np.float64)
        w = k ** -p["s"]
        x = r.choice(p["codes"], n, p=w / w.sum())
    elif family == "outliers":
        x = np.rint(r.normal(p["mu"], p["sigma"], n))
        m = r.random(n) < p["frac"]
        x[m] = r.integers(0, 1 << 31, int(m.sum()))
    elif family == "uniform":
        x = r.integers(0, 1 << p["bits"], n)
    elif family == "sorted":
        x = np.cumsum(r.integers(0, (1 << p["step_bits"]) + 1, n))
    elif family == "near_constant":
        x = np.full(n, int(r.integers(0, 1 << p["bits"])), np.int64)
        m = r.random(n) < p["frac"]
        x[m] = r.integers(0, 1 << p["bits"], int(m.sum()))
    elif family == "periodic":
        t = np.arange(n)
        x = np.rint(p["amp"] * np.sin(2 * np.pi * t / p["period"]) + r.normal(0, p["noise"], n))
    else:
        x = np.cumsum(r.integers(-p["step"], p["step"] + 1, n))
    return np.clip(x, -(1 << 31), (1 << 31) - 1).astype(np.int32)


def random_blocks(count: int, n: int, seed: int) -> Iterator[Tuple[str, np.ndarray]]:
    """Independent blocks, each with its own family and parameters."""
    r = np.random.default_rng(seed)
    for _ in range(count):
        fam = FAMILIES[int(r.integers(0, len(FAMILIES)))]
        yield fam, block(fam, _params(fam, r), n, r)


def regime_stream(count: int, n: int, regime: int, seed: int) -> Iterator[Tuple[int, str, np.ndarray]]:
    """A stream whose family and parameters change every `regime` blocks."""
    r = np.random.default_rng(seed)
    fam, p = None, None
    for i in range(count):
        if i % regime == 0:
            fam = FAMILIES[int(r.integers(0, len(FAMILIES)))]
            p = _params(fam, r)
        yield i // regime, fam, block(fam, p, n, r)


# ----------------------------------------------------------- float columns
# Float and double blocks for the float planner: prices and sensor readings
# at a few decimals, full-precision noise and drifts, values with a maximum
# precision (CudaArrayGenerator::CreateRandomFloatsWithMaxPrecision), runs,
# integer counts stored as floats, near-constant columns and outliers.
FLOAT_FAMILIES = ["price_walk", "sensor", "random_full", "max_precision", "step", "int_valued", "near_constant",
                  "outliers", "drift"]


def _float_params(family: str, r: np.random.Generator) -> dict:
    p = {"double": bool(r.integers(0, 2)), "decimals": int(r.integers(0, 5))}
    if family == "price_walk":
        p.update(start=float(np.exp(r.uniform(0, np.log(1e4)))), sigma=float(np.exp(r.uniform(np.log(1e-3), 2))))
    elif family == "sensor":
        p.update(period=float(np.exp(r.uniform(np.log(8), np.log(1e5)))), amp=float(np.exp(r.uniform(-3, 6))),
                 noise=float(np.exp(r.uniform(-8, 0))), offset=float(r.uniform(-100, 100)))
    elif family == "random_full":
        p.update(scale=float(np.exp(r.uniform(-5, 15))))
    elif family == "max_precision":
        p.update(prec=int(r.integers(0, 4)), scale=float(np.exp(r.uniform(0, 10))))
    elif family == "step":
        p.update(mean_run=float(np.exp(r.uniform(np.log(2), np.log(5000)))), scale=float(np.exp(r.uniform(0, 12))))
    elif family == "int_valued":
        p.update(s=float(r.uniform(1.05, 2.5)), scale=int(r.integers(1, 1000)))
    elif family == "near_constant":
        p.update(frac=float(r.uniform(0, 0.2)), value=float(r.uniform(-1e3, 1e3)))
    elif family == "outliers":
        p.update(mu=float(r.uniform(-1e3, 1e3)), sigma=float(np.exp(r.uniform(-3, 5))),
                 frac=float(r.uniform(0, 0.05)))
    else:
        p.update(step=float(np.exp(r.uniform(-12, 0))))
    return p


def float_block(family: str, p: dict, n: int, r: np.random.Generator) -> np.ndarray:
    """One float32 or float64 block (p["double"]) of `n` rows."""
    d = p["decimals"]
    if family == "price_walk":
        x = np.round(p["start"] + np.cumsum(r.normal(0, p["sigma"], n)), d)
    elif family == "sensor":
        t = np.arange(n)
        x = np.round(p["offset"] + p["amp"] * np.sin(2 * np.pi * t / p["period"]) + r.normal(0, p["noise"], n), d)
    elif family == "random_full":
        x = r.normal(0, p["scale"], n)
    elif family == "max_precision":
        pr = r.integers(0, p["prec"] + 1, n)
        x = np.round(r.random(n) * p["scale"] * 10.0 ** pr) / 10.0 ** pr
    elif family == "step":
        runs = r.geometric(1.0 / p["mean_run"], n // 2 + 2)
        x = np.resize(np.repeat(np.round(r.random(runs.size) * p["scale"], d), runs)[:n], n)
    elif family == "int_valued":
        x = (r.zipf(p["s"], n) % 100000).astype(np.float64) * p["scale"]
    elif family == "near_constant":
        x = np.full(n, round(p["value"], d))
        m = r.random(n) < p["frac"]
        x[m] = np.round(r.uniform(-1e3, 1e3, int(m.sum())), d)
    elif family == "outliers":
        x = np.round(r.normal(p["mu"], p["sigma"], n), d)
        m = r.random(n) < p["frac"]
        x[m] = r.uniform(-1e9, 1e9, int(m.sum()))
    else:
        x = np.cumsum(r.normal(0, p["step"], n)) + 20.0
    return x.astype(np.float64 if p["double"] else np.float32)


def random_float_blocks(count: int, n: int, seed: int) -> Iterator[Tuple[str, np.ndarray]]:
    r = np.random.default_rng(seed)
    for _ in range(count):
        fam = FLOAT_FAMILIES[int(r.integers(0, len(FLOAT_FAMILIES)))]
        yield fam, float_block(fam, _float_params(fam, r), n, r)


def float_regime_stream(count: int, n: int, regime: int, seed: int) -> Iterator[Tuple[int, str, np.ndarray]]:
    r = np.random.default_rng(seed)
    fam, p = None, None
    for i in range(count):
        if i % regime == 0:
            fam = FLOAT_FAMILIES[int(r.integers(0, len(FLOAT_FAMILIES)))]
            p = _float_params(fam, r)
        yield i // regime, fam, float_block(fam, p, n, r)
