"""The thesis' "optimizer vs single scheme" experiment on this library
(reference docs/mgr/CompressionOptimizerCUDA.tex:1276-1304, with the data
sets of optimizer/optimizer_unittest.cpp:347-486).

A column arrives in parts whose character changes every few parts.  The
stateful optimizer (CompressionOptimizer::CompressData: the exhaustive phase
on the first part and every 10th, the edge statistics and tree repair in
between) compresses part after part; the single scheme is the tree the
exhaustive planner picks for the first part, applied to the whole column.
The gain is the ratio of the two compression ratios.

The generators restate CudaArrayGenerator's host patterns
(util/generator/cuda_array_generator.cpp:69-166,
cuda_array_generator_dev.cu:67-81) with a seeded NumPy generator in place
of rand() / cuRAND; the reference's "time from the test file" parts (its
sample_data/info.log is not in its tree) are synthetic timestamps.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, List, Sequence, Tuple

import numpy as np

from . import datasets as D


def pattern_a(npdt, part: int, length: int, step, mn, mx, size: int) -> np.ndarray:
    """GetFakeDataWithPatternA: the level rises by `step` every `length` rows
    (a spike to max at each boundary) and bounces between min and max."""
    out = np.empty(size, dtype=npdt)
    old_min = npdt(mn)
    start = part * size
    cur = npdt(mn + (start // length) * step)
    st = npdt(step)
    mx = npdt(mx)
    i = 0
    while i < size:
        level = (cur if cur > old_min else old_min) if cur < mx else mx
        nxt = i + (-(start + i)) % length  # next row on a length boundary
        if nxt == i:  # boundary row: the spike, then the level moves
            out[i] = mx
            cur = npdt(cur + st)
            if cur > mx or cur < old_min:
                st = npdt(-st)
            i += 1
            continue
        end = min(nxt, size)
        out[i:end] = level
        i = end
    return out


def pattern_b(npdt, part: int, length: int, mn, mx, size: int, rng: np.random.Generator) -> np.ndarray:
    """GetFakeDataWithPatternB: alternating spans of noisy extremes
    (max - r, min + r, r in 0..5) and a triangle from max down and back."""
    start = part * size
    idx = np.arange(size, dtype=np.int64)
    g = start + idx
    r = rng.integers(0, 6, size)
    noisy = np.where(idx % 2 == 0, mx - r, mn + r)
    if np.issubdtype(np.dtype(npdt), np.integer):
        step = (int(mx) - int(mn)) // (length // 2)
    else:
        step = (mx - mn) / (length // 2)
    x = g % length
    tri = np.where(x < length // 2, mx - x * step, mx - (length - x) * step)
    return np.where((g // length) % 2 == 0, noisy, tri).astype(npdt)


def fake_time(size: int, rng: np.random.Generator, start: int = 0, flatness: float = 0.5) -> np.ndarray:
    """GetFakeDataForTime: the time advances by 1 or 2 on a (1 - flatness)
    share of the rows."""
    step = 1 + rng.integers(0, 2, size)
    move = rng.integers(0, 100, size) > flatness * 100
    return (start + np.cumsum(np.where(move, step, 0))).astype(np.int64)


def floats_max_precision(size: int, prec: int, rng: np.random.Generator) -> np.ndarray:
    """CreateRandomFloatsWithMaxPrecision: uniform [0, 1) floats rounded to a
    random precision in 0..prec."""
    x = rng.random(size, dtype=np.float32)
    p = rng.integers(0, prec + 1, size)
    return (np.round(x.astype(np.float64) * 10.0 ** p) / 10.0 ** p).astype(np.float32)


def data_sets(size: int) -> Dict[str, Tuple[int, List[Callable[[int, np.random.Generator], np.ndarray]]]]:
    """The five mixes of the thesis' table (DataType, the regime generators
    taking turns); a generator gets the part's index and a generator."""
    from . import D_FLOAT, D_INT

    i32, f32 = np.int32, np.float32
    return {
        "B": (D_INT, [lambda k, r: pattern_b(i32, k, 3 * size, -10_000, 10_000, size, r)]),
        "A+B": (D_INT, [lambda k, r: pattern_a(i32, k, 10, 1, -100, 1000, size),
                        lambda k, r: pattern_b(i32, k, 3 * size, 10_000, 10_000_000, size, r)]),
        "Time+A": (D_INT, [lambda k, r: (1_400_000_000 + fake_time(size, r, k * size)).astype(i32),
                           lambda k, r: pattern_a(i32, k, 100, 1, -100, 1_000_000, size)]),
        "A+B+MaxPrec": (D_FLOAT, [lambda k, r: pattern_a(f32, k, 100, 1.0, -100.0, 1e6, size),
                                  lambda k, r: pattern_b(f32, k, 3 * size, -1e4, 1e4, size, r),
                                  lambda k, r: floats_max_precision(size, 2, r)]),
        "Time+Cons+Rand": (D_INT, [lambda k, r: D.timestamps(size, seed=1000 + k),
                                   lambda k, r: np.arange(size, dtype=i32),
                                   lambda k, r: r.integers(-2**31, 2**31 - 1, size).astype(i32)]),
    }


def make_parts(gens: Sequence[Callable], parts: int, per_regime: int, seed: int) -> List[np.ndarray]:
    rng = np.random.default_rng(seed)
    return [gens[(k // per_regime) % len(gens)](k, rng) for k in range(parts)]


def optimizer_vs_single_scheme(ctx, size: int = 10_000, parts: int = 75, per_regime: int = 5,
                               seed: int = 7) -> Dict[str, dict]:
    """COpt (the optimizer part by part), CSch (the first part's best tree on
    the whole column) and their ratio, per data set; every blob is decoded
    and checked."""
    from . import CompressionOptimizer, CompressionTree, NP_DTYPE

    out = {}
    for name, (dtype, gens) in data_sets(size).items():
        data = make_parts(gens, parts, per_regime, seed)
        raw = sum(p.nbytes for p in data)
        opt = CompressionOptimizer(ctx)
        t0 = time.perf_counter()
        blobs = [opt.CompressData(ctx.upload(p), dtype) for p in data]
        ctx.sync()
        t_opt = time.perf_counter() - t0
        c_opt = raw / sum(b.nbytes for b in blobs)
        exact = all(CompressionTree(ctx, None, dtype).Decompress(b).to_bytes() == p.tobytes()
                    for b, p in zip(blobs, data))
        del blobs
        first, _, _ = CompressionOptimizer(ctx).PlanExhaustive(ctx.upload(data[0]), dtype)
        whole = np.concatenate(data).astype(NP_DTYPE[dtype])
        tree = CompressionTree(ctx, first.path, dtype)
        blob = tree.Compress(ctx.upload(whole))
        c_single = whole.nbytes / blob.nbytes
        exact = exact and tree.Decompress(blob).to_bytes() == whole.tobytes()
        out[name] = {"parts": parts, "rows_per_part": size, "regime_every": per_regime,
                     "optimizer_ratio": round(c_opt, 4), "single_scheme_ratio": round(c_single, 4),
                     "gain": round(c_opt / c_single, 3), "single_scheme": tree.path,
                     "final_tree": opt.GetOptimalTree(), "optimizer_gbps": round(raw / t_opt / 1e9, 4),
                     "bit_exact": bool(exact)}
    gains = [v["gain"] for v in out.values()]
    out["mean_gain"] = round(float(np.mean(gains)), 3)
    return out
