"""Chooses the learned planner's candidate trees (the network's 48 outputs).

The exhaustive planner (CompressionOptimizer phase 1, reference
optimizer/compression_optimizer.cpp:16-36 via FullStatisticsUpdate) is run on
seeded blocks of every synthetic family; the trees it picks, plus the
current candidate list, form the pool.  Every pool tree compresses every
block on the GPU (exact serialized sizes), and trees are then added greedily,
each time the one that most lowers the sum over blocks of log2(best size so
far), until 48 are chosen.  The result is printed as JSON pre-order paths
(the list compiled into learned.cpp) and written to --out.

    python -m gputimeseriescompressionoptimizer_amd.select_candidates --per-family 64 --out cands.json
    (--big-rows 1048576 --big-per-family 16: also blocks of 1M rows)
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

import gputimeseriescompressionoptimizer_amd as g
from gputimeseriescompressionoptimizer_amd import planner_data as P


def greedy(sizes: np.ndarray, k: int) -> list:
    """Indices of k columns of sizes[blocks, trees] chosen greedily to
    minimise sum_b log2(min over chosen of sizes[b, t])."""
    L = np.log2(np.maximum(sizes, 1.0))
    chosen: list = []
    best = np.full(L.shape[0], np.inf)
    for _ in range(min(k, L.shape[1])):
        cand = np.minimum(best[:, None], L)
        cost = np.where(np.isinf(cand), 1e9, cand).sum(0)
        cost[chosen] = np.inf
        t = int(np.argmin(cost))
        chosen.append(t)
        best = np.minimum(best, L[:, t])
    return chosen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-family", type=int, default=64)
    ap.add_argument("--rows", type=int, default=65536)
    ap.add_argument("--big-rows", type=int, default=0, help="also blocks of this many rows (0: none)")
    ap.add_argument("--big-per-family", type=int, default=0)
    ap.add_argument("--k", type=int, default=g.MLP_OUTPUTS)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default="candidates.json")
    ap.add_argument("--float", action="store_true", help="float32 / float64 blocks of the float families")
    args = ap.parse_args()
    cls = g.PLANNER_FLOAT if args.float else g.PLANNER_INT
    families = P.FLOAT_FAMILIES if args.float else P.FAMILIES
    make = (lambda f, r, n: P.float_block(f, P._float_params(f, r), n, r)) if args.float else \
        (lambda f, r, n: P.block(f, P._params(f, r), n, r))
    sizes_of = [(args.rows, args.per_family)] + ([(args.big_rows, args.big_per_family)] if args.big_rows else [])
    r = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    with g.Context(0) as ctx:
        pool = [tuple(p) for p in g.planner_candidates(cls)]
        blocks = []
        for fam in families:
            for rows, count in sizes_of:
                for _ in range(count):
                    x = make(fam, r, rows)
                    buf = ctx.upload(x)
                    tree, _, _ = g.CompressionOptimizer(ctx).PlanExhaustive(buf, g.DTYPE_OF_NP[x.dtype])
                    if tuple(tree.path) not in pool:
                        pool.append(tuple(tree.path))
                    blocks.append((fam, x))
            print(f"[exhaustive] {fam}: pool {len(pool)} trees, {time.perf_counter() - t0:.0f} s", flush=True)
        sizes = np.zeros((len(blocks), len(pool)))
        for b, (_, x) in enumerate(blocks):
            buf = ctx.upload(x)
            dt = g.DTYPE_OF_NP[x.dtype]
            for t, p in enumerate(pool):
                try:
                    sizes[b, t] = g.CompressionTree(ctx, list(p), dt).Compress(buf).nbytes
                except g.DdjError:  # a tree this type cannot take (AFL on double)
                    sizes[b, t] = 256.0 * x.nbytes
            if (b + 1) % 64 == 0:
                print(f"[sizes] {b + 1}/{len(blocks)} blocks, {time.perf_counter() - t0:.0f} s", flush=True)
    pick = greedy(sizes, args.k)
    chosen = [list(pool[i]) for i in pick]
    best_all = sizes.min(1).sum()
    best_k = sizes[:, pick].min(1).sum()
    old = sizes[:, : len(g.planner_candidates(cls))].min(1).sum()
    res = {"candidates": chosen, "pool": len(pool), "blocks": len(blocks),
           "bytes_best_of_pool": best_all, "bytes_best_of_chosen": best_k, "bytes_best_of_previous": old,
           "chosen_vs_pool": best_k / best_all, "previous_vs_pool": old / best_all}
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps({k: v for k, v in res.items() if k != "candidates"}))
    print(json.dumps(chosen))


if __name__ == "__main__":
    main()
