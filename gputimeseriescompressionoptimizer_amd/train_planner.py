"""Trains the learned planner on the GPU and evaluates it.

1. Labels: `--blocks` random int32 blocks of `--rows` rows (planner_data
   families with random parameters; with `--float`, float32 and float64
   blocks of the float families for the float network); for each, the 32
   features (one statistics pass) and the exact blob size of all 48
   candidate trees.
2. Training: the network's own HIP forward/backward/Adam kernels, MSE on
   log2(blob bytes / input bytes).
3. Evaluation on a stream whose family changes every `--regime` blocks:
   compressed bytes with the learned choice against the best of the 48
   candidates (every block) and against the exhaustive planner
   (CompressionOptimizer phase 1, every `--exhaustive-every`-th block), and
   the planning rate of each.

    python -m gputimeseriescompressionoptimizer_amd.train_planner --out-dir /tmp/planner
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import gputimeseriescompressionoptimizer_amd as g  # noqa: E402
from gputimeseriescompressionoptimizer_amd import planner_data as PD  # noqa: E402

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_weights.npy")
WEIGHTS_FLOAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_weights_float.npy")


def dtype_of(x: np.ndarray) -> int:
    return g.DTYPE_OF_NP[x.dtype]


def label_blocks(ctx, blocks, log_every=500, cls=g.PLANNER_INT):
    net = g.LearnedPlanner(ctx, cls=cls)
    X, T, fams = [], [], []
    t0 = time.perf_counter()
    for i, (fam, x) in enumerate(blocks):
        buf = ctx.upload(x)
        X.append(net.Features(buf, dtype_of(x)))
        T.append(net.Labels(buf, dtype_of(x)))
        fams.append(fam)
        if log_every and (i + 1) % log_every == 0:
            print(f"[label] {i + 1} blocks, {time.perf_counter() - t0:.1f} s", flush=True)
    return np.array(X, np.float32), np.array(T, np.float32), fams


def evaluate(ctx, net, stream, exhaustive_every: int, min_margin: float = 0.0, log_every: int = 0,
             check_within: float = 0.3, max_checked: int = 4):
    """Per block: learned choice (timed), the checked learned choice (exact
    sizes for the candidates predicted within `check_within` of the best;
    timed), best-of-48 size, and every k-th block the exhaustive planner's
    choice (timed) and its size."""
    paths = g.planner_candidates(net.cls)
    learned_bytes = best_bytes = raw = checked_bytes = 0.0
    ex_learned = ex_checked = ex_bytes = 0.0
    t_learned = t_checked = t_ex = 0.0
    n_ex = n = fallbacks = n_checked = 0
    lab = g.LearnedPlanner(ctx, cls=net.cls)
    for i, (_, fam, x) in enumerate(stream):
        dt = dtype_of(x)
        buf = ctx.upload(x)
        ctx.sync()
        t0 = time.perf_counter()
        path, fb, _ = net.Plan(buf, dt, min_margin=min_margin)
        t_learned += time.perf_counter() - t0
        fallbacks += int(fb)
        ctx.sync()
        t0 = time.perf_counter()
        cpath, k = net.PlanChecked(buf, dt, within=check_within, max_checked=max_checked)
        t_checked += time.perf_counter() - t0
        n_checked += int(k > 0)
        labels = lab.Labels(buf, dt)
        sizes = np.exp2(labels.astype(np.float64)) * x.nbytes
        idx = paths.index(path) if path in paths else -1
        chosen = sizes[idx] if idx >= 0 else g.CompressionTree(ctx, path, dt).Compress(buf).nbytes
        cchosen = sizes[paths.index(cpath)]
        learned_bytes += chosen
        checked_bytes += cchosen
        best_bytes += sizes.min()
        raw += x.nbytes
        n += 1
        if exhaustive_every and i % exhaustive_every == 0:
            opt = g.CompressionOptimizer(ctx)
            ctx.sync()
            t0 = time.perf_counter()
            tree, _, _ = opt.PlanExhaustive(buf, dt)
            t_ex += time.perf_counter() - t0
            ex_bytes += tree.Compress(buf).nbytes
            ex_learned += chosen
            ex_checked += cchosen
            n_ex += 1
        if log_every and (i + 1) % log_every == 0:
            print(f"[eval] {i + 1} blocks", flush=True)
    return {
        "blocks": n,
        "learned_bytes": learned_bytes,
        "best_of_48_bytes": best_bytes,
        "input_bytes": raw,
        "learned_vs_best_of_48": learned_bytes / best_bytes,
        "exhaustive_blocks": n_ex,
        "learned_vs_exhaustive": ex_learned / ex_bytes if ex_bytes else None,
        "learned_blocks_per_s": n / t_learned if t_learned else None,
        "exhaustive_blocks_per_s": n_ex / t_ex if t_ex else None,
        "planning_speedup": (n / t_learned) / (n_ex / t_ex) if t_ex and t_learned else None,
        "fallbacks": fallbacks,
        "checked": {"within_log2": check_within, "max_checked": max_checked, "blocks_checked": n_checked,
                    "vs_best_of_48": checked_bytes / best_bytes,
                    "vs_exhaustive": ex_checked / ex_bytes if ex_bytes else None,
                    "blocks_per_s": n / t_checked if t_checked else None,
                    "planning_speedup": (n / t_checked) / (n_ex / t_ex) if t_ex and t_checked else None},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=6000)
    ap.add_argument("--rows", type=int, default=65536)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-blocks", type=int, default=1000)
    ap.add_argument("--regime", type=int, default=500)
    ap.add_argument("--exhaustive-every", type=int, default=20)
    ap.add_argument("--out-dir", default=".")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--eval-only", metavar="WEIGHTS", help="skip labelling and training; evaluate these weights")
    ap.add_argument("--no-eval", action="store_true", help="label and train only")
    ap.add_argument("--float", action="store_true", help="the float network: float32 / float64 blocks")
    ap.add_argument("--dataset", help="label set (npz with X, T) to train on instead of labelling")
    ap.add_argument("--big-blocks", type=int, default=0, help="also label this many blocks of --big-rows rows")
    ap.add_argument("--big-rows", type=int, default=1 << 20)
    ap.add_argument("--big-eval-blocks", type=int, default=0, help="also evaluate on a stream of --big-rows blocks")
    args = ap.parse_args()
    os.makedirs(args.out_dir, exist_ok=True)
    ctx = g.Context(0)
    cls = g.PLANNER_FLOAT if args.float else g.PLANNER_INT
    stream = PD.float_regime_stream if args.float else PD.regime_stream
    if args.eval_only:
        net = g.LearnedPlanner(ctx, cls=cls)
        net.Load(args.eval_only)
        ev = evaluate(ctx, net, stream(args.eval_blocks, args.rows, args.regime, args.seed + 1000),
                      args.exhaustive_every, log_every=500)
        res = {"weights": os.path.basename(args.eval_only), "rows_per_block": args.rows, "regime": args.regime,
               "eval": ev}
        with open(os.path.join(args.out_dir, "planner_eval.json"), "w") as f:
            json.dump(res, f, indent=1)
        print(json.dumps(res))
        return
    if args.dataset:
        d = np.load(args.dataset)
        X, T = d["X"], d["T"]
    else:
        gen = PD.random_float_blocks if args.float else PD.random_blocks
        X, T, fams = label_blocks(ctx, gen(args.blocks, args.rows, args.seed), cls=cls)
        if args.big_blocks:  # blocks of both sizes, shuffled together
            Xb, Tb, fb = label_blocks(ctx, gen(args.big_blocks, args.big_rows, args.seed + 1), cls=cls)
            order = np.random.default_rng(args.seed).permutation(len(X) + len(Xb))
            X, T = np.concatenate([X, Xb])[order], np.concatenate([T, Tb])[order]
            fams = [(fams + fb)[i] for i in order]
        np.savez_compressed(os.path.join(args.out_dir, "planner_dataset.npz"), X=X, T=T, fams=np.array(fams))
    n_val = max(1, len(X) // 10)
    Xtr, Ttr, Xva, Tva = X[n_val:], T[n_val:], X[:n_val], T[:n_val]
    net = g.LearnedPlanner(ctx, seed=args.seed, cls=cls)
    t0 = time.perf_counter()
    history = []
    for e in range(0, args.epochs, 10):
        loss = net.Train(Xtr, Ttr, epochs=min(10, args.epochs - e), batch=args.batch, lr=args.lr, seed=args.seed + e)
        Yva = net.Predict(Xva)
        vloss = float(np.mean((Yva - Tva) ** 2))
        hit = float(np.mean(np.argmin(Yva, 1) == np.argmin(Tva, 1)))
        pick = np.exp2(Tva[np.arange(len(Tva)), np.argmin(Yva, 1)].astype(np.float64))
        size_ratio = float(pick.sum() / np.exp2(Tva.min(1).astype(np.float64)).sum())
        history.append({"epoch": e + 10, "train_mse": loss, "val_mse": vloss, "val_top1": hit,
                        "val_bytes_vs_best_of_48": size_ratio})
        print(f"[train] epoch {e + 10} train {loss:.4f} val {vloss:.4f} top1 {hit:.3f} "
              f"bytes/best {size_ratio:.4f}", flush=True)
    train_s = time.perf_counter() - t0
    weights = os.path.join(args.out_dir, "planner_weights.npy")
    net.Save(weights)
    if args.no_eval:
        with open(os.path.join(args.out_dir, "planner_train.json"), "w") as f:
            json.dump({"train_blocks": len(Xtr), "rows_per_block": args.rows, "epochs": args.epochs,
                       "train_s": train_s, "history": history}, f, indent=1)
        return
    ev = evaluate(ctx, net, stream(args.eval_blocks, args.rows, args.regime, args.seed + 1000),
                  args.exhaustive_every, log_every=50)
    res = {"train_blocks": len(Xtr), "rows_per_block": args.rows, "epochs": args.epochs, "train_s": train_s,
           "history": history, "eval": ev}
    if args.big_blocks:
        res["big_blocks"] = args.big_blocks
        res["big_rows"] = args.big_rows
    if args.big_eval_blocks:
        res["eval_big_rows"] = evaluate(ctx, net, stream(args.big_eval_blocks, args.big_rows, args.regime,
                                                         args.seed + 2000), args.exhaustive_every, log_every=50)
    with open(os.path.join(args.out_dir, "planner_eval.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
