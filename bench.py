"""Benchmark: compress / decompress input GB/s of the single-scheme GPU
workload (four 64M-row int32 columns, each through its fixed tree), one GPU.

A step is one compression pass over the four resident columns (every tree
node's kernels, size read-backs and blob assembly included).  Decompression
of the four blobs is timed separately with the same step count.  The
steps behind `value` run without instrumentation; the same number of steps
is then run again with HIP events around every launch on the library's
stream, and the dominant kernel of that compress region is set against the
8 TB/s HBM peak (its average launch time, its algorithmic bytes).  CPU
baselines: the oracle (the reference's codecs restated on the host)
compresses a bounded prefix of the same columns on one and on all host
threads, and the reference's only CPU path, its FL host loops, packs the
1M-row plumbing column and the 64M timestamp deltas.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "compress/decompress input GB/s at 1 GPU + ratio gap vs exhaustive planner"
HBM_PEAK_GBPS = 8000.0
# HBM bytes per launch of each kernel on this workload, measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by a copy kernel of known
# size (the summary and its commands are in this file).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_int_columns_64M.txt")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        return None
    return None


def columns(rows: int):
    from gputimeseriescompressionoptimizer_amd import datasets as D
    import gputimeseriescompressionoptimizer_amd as g

    return [
        ("timestamps delta+afl", D.timestamps(rows, seed=11), [g.DELTA, g.AFL, g.NONE]),
        ("step function rle+afl", D.step_function(rows, seed=12), [g.RLE, g.AFL, g.NONE, g.AFL, g.NONE]),
        ("zipf(1.2) over 50 codes dict", D.zipf_codes(rows, seed=13), [g.DICT, g.NONE, g.NONE]),
        ("N(1000,50)+1% outliers patch+afl", D.normal_with_outliers(rows, seed=14),
         [g.PATCH, g.AFL, g.NONE, g.AFL, g.NONE]),
    ]


def cpu_threads() -> int:
    """Host threads for the all-cores CPU baselines: the process's CPU
    affinity, at most 16 (one GPU's share of its host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(cols, sample_rows: int, budget_s: float = 10.0):
    """The oracle (the reference's codecs restated in C++) compressing the
    first `sample_rows` rows of each column with the same trees: on one
    thread, and on all host threads (each column sample cut into one part
    per thread, as the column API's batches are independent)."""
    from oracle import cpu_baseline as CB

    columns = [(path, x[:sample_rows]) for _, x, path in cols]
    v1, reps1, el1 = CB.oracle_trees(columns, budget_s / 2, 1)
    nt = cpu_threads()
    vn, repsn, eln = CB.oracle_trees(columns, budget_s / 2, nt)
    return {"value": round(vn / 1e9, 4), "unit": "GB/s", "cores": nt, "kind": "port",
            "value_1_core": round(v1 / 1e9, 4),
            "sample": f"first {sample_rows} rows of each of the 4 columns through the same trees: "
                      f"{reps1} passes in {el1:.1f} s on 1 thread, {repsn} passes in {eln:.1f} s on {nt} threads",
            "host": CB.host_cores()}


def float_section(ctx, rows: int, steps: int):
    """The float price columns (64M rows): GFC against floatToInt cascades,
    compress and decompress GB/s with inputs resident in HBM, ratios, and a
    bit-exact round-trip check of every tree."""
    import numpy as np
    import gputimeseriescompressionoptimizer_amd as g
    from gputimeseriescompressionoptimizer_amd import datasets as D

    cols = [
        ("float64 walk", D.price_walk_f64(rows, seed=21), g.D_DOUBLE,
         {"gfc": [g.GFC, g.NONE, g.NONE], "floatToInt+delta+afl": [g.FLOAT_TO_INT, g.DELTA, g.AFL, g.NONE]}),
        ("float32 walk", D.price_walk_f32(rows, seed=22), g.D_FLOAT,
         {"gfc": [g.GFC, g.NONE, g.NONE],
          "floatToInt+scale+delta+afl": [g.FLOAT_TO_INT, g.SCALE, g.DELTA, g.AFL, g.NONE]}),
    ]
    out = {}
    for name, x, dt, trees in cols:
        buf = ctx.upload(x)
        for tname, path in trees.items():
            tree = g.CompressionTree(ctx, path, dt)
            blob = tree.Compress(buf)
            back = tree.Decompress(blob)
            exact = back.to_bytes() == x.tobytes()
            del back
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                blob = tree.Compress(buf)
            ctx.sync()
            tc = (time.perf_counter() - t0) / steps
            t0 = time.perf_counter()
            for _ in range(steps):
                back = tree.Decompress(blob)
                del back
            ctx.sync()
            td = (time.perf_counter() - t0) / steps
            # one more round trip with HIP events around every launch: the
            # largest kernels' device time, in microseconds
            ctx.profile(True)
            ctx.profile_reset()
            blob = tree.Compress(buf)
            back = tree.Decompress(blob)
            del back
            ctx.sync()
            ktab = ctx.kernel_table()
            ctx.profile(False)
            top = sorted(ktab.items(), key=lambda kv: -kv[1][0])[:6]
            out[f"{name} {tname}"] = {"compress_gbps": round(x.nbytes / tc / 1e9, 2),
                                      "decompress_gbps": round(x.nbytes / td / 1e9, 2),
                                      "ratio": round(x.nbytes / blob.nbytes, 3), "bit_exact": exact,
                                      "kernels_us": {k: round(v[0] * 1e3, 1) for k, v in top}}
            del blob
        del buf
    return out


def batched_section(ctx, n_cols: int, min_rows: int, max_rows: int, workers: int):
    """Many columns of log-uniform lengths, mixed int32 timestamps, float64
    walk prices, int32 Zipf volumes and int32 50-code categoricals, each type
    with the tree the exhaustive planner picks on a 10,000-row sample:
    compressed one column after another (ddj_tree_compress per column)
    against together (ddj_batch_compress)."""
    import numpy as np
    import gputimeseriescompressionoptimizer_amd as g
    from gputimeseriescompressionoptimizer_amd import datasets as D

    r = np.random.default_rng(33)
    lengths = np.floor(np.exp(r.uniform(np.log(min_rows), np.log(max_rows), n_cols))).astype(np.int64)
    kinds = ["timestamps", "walk_f64", "zipf_volumes", "categorical50"]

    def make(kind, n, seed):
        rr = np.random.default_rng(seed)
        if kind == "timestamps":
            return D.timestamps(n, seed=seed), g.D_INT
        if kind == "walk_f64":
            return D.price_walk_f64(n, seed=seed), g.D_DOUBLE
        if kind == "zipf_volumes":
            return ((rr.zipf(1.5, n) % 100000) * 100).astype(np.int32), g.D_INT
        return rr.integers(0, 50, n).astype(np.int32), g.D_INT

    trees = {}
    for k in kinds:
        x, dt = make(k, 10_000, 1)
        tree, _, _ = g.CompressionOptimizer(ctx).PlanExhaustive(ctx.upload(x), dt)
        trees[k] = tree.path
    cols, raw = [], 0
    for i, n in enumerate(lengths):
        k = kinds[i % 4]
        x, dt = make(k, int(n), 100 + i)
        cols.append((ctx.upload(x), dt, trees[k]))
        raw += x.nbytes
        del x
    ctx.sync()

    def per_column():
        return [g.CompressionTree(ctx, path, dt).Compress(buf) for buf, dt, path in cols]

    t0 = time.perf_counter()
    config, _ = g.batch_autotune(ctx, cols) if workers == 0 else ({"workers": workers}, 0.0)
    tune_s = time.perf_counter() - t0

    def together():
        return g.batch_compress(ctx, cols, config=config)

    a = per_column()
    b = together()
    same = all(x.to_bytes() == y.to_bytes() for x, y in zip(a, b))
    out_bytes = sum(x.nbytes for x in a)
    del a, b
    times = {}
    for name, fn in (("per_column", per_column), ("batched", together)):
        best = None
        for _ in range(3):  # best of three runs of the whole batch
            ctx.sync()
            t0 = time.perf_counter()
            res = fn()
            ctx.sync()
            dt_run = time.perf_counter() - t0
            best = dt_run if best is None or dt_run < best else best
            del res
        times[name] = best
    return {"columns": n_cols, "rows": int(lengths.sum()), "input_bytes": raw,
            "rows_min_max": [int(lengths.min()), int(lengths.max())], "trees": trees,
            "config": config, "autotune_s": round(tune_s, 2),
            "per_column_gbps": round(raw / times["per_column"] / 1e9, 2),
            "batched_gbps": round(raw / times["batched"] / 1e9, 2),
            "speedup": round(times["per_column"] / times["batched"], 2), "workers": config.get("workers"),
            "ratio": round(raw / out_bytes, 3), "blobs_identical": same}


def fl_plumbing_section(ctx):
    """The reference's CPU-runnable case: 1,000,000 int32 uniform in [0,1024)
    FL-packed to 10 bits and unpacked, on the GPU and with the reference's
    only CPU path (its FL host loops run_afl_compress_cpu<int,1> /
    run_afl_decompress_cpu<int,1>, restated and built with g++ -O3 in
    oracle/) on one host thread and on all host threads.  The same for the
    FL stage of the 64M-row timestamp column (its deltas at their width)."""
    import numpy as np
    import gputimeseriescompressionoptimizer_amd as g
    from gputimeseriescompressionoptimizer_amd import datasets as D
    from oracle import cpu_baseline as CB

    nt = cpu_threads()

    def one(x, bits, reps_gpu, reps_cpu):
        buf = ctx.upload(x)
        packed = ctx.fl_pack(buf, g.D_INT, bits, x.size)
        exact = np.array_equal(ctx.fl_unpack(packed, g.D_INT, bits, x.size).to_numpy(np.int32), x)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps_gpu):
            packed = ctx.fl_pack(buf, g.D_INT, bits, x.size)
        ctx.sync()
        t_pack = (time.perf_counter() - t0) / reps_gpu
        t0 = time.perf_counter()
        for _ in range(reps_gpu):
            back = ctx.fl_unpack(packed, g.D_INT, bits, x.size)
            del back
        ctx.sync()
        t_unpack = (time.perf_counter() - t0) / reps_gpu
        stream = packed.to_bytes()
        out = {"rows": int(x.size), "bits": bits, "gpu_pack_gbps": round(x.nbytes / t_pack / 1e9, 2),
               "gpu_unpack_gbps": round(x.nbytes / t_unpack / 1e9, 2), "round_trip_exact": exact}
        for label, threads in (("cpu_1_thread", 1), ("cpu_all_threads", nt)):
            ps, us, cpu_stream, cpu_out = CB.fl_loops(x, bits, threads, reps_cpu)
            out[label] = {"threads": threads, "pack_gbps": round(x.nbytes / ps / 1e9, 3),
                          "unpack_gbps": round(x.nbytes / us / 1e9, 3),
                          "stream_identical_to_gpu": cpu_stream.tobytes() == stream,
                          "round_trip_exact": bool(np.array_equal(cpu_out, x))}
        return out

    res = {"cfg0_uniform_1m": one(D.uniform_int(1_000_000, 1024, seed=0), 10, 200, 10)}
    ts = D.timestamps(1 << 26, seed=11)
    d = np.diff(ts.astype(np.int64)).astype(np.int32)
    res["timestamp_deltas_64m"] = one(d, int(d.max()).bit_length(), 20, 5)
    res["cpu"] = {"compiler": "g++ -O3", "isa": CB.isa_level(), **CB.host_cores()}
    return res


def learned_planner_section(ctx, blocks: int, rows: int, regime: int, exhaustive_every: int, big_blocks: int = 0,
                            big_rows: int = 1 << 20):
    """Learned vs exhaustive planning on streams of blocks whose generator
    family changes every `regime` blocks: int32 blocks with the integer
    network, float32/float64 blocks with the float network (weights trained
    by gputimeseriescompressionoptimizer_amd.train_planner on 65,536-row
    blocks), and `big_blocks` int32 blocks of `big_rows` rows (the 1M-row
    blocks of the learned-planner workload) with the integer network."""
    import gputimeseriescompressionoptimizer_amd as g
    from gputimeseriescompressionoptimizer_amd import planner_data as PD
    from gputimeseriescompressionoptimizer_amd.train_planner import WEIGHTS, WEIGHTS_FLOAT, evaluate

    out = {}
    for name, weights, cls, stream in (("int32", WEIGHTS, g.PLANNER_INT, PD.regime_stream),
                                       ("float", WEIGHTS_FLOAT, g.PLANNER_FLOAT, PD.float_regime_stream)):
        if not os.path.exists(weights):
            continue
        net = g.LearnedPlanner(ctx, cls=cls)
        net.Load(weights)
        ev = evaluate(ctx, net, stream(blocks, rows, regime, seed=4242), exhaustive_every)
        ev["stream"] = f"{blocks} blocks x {rows} rows, family changes every {regime} blocks"
        ev["ratio_gap_vs_exhaustive"] = (ev["learned_vs_exhaustive"] - 1.0) if ev["learned_vs_exhaustive"] else None
        out[name] = ev
        if name == "int32" and big_blocks > 0:
            ev = evaluate(ctx, net, PD.regime_stream(big_blocks, big_rows, regime, seed=4243), exhaustive_every)
            ev["stream"] = f"{big_blocks} blocks x {big_rows} rows, family changes every {regime} blocks"
            ev["ratio_gap_vs_exhaustive"] = (ev["learned_vs_exhaustive"] - 1.0) if ev["learned_vs_exhaustive"] else None
            out["int32_1m_rows"] = ev
    return out or None


def access_section(ctx, cols, rows: int, steps: int):
    """The secondary paths on the benchmark's columns: unique -> none against
    dict on the Zipf column (its codes), the 32-lane FL layout against the
    contiguous one (the reference's BM_Pure_Afl_Aligned shape: int32 at 10
    bits), random-access fetches and a partial decode of the timestamp
    column's delta -> afl blob."""
    import numpy as np
    import gputimeseriescompressionoptimizer_amd as g
    from gputimeseriescompressionoptimizer_amd import datasets as D

    def timeit(fn, reps):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
            del r
        ctx.sync()
        return (time.perf_counter() - t0) / reps

    out = {}
    # unique -> none vs dict -> none, none on the Zipf codes
    zipf = next(x for name, x, _ in cols if name.startswith("zipf"))
    buf = ctx.upload(zipf)
    for name, path in (("unique", [g.UNIQUE, g.NONE]), ("dict", [g.DICT, g.NONE, g.NONE])):
        tree = g.CompressionTree(ctx, path, g.D_INT)
        blob = tree.Compress(buf)
        exact = tree.Decompress(blob).to_bytes() == zipf.tobytes()
        tc = timeit(lambda: tree.Compress(buf), steps)
        td = timeit(lambda: tree.Decompress(blob), steps)
        out[f"zipf_{name}"] = {"compress_gbps": round(zipf.nbytes / tc / 1e9, 2),
                               "decompress_gbps": round(zipf.nbytes / td / 1e9, 2),
                               "ratio": round(zipf.nbytes / blob.nbytes, 3), "bit_exact": exact}
        del blob
    del buf
    # 32-lane vs contiguous FL: int32 at 10 bits
    x = D.uniform_int(rows, 1024, seed=0)
    xb = ctx.upload(x)
    lanes = ctx.fl_pack_lanes(xb, g.D_INT, 10, rows)
    flat = ctx.fl_pack(xb, g.D_INT, 10, rows)
    exact = np.array_equal(ctx.fl_unpack_lanes(lanes, g.D_INT, 10, rows).to_numpy(np.int32), x)
    res = {"rows": rows, "bits": 10, "round_trip_exact": exact}
    for name, pack, unpack in (("lanes32", lambda: ctx.fl_pack_lanes(xb, g.D_INT, 10, rows),
                                lambda: ctx.fl_unpack_lanes(lanes, g.D_INT, 10, rows)),
                               ("contiguous", lambda: ctx.fl_pack(xb, g.D_INT, 10, rows),
                                lambda: ctx.fl_unpack(flat, g.D_INT, 10, rows))):
        res[f"{name}_pack_gbps"] = round(x.nbytes / timeit(pack, steps) / 1e9, 2)
        res[f"{name}_unpack_gbps"] = round(x.nbytes / timeit(unpack, steps) / 1e9, 2)
    out["fl_layouts"] = res
    # random access into the contiguous stream, and a row range of a delta blob
    r = np.random.default_rng(5)
    k = 1 << 20
    idx = ctx.upload(r.integers(0, rows, k).astype(np.uint64))
    got = ctx.fl_fetch(flat, g.D_INT, 10, rows, idx).to_numpy(np.int32)
    ok = np.array_equal(got, x[idx.to_numpy(np.uint64).astype(np.int64)])
    tf = timeit(lambda: ctx.fl_fetch(flat, g.D_INT, 10, rows, idx), steps)
    ts = next(x for name, x, _ in cols if name.startswith("timestamps"))
    tree = g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_INT)
    blob = tree.Compress(ctx.upload(ts))
    b0, cnt = len(ts) // 2, 1 << 20
    part_ok = tree.DecompressRows(blob, b0, cnt).to_bytes() == ts[b0:b0 + cnt].tobytes()
    tr = timeit(lambda: tree.DecompressRows(blob, b0, cnt), steps)
    tfull = timeit(lambda: tree.Decompress(blob), steps)
    out["random_access"] = {"fetch_rows": k, "fetch_mrows_per_s": round(k / tf / 1e6, 1), "fetch_exact": ok,
                            "range_rows": cnt, "range_begin": b0, "range_ms": round(tr * 1e3, 4),
                            "full_decode_ms": round(tfull * 1e3, 4), "range_exact": part_ok}
    return out


def copy_calibration():
    import numpy as np
    import gputimeseriescompressionoptimizer_amd as g

    with g.Context(0) as ctx:
        x = np.arange(1 << 26, dtype=np.int32)
        buf = ctx.upload(x)
        tree = g.CompressionTree(ctx, [g.NONE], g.D_INT)
        for _ in range(5):
            blob = tree.Compress(buf)
            back = tree.Decompress(blob)
            ctx.sync()
        assert blob.nbytes == x.nbytes + 16 and back.nbytes == x.nbytes
        print(json.dumps({"copy_calibration": {"column_bytes": x.nbytes, "blob_bytes": blob.nbytes, "passes": 5}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=1 << 26)
    ap.add_argument("--cpu-sample", type=int, default=1 << 21)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--planner-blocks", type=int, default=1000)
    ap.add_argument("--planner-regime", type=int, default=500, help="blocks per regime of the planner's stream")
    ap.add_argument("--planner-rows", type=int, default=65536)
    ap.add_argument("--planner-big-blocks", type=int, default=1000,
                    help="int32 blocks of --planner-big-rows rows also planned by the integer network (0: none)")
    ap.add_argument("--planner-big-rows", type=int, default=1 << 20)
    ap.add_argument("--no-planner", action="store_true")
    ap.add_argument("--no-experiment", action="store_true",
                    help="skip the optimizer-vs-single-scheme experiment on regime-shifting columns")
    ap.add_argument("--experiment-rows", type=int, default=10_000,
                    help="rows per part in that experiment (the reference's unittest size)")
    ap.add_argument("--float-rows", type=int, default=1 << 26)
    ap.add_argument("--batch-columns", type=int, default=256)
    ap.add_argument("--batch-max-rows", type=int, default=16 << 20)
    ap.add_argument("--batch-workers", type=int, default=0, help="0: autotune the batch configuration")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-float", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip unique / 32-lane FL / random access on the benchmark's columns")
    ap.add_argument("--compress-only", action="store_true",
                    help="warmup and timed compress steps only (for hardware-counter runs)")
    ap.add_argument("--no-concurrent", action="store_true",
                    help="skip the four columns through the batch API (for kernel traces of the steps alone)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run plain copies of known size (for scaling hardware counters): a [none] "
                         "tree over a 64M-row int32 column, 5 times; its blob assembly is one "
                         "k_gather_copy reading the 256 MiB column (16-byte aligned), and decompressing "
                         "the blob copies its 256 MiB of data into a fresh (aligned) column")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (columns would shard "
              "across GPUs with a host-side scan of blob sizes)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import numpy as np
    import gputimeseriescompressionoptimizer_amd as g

    cols = columns(args.rows)
    ctx = g.Context(0)
    dev = [(name, ctx.upload(x), path, x) for name, x, path in cols]
    in_bytes = sum(x.nbytes for _, x, _ in cols)
    trees = [g.CompressionTree(ctx, path, g.D_INT) for _, _, path, _ in dev]

    def compress_all():
        return [t.Compress(buf) for t, (_, buf, _, _) in zip(trees, dev)]

    def decompress_all(blobs):
        return [t.Decompress(b) for t, b in zip(trees, blobs)]

    def timed(fn, profile):
        """Seconds per step over args.steps steps; with profile, also the
        per-kernel HIP-event table of the same steps."""
        ctx.sync()
        if profile:
            ctx.profile(True)
            ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.sync()
        dt = (time.perf_counter() - t0) / args.steps
        tab = None
        if profile:
            tab = ctx.kernel_table()
            ctx.profile(False)
        return dt, tab

    for _ in range(args.warmup):
        blobs = compress_all()
    holder = {}

    def comp_step():
        holder["blobs"] = compress_all()

    if args.compress_only:
        t_comp, _ = timed(comp_step, False)
        print(json.dumps({"compress_only": True, "ms_per_step": round(t_comp * 1e3, 3), "steps": args.steps}))
        return
    t_comp, _ = timed(comp_step, False)           # value: no events in the way
    t_comp_prof, ktab = timed(comp_step, True)    # kernel table: HIP events around every launch
    blobs = holder["blobs"]

    # the same job with the four columns overlapped (ddj_batch_compress,
    # worker streams, autotuned): blobs must be identical
    concurrent = None
    if not args.no_concurrent:
        bcols = [(buf, g.D_INT, path) for _, buf, path, _ in dev]
        bcfg, _ = g.batch_autotune(ctx, bcols)
        same_concurrent = all(a.to_bytes() == b.to_bytes()
                              for a, b in zip(g.batch_compress(ctx, bcols, config=bcfg), blobs))

        def conc_step():
            holder["cblobs"] = g.batch_compress(ctx, bcols, config=bcfg)

        t_conc, _ = timed(conc_step, False)
        holder.pop("cblobs", None)
        concurrent = {"compress_gbps": round(in_bytes / t_conc / 1e9, 3), "ms_per_step": round(t_conc * 1e3, 3),
                      "config": bcfg, "blobs_identical": same_concurrent}

    # decompression of the same blobs, verified once
    outs = decompress_all(blobs)
    for (_, _, _, x), o in zip(dev, outs):
        if not np.array_equal(o.to_numpy(np.int32), x):
            raise SystemExit("round trip mismatch")
    del outs

    def dec_step():
        outs = decompress_all(blobs)
        del outs

    t_dec, _ = timed(dec_step, False)
    t_dec_prof, dtab = timed(dec_step, True)

    ratios = {name: round(x.nbytes / b.nbytes, 3) for (name, _, _, x), b in zip(dev, blobs)}
    out_bytes = sum(b.nbytes for b in blobs)

    # whole-tree rooflines: (input + blob bytes) / time of each column's tree
    # alone, and of the whole step, against the HBM peak (intermediate
    # buffers do not count, so only fusion moves these figures)
    per_tree = {}
    for t, (name, buf, _, x), b in zip(trees, dev, blobs):
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t.Compress(buf)
        ctx.sync()
        tc = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t.Decompress(b)
        ctx.sync()
        td = (time.perf_counter() - t0) / args.steps
        algo = x.nbytes + b.nbytes
        per_tree[name] = {"compress_ms": round(tc * 1e3, 4), "decompress_ms": round(td * 1e3, 4),
                          "compress_frac": round(algo / tc / 1e9 / HBM_PEAK_GBPS, 4),
                          "decompress_frac": round(algo / td / 1e9 / HBM_PEAK_GBPS, 4)}
    tree_roofline = {"algorithmic_bytes_per_step": in_bytes + out_bytes,
                     "compress_frac": round((in_bytes + out_bytes) / t_comp / 1e9 / HBM_PEAK_GBPS, 4),
                     "decompress_frac": round((in_bytes + out_bytes) / t_dec / 1e9 / HBM_PEAK_GBPS, 4),
                     "per_tree": per_tree}

    # dominant kernel of the compress step (largest total device time)
    dom = max(ktab.items(), key=lambda kv: kv[1][0])
    dname, (dms, dlaunch, dbytes) = dom
    achieved = dbytes / (dms / 1e3) / 1e9 if dms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(dname),
                "avg_launch_us": round(dms / dlaunch * 1e3, 2), "launches_per_step": dlaunch / args.steps,
                "algorithmic_bytes_per_launch": int(dbytes / dlaunch)}
    def table(tab):
        return {k: {"ms_per_step": round(v[0] / args.steps, 4), "launches_per_step": v[1] / args.steps,
                    "gbps": round(v[2] / (v[0] / 1e3) / 1e9, 1) if v[0] > 0 and v[2] > 0 else None}
                for k, v in sorted(tab.items(), key=lambda kv: -kv[1][0])}

    kernels = table(ktab)
    dkernels = table(dtab)

    res = {
        "metric": METRIC,
        "value": round(in_bytes / t_comp / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_comp * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: seeded numpy generators of the four single-scheme columns",
        "config": {"workload": "single-scheme GPU kernels, 4 int32 columns x %d rows: timestamps delta+afl, "
                               "step function rle+afl, zipf(1.2)/50 codes dict, N(1000,50)+1%% outliers "
                               "patch+afl" % args.rows,
                   "rows_per_column": args.rows, "input_bytes": in_bytes},
        "decompress_gbps": round(in_bytes / t_dec / 1e9, 3),
        "decompress_ms_per_step": round(t_dec * 1e3, 3),
        "compression_ratio": round(in_bytes / out_bytes, 3),
        "ratios": ratios,
        "roofline": roofline,
        "tree_roofline": tree_roofline,
        "profiled_ms_per_step": {"compress": round(t_comp_prof * 1e3, 3), "decompress": round(t_dec_prof * 1e3, 3)},
        "concurrent_columns": concurrent,
        "kernels": kernels,
        "decompress_kernels": dkernels,
    }
    def step(name):
        print(f"[bench] {name}", file=sys.stderr, flush=True)

    step("fl plumbing")
    res["fl_plumbing"] = fl_plumbing_section(ctx)
    if not args.no_secondary:
        step("secondary paths")
        res["secondary_paths"] = access_section(ctx, cols, args.rows, max(1, args.steps // 2))
    if not args.no_float:
        step("float columns")
        res["float_columns"] = float_section(ctx, args.float_rows, max(1, args.steps // 2))
    if not args.no_batch:
        step("batched columns")
        res["batched_columns"] = batched_section(ctx, args.batch_columns, 4096, args.batch_max_rows,
                                                 args.batch_workers)
    if not args.no_planner:
        step("learned planner")
        res["learned_planner"] = learned_planner_section(ctx, args.planner_blocks, args.planner_rows,
                                                         regime=max(1, min(args.planner_regime, args.planner_blocks)),
                                                         exhaustive_every=10, big_blocks=args.planner_big_blocks,
                                                         big_rows=args.planner_big_rows)
    if not args.no_experiment:
        step("optimizer vs single scheme")
        from gputimeseriescompressionoptimizer_amd.experiments import optimizer_vs_single_scheme

        res["optimizer_vs_single_scheme"] = optimizer_vs_single_scheme(ctx, size=args.experiment_rows)
    ctx.close()
    if not args.no_cpu:
        step("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(cols, min(args.cpu_sample, args.rows))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
