"""Random access and the interleaved FL layout on the GPU, against the oracle.

* ddj_fl_fetch: values at arbitrary rows of contiguous (AflEncoding's) and
  32-lane streams, including rows whose bits straddle words, the first and
  last rows, and rows past the end (an error, after which the library keeps
  working).
* ddj_fl_pack_lanes / ddj_fl_unpack_lanes: byte parity with the oracle's
  run_afl_compress_cpu<T,32> restatement and round trips.
* ddj_tree_decompress_rows: any row range of a blob equals the same slice of
  the column, for trees decoded partially (none, afl, delta -> afl, scale)
  and trees sliced after a whole decode (rle, dict, patch, gfc).
"""
import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g
from gputimeseriescompressionoptimizer_amd import datasets as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INT_TYPES = [g.D_CHAR, g.D_SHORT, g.D_INT, g.D_TIME]


def column(dtype, n, bits, seed):
    npdt = np.dtype(g.NP_DTYPE[dtype])
    r = np.random.default_rng(seed)
    u = r.integers(0, 2**bits - 1, n, dtype=np.uint64, endpoint=True) if bits < 64 else \
        r.integers(0, 2**63, n, dtype=np.uint64)
    return u.astype(npdt.str.replace("i", "u")).view(npdt)


def widths(dtype):
    W = 8 * np.dtype(g.NP_DTYPE[dtype]).itemsize
    return sorted({1, 3, 7, W // 2 + 1, W - 1, W})


def rows_for(n, seed, k=3000):
    r = np.random.default_rng(seed)
    return np.concatenate([[0, n - 1, n // 2], r.integers(0, n, k)]).astype(np.uint64)


@pytest.mark.parametrize("dtype", INT_TYPES)
@pytest.mark.parametrize("n", [1, 33, 4097, 300_001])
def test_fetch_contiguous_stream(ctx, dtype, n):
    for bits in widths(dtype):
        print(f"[fetch] dtype {dtype} n {n} bits {bits}", flush=True)
        x = column(dtype, n, bits, seed=bits)
        stream = ctx.fl_pack(ctx.upload(x), dtype, bits, n)
        assert stream.to_bytes() == O.fl_pack(dtype, bits, x).tobytes()
        rows = rows_for(n, bits)
        got = ctx.fl_fetch(stream, dtype, bits, n, ctx.upload(rows)).to_numpy(g.NP_DTYPE[dtype])
        assert np.array_equal(got, O.fl_fetch(dtype, bits, 1, O.fl_pack(dtype, bits, x), n, rows))
        assert np.array_equal(got, x[rows.astype(np.int64)])


def test_fetch_from_an_afl_leaf_inside_a_blob(ctx):
    """The FL data of an afl -> none blob, addressed at its (unaligned) byte
    offset inside the blob, fetched directly."""
    x = D.uniform_int(100_000, 1 << 13, seed=5)
    enc = g.Encoding(ctx, g.AFL).Encode(ctx.upload(x), g.D_INT)
    meta = enc[0].to_bytes()
    bits = meta[0]
    blob = g.CompressionTree(ctx, [g.AFL, g.NONE], g.D_INT).Compress(ctx.upload(x))
    data_off = 8 + len(meta) + 8 + 8  # afl header + meta, none header + size_t
    data = blob.view(data_off, blob.nbytes - data_off)
    assert data.to_bytes() == enc[1].to_bytes()
    rows = rows_for(x.size, 1)
    got = ctx.fl_fetch(data, g.D_INT, bits, x.size, ctx.upload(rows)).to_numpy(np.int32)
    assert np.array_equal(got, x[rows.astype(np.int64)])


@pytest.mark.parametrize("dtype", INT_TYPES)
@pytest.mark.parametrize("n", [1, 31, 1000, 32 * 64 * 3, 300_001])
def test_lanes_layout_matches_oracle(ctx, dtype, n):
    for bits in widths(dtype):
        print(f"[lanes] dtype {dtype} n {n} bits {bits}", flush=True)
        x = column(dtype, n, bits, seed=bits + 1)
        want = O.fl_pack_lanes(dtype, bits, x)
        stream = ctx.fl_pack_lanes(ctx.upload(x), dtype, bits, n)
        assert stream.to_bytes() == want.tobytes()
        assert np.array_equal(ctx.fl_unpack_lanes(stream, dtype, bits, n).to_numpy(g.NP_DTYPE[dtype]), x)
        rows = rows_for(n, bits, 1000)
        got = ctx.fl_fetch(stream, dtype, bits, n, ctx.upload(rows), lanes=32).to_numpy(g.NP_DTYPE[dtype])
        assert np.array_equal(got, x[rows.astype(np.int64)])


def test_lanes_layout_at_a_reference_benchmark_size(ctx):
    """2^21 int32 rows at 10 bits (afl_encoding_benchmark.cpp:33-47): the
    stream is ceil(b*n/W) words, byte-identical to the oracle."""
    n = 1 << 21
    x = D.uniform_int(n, 1024, seed=3)
    stream = ctx.fl_pack_lanes(ctx.upload(x), g.D_INT, 10, n)
    assert stream.nbytes == (10 * n + 31) // 32 * 4
    assert stream.to_bytes() == O.fl_pack_lanes(g.D_INT, 10, x).tobytes()
    assert np.array_equal(ctx.fl_unpack_lanes(stream, g.D_INT, 10, n).to_numpy(np.int32), x)


def test_fetch_row_past_the_end_is_an_error(ctx):
    x = column(g.D_INT, 1000, 9, seed=1)
    stream = ctx.fl_pack(ctx.upload(x), g.D_INT, 9, 1000)
    with pytest.raises(g.DdjError) as e:
        ctx.fl_fetch(stream, g.D_INT, 9, 1000, ctx.upload(np.array([3, 1000], np.uint64)))
    assert e.value.status == 1
    got = ctx.fl_fetch(stream, g.D_INT, 9, 1000, ctx.upload(np.array([3, 999], np.uint64))).to_numpy(np.int32)
    assert np.array_equal(got, x[[3, 999]])
    with pytest.raises(g.DdjError):  # a stream too short for n values
        ctx.fl_fetch(stream, g.D_INT, 9, 2000, ctx.upload(np.array([3], np.uint64)))


ROW_TREES = {
    "none int32": (g.D_INT, [g.NONE], lambda n, s: D.uniform_int(n, 1 << 20, seed=s)),
    "afl int32": (g.D_INT, [g.AFL, g.NONE], lambda n, s: D.uniform_int(n, 1 << 11, seed=s)),
    "afl int16": (g.D_SHORT, [g.AFL, g.NONE],
                  lambda n, s: np.random.default_rng(s).integers(0, 3000, n).astype(np.int16)),
    "delta-afl int32 timestamps": (g.D_INT, [g.DELTA, g.AFL, g.NONE], lambda n, s: D.timestamps(n, seed=s)),
    "delta-afl int64 walk": (g.D_TIME, [g.DELTA, g.AFL, g.NONE],
                             lambda n, s: np.cumsum(np.random.default_rng(s).integers(-9, 10, n)).astype(np.int64)),
    "delta-afl int8 wrap": (g.D_CHAR, [g.DELTA, g.AFL, g.NONE],
                            lambda n, s: (np.arange(n) * 7 + s).astype(np.int64).astype(np.int8)),
    "scale-afl int32": (g.D_INT, [g.SCALE, g.AFL, g.NONE],
                        lambda n, s: np.random.default_rng(s).integers(-5000, 5000, n).astype(np.int32)),
    "scale-delta-afl int32": (g.D_INT, [g.SCALE, g.DELTA, g.AFL, g.NONE], lambda n, s: D.timestamps(n, seed=s)),
    "rle-afl int32": (g.D_INT, [g.RLE, g.AFL, g.NONE, g.AFL, g.NONE], lambda n, s: D.step_function(n, seed=s)),
    "dict int32": (g.D_INT, [g.DICT, g.NONE, g.NONE], lambda n, s: D.zipf_codes(n, seed=s)),
    "patch-afl int32": (g.D_INT, [g.PATCH, g.AFL, g.NONE, g.AFL, g.NONE],
                        lambda n, s: D.normal_with_outliers(n, seed=s)),
    "gfc float64": (g.D_DOUBLE, [g.GFC, g.NONE, g.NONE], lambda n, s: D.price_walk_f64(n, seed=s)),
}


@pytest.mark.parametrize("name", list(ROW_TREES))
def test_decompress_rows_is_a_slice(ctx, name):
    dtype, path, make = ROW_TREES[name]
    n = 100_003
    x = make(n, 9)
    blob = g.CompressionTree(ctx, path, dtype).Compress(ctx.upload(x))
    tree = g.CompressionTree(ctx, path, dtype)
    r = np.random.default_rng(4)
    ranges = [(0, n), (0, 1), (n - 1, 1), (0, 0), (n, 0), (4095, 2), (4096, 4097), (12345, 8191), (n - 4097, 4097)]
    ranges += [(int(b), int(r.integers(0, n - b + 1))) for b in r.integers(0, n, 6)]
    for b, c in ranges:
        print(f"[rows] {name} begin {b} count {c}", flush=True)
        got = tree.DecompressRows(blob, b, c)
        assert got.nbytes == c * x.itemsize
        assert got.to_bytes() == x[b:b + c].tobytes(), (b, c)
    for b, c in [(n, 1), (n - 5, 6), (n + 10, 0)]:
        with pytest.raises(g.DdjError) as e:
            tree.DecompressRows(blob, b, c)
        assert e.value.status == 1
    assert tree.Decompress(blob).to_bytes() == x.tobytes()


def test_decompress_rows_long_delta_column(ctx):
    """A 16.8M-row delta column: ranges deep inside sum millions of packed
    deltas before them."""
    n = 1 << 24
    x = D.timestamps(n, seed=31)
    tree = g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_INT)
    blob = tree.Compress(ctx.upload(x))
    for b, c in [(0, 10), (n - 10, 10), (7_777_777, 1_000_000), (1 << 23, 4096), (5, n - 5)]:
        print(f"[rows] long delta begin {b} count {c}", flush=True)
        assert tree.DecompressRows(blob, b, c).to_bytes() == x[b:b + c].tobytes()
