"""ddj_tree_decompress_noalloc (no device allocation, fixed per-GPU scratch)
against ddj_tree_decompress on the same blobs: two calls in a row, the
second with a smaller column, into the same caller buffer."""
import ctypes

import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g
from gputimeseriescompressionoptimizer_amd import datasets as D

pytestmark = pytest.mark.gpu

CASES = {
    "delta-afl int32 timestamps": (g.D_INT, [g.DELTA, g.AFL, g.NONE], lambda n, s: D.timestamps(n, seed=s)),
    "delta-afl int64 walk": (g.D_TIME, [g.DELTA, g.AFL, g.NONE],
                             lambda n, s: np.cumsum(np.random.default_rng(s).integers(-9, 10, n)).astype(np.int64)),
    "afl int16": (g.D_SHORT, [g.AFL, g.NONE],
                  lambda n, s: np.random.default_rng(s).integers(0, 3000, n).astype(np.int16)),
    "none float64": (g.D_DOUBLE, [g.NONE], lambda n, s: D.price_walk_f64(n, seed=s)),
}


def noalloc(ctx, blob, out_dev, out_cap):
    ob, dt = ctypes.c_size_t(0), ctypes.c_int(-1)
    st = ctx._L.ddj_tree_decompress_noalloc(ctx._h, blob.dev or None, blob.nbytes, out_dev, out_cap,
                                            ctypes.byref(ob), ctypes.byref(dt))
    return st, ob.value, dt.value


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_matches_tree_decompress(ctx, name):
    dtype, path, make = CASES[name]
    es = np.dtype(g.NP_DTYPE[dtype]).itemsize
    sizes = (1_000_003, 4097)
    out = ctx.alloc(sizes[0] * es)
    for i, n in enumerate(sizes):
        print(f"[noalloc] {name} call {i} rows {n}", flush=True)
        x = make(n, 100 + i)
        tree = g.CompressionTree(ctx, path, dtype)
        blob = tree.Compress(ctx.upload(x))
        want = tree.Decompress(blob).to_bytes()
        assert want == x.tobytes()
        st, nbytes, dt = noalloc(ctx, blob, out.dev, out.nbytes)
        assert st == 0, ctx._L.ddj_last_error(ctx._h)
        assert (nbytes, dt) == (len(want), dtype)
        assert out.to_bytes()[:nbytes] == want


def test_noalloc_small_and_empty_columns(ctx):
    out = ctx.alloc(4 * 33)
    for n in (1, 2, 33):
        x = D.timestamps(n, seed=n)
        tree = g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_INT)
        blob = tree.Compress(ctx.upload(x))
        st, nbytes, dt = noalloc(ctx, blob, out.dev, out.nbytes)
        assert st == 0 and nbytes == x.nbytes, (n, ctx._L.ddj_last_error(ctx._h))
        assert out.to_bytes()[:nbytes] == x.tobytes()


def test_noalloc_rejects_what_it_cannot_do(ctx):
    x = D.step_function(10_000, seed=1)
    out = ctx.alloc(x.nbytes)
    rle = g.CompressionTree(ctx, [g.RLE, g.AFL, g.NONE, g.AFL, g.NONE], g.D_INT).Compress(ctx.upload(x))
    assert noalloc(ctx, rle, out.dev, out.nbytes)[0] == 2  # DDJ_EUNSUPPORTED
    fl = g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_INT).Compress(ctx.upload(x))
    assert noalloc(ctx, fl, out.dev, out.nbytes - 4)[0] == 1  # DDJ_EINVAL: capacity
    # a delta column past the scratch's 65,535 tiles (268,431,361 rows)
    n = (1 << 28) + 2
    big = np.arange(n, dtype=np.int64).astype(np.int8)  # every delta 1: one bit per row
    dbuf = ctx.upload(big)
    blob = g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_CHAR).Compress(dbuf)
    del dbuf
    obig = ctx.alloc(n)
    assert noalloc(ctx, blob, obig.dev, obig.nbytes)[0] == 1  # DDJ_EINVAL: too long for the scratch
    # the library still works afterwards
    assert g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_INT).Decompress(fl).to_bytes() == x.tobytes()
