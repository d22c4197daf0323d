"""The benchmark's columns at their full size, checked against the oracle
byte for byte: the four single-scheme int32 trees at 2^26 rows (past row
67,107,840, where the reference's default launch shape stops covering the
column) and at 64 * 10^6 rows, and the two GFC float price columns at 2^26
rows.  Blobs are compared through SHA-256 digests; each decompresses to the
input."""
import hashlib

import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g
from gputimeseriescompressionoptimizer_amd import datasets as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INT_TREES = {
    "timestamps delta+afl": (lambda n: D.timestamps(n, seed=11), [g.DELTA, g.AFL, g.NONE]),
    "step function rle+afl": (lambda n: D.step_function(n, seed=12), [g.RLE, g.AFL, g.NONE, g.AFL, g.NONE]),
    "zipf codes dict": (lambda n: D.zipf_codes(n, seed=13), [g.DICT, g.NONE, g.NONE]),
    "outliers patch+afl": (lambda n: D.normal_with_outliers(n, seed=14), [g.PATCH, g.AFL, g.NONE, g.AFL, g.NONE]),
}
CASES = [(name, n) for n in (1 << 26, 64_000_000) for name in INT_TREES]


def check(ctx, x, dtype, path):
    print(f"[full size] {x.dtype} {x.size} rows, tree {path}", flush=True)
    tree = g.CompressionTree(ctx, path, dtype)
    buf = ctx.upload(x)
    blob = tree.Compress(buf)
    got = blob.to_bytes()
    want = O.compress(O.tree_from_path(path, dtype), x)
    assert len(got) == len(want)
    assert hashlib.sha256(got).digest() == hashlib.sha256(want).digest()
    del want, got
    back = tree.Decompress(blob).to_bytes()
    assert back == x.tobytes()


@pytest.mark.parametrize("name,n", CASES, ids=[f"{a}-{b}" for a, b in CASES])
def test_int_tree_full_size(ctx, name, n):
    make, path = INT_TREES[name]
    check(ctx, make(n), g.D_INT, path)


@pytest.mark.parametrize("dtype", [g.D_DOUBLE, g.D_FLOAT])
def test_gfc_price_walk_full_size(ctx, dtype):
    x = D.price_walk_f64(1 << 26, seed=21) if dtype == g.D_DOUBLE else D.price_walk_f32(1 << 26, seed=22)
    check(ctx, x, dtype, [g.GFC, g.NONE, g.NONE])
