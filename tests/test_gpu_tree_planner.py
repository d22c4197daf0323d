"""Cascades and planner on the GPU against the oracle.

* trees: the serialized blob of every tree is byte-identical to the oracle's,
  decompresses to the input (also from a misaligned address), node ratios
  match, and the size predictor equals the blob size;
* planner: the exhaustive phase-1 planner picks the same tree with the same
  ratio from the same number of candidates; the 30-path brute force picks the
  same tree; the stateful optimizer emits the oracle's blobs part after part.
"""
import struct

import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g
from gputimeseriescompressionoptimizer_amd import datasets as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INT_COLUMNS = {
    "timestamps": lambda n: D.timestamps(n, seed=21),
    "step": lambda n: D.step_function(n, seed=22, mean_run=50),
    "zipf": lambda n: D.zipf_codes(n, seed=23),
    "outliers": lambda n: D.normal_with_outliers(n, seed=24),
}

# trees hand-built in the reference's tests (tree/compression_tree_unittest.cpp)
HAND_TREES = [
    ([g.DELTA, g.NONE], g.D_INT),
    ([g.DELTA, g.SCALE, g.NONE], g.D_INT),
    ([g.DICT, g.SCALE, g.DELTA, g.NONE, g.DELTA, g.NONE], g.D_INT),
    ([g.SCALE, g.DELTA, g.PATCH, g.AFL, g.NONE, g.AFL, g.NONE], g.D_INT),
    ([g.DELTA, g.AFL, g.NONE], g.D_INT),
    ([g.RLE, g.AFL, g.NONE, g.DELTA, g.AFL, g.NONE], g.D_INT),
    ([g.CONST, g.SCALE, g.AFL, g.NONE], g.D_INT),
    ([g.GFC, g.NONE, g.NONE], g.D_DOUBLE),
    ([g.GFC, g.NONE, g.NONE], g.D_FLOAT),
    ([g.FLOAT_TO_INT, g.DELTA, g.AFL, g.NONE], g.D_DOUBLE),
    ([g.FLOAT_TO_INT, g.SCALE, g.DELTA, g.AFL, g.NONE], g.D_FLOAT),
    ([g.PATCH, g.CONST, g.NONE, g.FLOAT_TO_INT, g.RLE, g.CONST, g.NONE, g.SCALE, g.AFL, g.NONE], g.D_FLOAT),
    ([g.AFL, g.NONE], g.D_FLOAT),
    ([g.DELTA, g.AFL, g.NONE], g.D_TIME),
    ([g.DICT, g.AFL, g.NONE, g.AFL, g.NONE], g.D_SHORT),
    ([g.RLE, g.NONE, g.NONE], g.D_CHAR),
]


def column(dtype, n, seed=0):
    if dtype == g.D_DOUBLE:
        return D.price_walk_f64(n, seed=seed)
    if dtype == g.D_FLOAT:
        return D.price_walk_f32(n, seed=seed)
    if dtype == g.D_TIME:
        return D.timestamps(n, seed=seed).astype(np.int64) * 1000
    if dtype == g.D_SHORT:
        return (D.zipf_codes(n, seed=seed) * 3).astype(np.int16)
    if dtype == g.D_CHAR:
        return np.repeat(np.arange(n // 9 + 1) % 100, 9)[:n].astype(np.int8)
    return D.step_function(n, seed=seed, mean_run=20)


def check_tree(ctx, path, dtype, x):
    ref_tree = O.tree_from_path(path, dtype)
    ref_tree.fix()
    want = O.compress(ref_tree, x)
    tree = g.CompressionTree(ctx, path, dtype)
    blob = tree.Compress(ctx.upload(x))
    got = blob.to_bytes()
    assert got == want, (path, dtype, len(got), len(want))
    # ratios of every node in pre-order
    ref_ratios = []

    def walk(nd):
        ref_ratios.append(nd.ratio)
        for ch in nd.children:
            walk(ch)

    walk(ref_tree)
    assert np.allclose(tree.node_ratios[: len(path)], ref_ratios[: len(path)], rtol=0, atol=1e-12)
    back = tree.Decompress(blob).to_bytes()
    assert back == O.decompress(want)
    return got


@pytest.mark.parametrize("name", sorted(INT_COLUMNS))
def test_generated_paths_match_oracle(ctx, name):
    """All 30 PathGenerator paths on each int column (compression_tree_unittest.cpp:344-536)."""
    x = INT_COLUMNS[name](6000)
    for path in O.generate_paths():
        got = check_tree(ctx, path, g.D_INT, x)
        assert np.array_equal(np.frombuffer(O.decompress(got), np.int32), x)
        # predicted size == actual (compression_tree_unittest.cpp:238-258)
        t = g.CompressionTree(ctx, path, g.D_INT)
        assert t.GetPredictedSizeAfterCompression(ctx.upload(x)) == len(got)


@pytest.mark.parametrize("path,dtype", HAND_TREES, ids=[str(i) for i in range(len(HAND_TREES))])
def test_hand_trees(ctx, path, dtype):
    for n in (1, 40, 5000, 70001):
        x = column(dtype, n, seed=n)
        got = check_tree(ctx, path, dtype, x)
        if g.FLOAT_TO_INT not in path:
            assert O.decompress(got) == x.tobytes()


def test_decompress_misaligned_blob(ctx):
    x = D.timestamps(20000, seed=5)
    for path in ([g.DELTA, g.NONE], [g.DELTA, g.AFL, g.NONE], [g.PATCH, g.NONE, g.NONE], [g.DICT, g.NONE, g.NONE],
                 [g.RLE, g.NONE, g.NONE]):
        tree = g.CompressionTree(ctx, path, g.D_INT)
        blob = tree.Compress(ctx.upload(x)).to_bytes()
        for off in (1, 2, 3, 5):
            big = ctx.upload(b"\xab" * off + blob + b"\xcd" * 7)
            out = tree.Decompress(big.view(off, len(blob)))
            assert np.array_equal(out.to_numpy(np.int32), x), (path, off)


@pytest.mark.parametrize("dtype", [g.D_CHAR, g.D_SHORT, g.D_INT, g.D_TIME, g.D_FLOAT, g.D_DOUBLE])
def test_merge_reads_rows_at_any_byte_offset(ctx, dtype):
    """Stencil outputs stored unchanged ([x, none, none] trees) sit after
    metadata of odd length; the merge stages them from any byte address,
    for every row width and every blob misalignment."""
    for n in (3001, 70_003):
        x = column(dtype, n, seed=n)
        for path in ([g.PATCH, g.NONE, g.NONE], [g.DICT, g.NONE, g.NONE], [g.CONST, g.NONE]):
            tree = g.CompressionTree(ctx, path, dtype)
            blob = tree.Compress(ctx.upload(x)).to_bytes()
            for off in (0, 1, 2, 3, 6):
                big = ctx.upload(b"\xab" * off + blob + b"\xcd" * 7)
                out = tree.Decompress(big.view(off, len(blob)))
                assert out.to_bytes() == x.tobytes(), (path, dtype, n, off)


@pytest.mark.parametrize("dtype", [g.D_CHAR, g.D_SHORT, g.D_INT, g.D_TIME, g.D_FLOAT, g.D_DOUBLE])
def test_dict_rows_written_in_place(ctx, dtype):
    """[dict, none, none]: the split's writing pass runs at blob assembly and
    writes the other rows straight into the blob (any byte offset); blobs
    equal the oracle's, also with no other rows and with no frequent rows."""
    for n in (1, 5000, 70_001):
        check_tree(ctx, [g.DICT, g.NONE, g.NONE], dtype, column(dtype, n, seed=n + 1))
    npdt = g.NP_DTYPE[dtype]
    check_tree(ctx, [g.DICT, g.NONE, g.NONE], dtype, np.array([3, 1, 3, 2, 1, 4] * 2000, dtype=npdt))
    check_tree(ctx, [g.DICT, g.NONE, g.NONE], dtype, np.arange(9000).astype(npdt))


@pytest.mark.parametrize("dtype", [g.D_CHAR, g.D_SHORT, g.D_INT, g.D_TIME])
def test_fl_leaves_packed_in_place(ctx, dtype):
    """Integer AFL leaves are packed straight into the blob at whatever byte
    offset the pieces before them leave (odd stencil lengths, 1- and 2-byte
    metadata); blobs equal the oracle's for every width."""
    npdt = g.NP_DTYPE[dtype]
    r = np.random.default_rng(dtype)
    for n in (1, 33, 4097, 70_001):
        cols = [r.integers(-50, 50, n).astype(npdt), np.cumsum(r.integers(0, 3, n)).astype(npdt),
                r.integers(0, 1 << (4 * np.dtype(npdt).itemsize), n).astype(npdt)]
        for x in cols:
            for path in ([g.AFL, g.NONE], [g.DELTA, g.AFL, g.NONE], [g.PATCH, g.AFL, g.NONE, g.AFL, g.NONE],
                         [g.RLE, g.AFL, g.NONE, g.AFL, g.NONE]):
                check_tree(ctx, path, dtype, x)


def test_corrupt_blob_is_an_error(ctx):
    x = D.timestamps(5000, seed=6)
    blob = g.CompressionTree(ctx, [g.DELTA, g.AFL, g.NONE], g.D_INT).Compress(ctx.upload(x)).to_bytes()
    for cut in (3, 9, len(blob) // 2, len(blob) - 1):
        with pytest.raises(g.DdjError):
            g.CompressionTree(ctx).Decompress(ctx.upload(blob[:cut]))
    bad = bytearray(blob)
    bad[0] = 77  # unknown encoding id in the root header
    with pytest.raises(g.DdjError):
        g.CompressionTree(ctx).Decompress(ctx.upload(bytes(bad)))


PLAN_CASES = [
    ("zipf", lambda: D.zipf_codes(3000, seed=31), g.D_INT),
    ("outliers", lambda: D.normal_with_outliers(3000, seed=32), g.D_INT),
    ("timestamps", lambda: D.timestamps(1500, seed=33), g.D_INT),
    ("step", lambda: D.step_function(1200, seed=34, mean_run=30), g.D_INT),
    ("walk_f64", lambda: D.price_walk_f64(3000, seed=35), g.D_DOUBLE),
    ("walk_f32", lambda: D.price_walk_f32(2000, seed=36), g.D_FLOAT),
    ("short", lambda: (D.zipf_codes(2500, seed=37) * 7).astype(np.int16), g.D_SHORT),
]


@pytest.mark.parametrize("name,gen,dtype", PLAN_CASES, ids=[c[0] for c in PLAN_CASES])
def test_exhaustive_planner_matches_oracle(ctx, name, gen, dtype):
    x = gen()
    print(f"[plan] {name} rows={x.size}", flush=True)
    tree, ratio, ncand = g.CompressionOptimizer(ctx).PlanExhaustive(ctx.upload(x), dtype)
    best, size, nc = O.plan_exhaustive(x, dtype)
    assert ncand == nc
    assert tree.path == best.preorder()
    assert ratio == best.ratio


@pytest.mark.parametrize("name", sorted(INT_COLUMNS))
def test_optimize30_matches_oracle(ctx, name):
    x = INT_COLUMNS[name](4000)
    tree = g.CompressionOptimizer(ctx).OptimizeTree(ctx.upload(x), g.D_INT)
    assert tree.path == O.optimize_tree(x, O.D_INT).preorder()


def test_compress_data_parts_match_oracle(ctx):
    """CompressData over successive parts: phase 1 on the first part and every
    10th, phase 2 statistics and phase 3 repair in between."""
    opt = g.CompressionOptimizer(ctx)
    ref = O.Optimizer()
    parts = [D.timestamps(4000, seed=40 + i) for i in range(4)] + \
            [D.zipf_codes(4000, seed=50 + i) for i in range(4)] + [D.step_function(4000, seed=60 + i) for i in range(4)]
    for i, p in enumerate(parts):
        blob = opt.CompressData(ctx.upload(p), g.D_INT).to_bytes()
        want = ref.compress_data(p, O.D_INT)
        assert opt.GetOptimalTree() == ref.tree.preorder(), i
        assert blob == want, i
        if O.decodable(ref.tree):
            assert np.array_equal(np.frombuffer(O.decompress(blob), np.int32), p), i


def _walk_planner_nodes(ctx, data: bytes, et: int, dt: int, level: int, seen: list):
    """Visit every (input, continuation) pair the exhaustive planner encodes
    (compression_optimizer.cpp:88-140) and compare the kernel's statistics and
    encoder outputs with the oracle's at that node."""
    n = len(data) // O.DTYPE_SIZE[dt]
    buf = ctx.upload(data) if data else ctx.alloc(0)
    o = O.stats(dt, data[: n * O.DTYPE_SIZE[dt]])
    if n:
        s = ctx.stats(buf, dt)
        bits = lambda v: struct.pack("<d", v)  # NaN minima compare by bit pattern
        assert (bits(s.min), bits(s.max), s.sorted, s.rlMetric, s.size) == \
               (bits(o.min), bits(o.max), o.sorted, o.rlMetric, o.size), (seen, dt, n)
    conts = O.get_continuations(et, dt, o, level)
    for c in conts:
        if c == O.NONE:
            continue
        want = O.encode(c, dt, data)
        got = [p.to_bytes() for p in g.Encoding(ctx, c).Encode(buf, dt)]
        assert got[: len(want)] == want, (seen + [c], dt, [len(x) for x in got], [len(x) for x in want])
        nres, rtypes = O.encoding_info(c, dt)
        for i in range(nres):
            _walk_planner_nodes(ctx, want[1 + i], c, rtypes[i], level + 1, seen + [(c, i)])


@pytest.mark.parametrize("name,gen,dtype", [c for c in PLAN_CASES if c[2] in (g.D_DOUBLE, g.D_FLOAT)],
                         ids=[c[0] for c in PLAN_CASES if c[2] in (g.D_DOUBLE, g.D_FLOAT)])
def test_planner_nodes_match_oracle(ctx, name, gen, dtype):
    """Every intermediate buffer of the floating-point planner cases: same
    statistics, same encoder outputs."""
    _walk_planner_nodes(ctx, O.as_bytes(gen()), O.NONE, dtype, 0, [])


@pytest.mark.parametrize("dtype", [g.D_FLOAT, g.D_DOUBLE])
def test_gfc_payload_written_in_place(ctx, dtype):
    """Float GFC payloads stored unchanged are encoded straight into the blob
    at whatever byte offset the pieces before them leave (a patch stencil of
    odd length before them); blobs equal the oracle's."""
    for n in (1, 33, 4097, 70_001):
        x = column(dtype, n, seed=n + 3)
        check_tree(ctx, [g.GFC, g.NONE, g.NONE], dtype, x)
        check_tree(ctx, [g.PATCH, g.GFC, g.NONE, g.NONE, g.GFC, g.NONE, g.NONE], dtype, x)
