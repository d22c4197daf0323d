"""Pins the oracle's interleaved (32-lane) FL layout and random-access fetch
(no GPU).

The 32-lane stream is checked two independent ways: lane by lane it must be
the contiguous FL stream (itself pinned by the LSB-first closed form in
test_oracle.py) of the lane's own values, and bit by bit it must follow the
layout's definition built here with plain integer arithmetic.  The fetch is
checked against plain indexing of the column.
"""
import numpy as np
import pytest

from oracle import oracle as O

TYPES = [(O.D_CHAR, np.int8), (O.D_SHORT, np.int16), (O.D_INT, np.int32), (O.D_TIME, np.int64)]


def column(npdt, n, bits, seed):
    r = np.random.default_rng(seed)
    hi = (1 << bits) - 1
    u = r.integers(0, hi, n, dtype=np.uint64, endpoint=True) if bits < 64 else r.integers(0, 2**63, n, dtype=np.uint64)
    return u.astype(np.dtype(npdt).newbyteorder("=").str.replace("i", "u")).view(npdt)


def lanes_by_bits(x, bits):
    """The 32-lane stream from its definition: row r of block k = r // (32W),
    lane l = r % 32, slot i = (r % (32W)) // 32 occupies lane bits
    [i*b, i*b + b), lane bit p being bit p % W of word k*32*b + l + 32*(p // W)."""
    W = 8 * x.itemsize
    n = x.size
    words = (n + 32 * W - 1) // (32 * W) * 32 * bits
    out = [0] * words
    ux = x.view(x.dtype.str.replace("i", "u")).astype(np.uint64)
    for r in range(n):
        v = int(ux[r]) & ((1 << bits) - 1)
        k, rin = divmod(r, 32 * W)
        l, i = rin % 32, rin // 32
        for t in range(bits):
            if (v >> t) & 1:
                p = i * bits + t
                out[k * 32 * bits + l + 32 * (p // W)] |= 1 << (p % W)
    return np.array(out, dtype=np.uint64).astype(x.dtype.str.replace("i", "u")).view(x.dtype)


@pytest.mark.parametrize("dt,npdt", TYPES)
@pytest.mark.parametrize("n", [1, 31, 33, 300, 1000])
def test_lanes_stream_is_bitwise_definition(dt, npdt, n):
    W = 8 * np.dtype(npdt).itemsize
    for bits in sorted({1, 3, W // 2 + 1, W - 1, W}):
        x = column(npdt, n, bits, seed=n * 100 + bits)
        got = O.fl_pack_lanes(dt, bits, x)
        assert got.tobytes() == lanes_by_bits(x, bits).tobytes(), (n, bits)
        assert np.array_equal(O.fl_unpack_lanes(dt, bits, got, n), x)


@pytest.mark.parametrize("dt,npdt", TYPES)
def test_each_lane_is_a_contiguous_fl_stream(dt, npdt):
    """Lane l of block k, read at stride 32, is the contiguous FL stream of
    values k*32*W + l + 32*i (i < W), the missing ones zero."""
    W = 8 * np.dtype(npdt).itemsize
    n = 2 * 32 * W + 77  # two whole blocks and a partial one
    for bits in (2, W // 2 + 3, W):
        x = column(npdt, n, bits, seed=bits)
        s = O.fl_pack_lanes(dt, bits, x)
        blocks = (n + 32 * W - 1) // (32 * W)
        padded = np.zeros(blocks * 32 * W, dtype=npdt)
        padded[:n] = x
        for k in range(blocks):
            for lane in range(32):
                vals = padded[k * 32 * W + lane:(k + 1) * 32 * W:32]
                assert vals.size == W
                words = s[k * 32 * bits + lane:(k + 1) * 32 * bits:32]
                assert words.tobytes() == O.fl_pack(dt, bits, vals).tobytes(), (bits, k, lane)


def test_lanes_stream_of_whole_blocks_is_the_reference_length():
    """For n a multiple of 32*W the stream is ceil(b*n/W) words, the
    comprLength the reference passes (afl_encoding_benchmark.cpp:36-47)."""
    for bits in (1, 10, 17, 32):
        n = 4 * 32 * 32
        x = column(np.int32, n, bits, seed=bits)
        assert O.fl_pack_lanes(O.D_INT, bits, x).size == (bits * n + 31) // 32


@pytest.mark.parametrize("lanes", [1, 32])
@pytest.mark.parametrize("dt,npdt", TYPES)
def test_fetch_is_indexing(lanes, dt, npdt):
    W = 8 * np.dtype(npdt).itemsize
    n = 3 * 32 * W + 5
    r = np.random.default_rng(7)
    for bits in (1, 5, W - 1, W):
        x = column(npdt, n, bits, seed=bits + lanes)
        stream = O.fl_pack(dt, bits, x) if lanes == 1 else O.fl_pack_lanes(dt, bits, x)
        rows = np.concatenate([[0, n - 1], r.integers(0, n, 500)]).astype(np.uint64)
        assert np.array_equal(O.fl_fetch(dt, bits, lanes, stream, n, rows), x[rows.astype(np.int64)])
