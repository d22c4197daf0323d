"""The oracle's FL restatement against the reference's own host loops
(run_afl_compress_cpu<T,1> / run_afl_decompress_cpu<T,1>,
src/compression/afl/afl_gpu.cu:6-65) compiled from the reference sources
into oracle/_ref/ by oracle/ref_build.py.  Widths b < W only: at b = W the
reference's host mask NBITSTOMASK(W) is undefined behaviour
(core/cuda_macros.cuh:6)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libddj_ref.so")


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (the reference sources are only present in the build container)")
    L = ctypes.CDLL(REF)
    L.ref_fl_pack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_ulong, ctypes.c_void_p,
                              ctypes.c_ulong]
    L.ref_fl_unpack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_ulong, ctypes.c_void_p]
    return L


@pytest.mark.parametrize("dt", [O.D_INT, O.D_TIME, O.D_SHORT, O.D_CHAR])
def test_fl_pack_matches_reference_loop(ref, dt):
    r = np.random.default_rng(11)
    npdt = O.NP_DTYPE[dt]
    es = np.dtype(npdt).itemsize
    W = 8 * es
    for b in range(1, W):
        for n in (1, 7, W, W + 3, 8 * W + 1, 3000):
            x = r.integers(0, 1 << b, n, dtype=np.int64).astype(npdt)
            words = (b * n + W - 1) // W
            out = np.zeros(words, npdt)
            assert ref.ref_fl_pack(es, b, x.ctypes.data, n, out.ctypes.data, words) == 0
            assert out.tobytes() == O.fl_pack(dt, b, x).tobytes(), (b, n)
            back = np.zeros(n, npdt)
            assert ref.ref_fl_unpack(es, b, out.ctypes.data, n, back.ctypes.data) == 0
            assert np.array_equal(back, O.fl_unpack(dt, b, out, n)), (b, n)
