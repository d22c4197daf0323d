"""The HIP library against the hand-computed known-answer vectors of
tests/pin_vectors.py (patch, outside patch, const, scale, floatToInt,
AFL-float, GFC): every output byte-identical, and the decode equal to the
input (or to the reference's lossy result where the vector says so)."""
import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g
from pin_vectors import VECTORS

pytestmark = pytest.mark.gpu

LOSSY = {"afl float32 negative zero counts as positive"}


@pytest.mark.parametrize("name,enc,dtype,x,want", VECTORS, ids=[v[0] for v in VECTORS])
def test_gpu_matches_hand_vector(ctx, name, enc, dtype, x, want):
    outs = g.Encoding(ctx, enc).Encode(ctx.upload(x), dtype)
    got = [o.to_bytes() for o in outs]
    assert len(got) == len(want)
    for i, (a, b) in enumerate(zip(got, want)):
        assert a == b, f"output {i}: {a.hex()} != {b.hex()}"
    back = g.Encoding(ctx, enc).Decode(outs, dtype).to_bytes()
    assert back == (np.abs(x).tobytes() if name in LOSSY else x.tobytes())
