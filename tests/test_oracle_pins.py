"""Pins the oracle's patch, outside-patch, const, scale, floatToInt, AFL-float
and GFC encoders to hand-computed known-answer vectors (tests/pin_vectors.py),
and checks each vector's decode restores the input (except where the
reference itself is lossy, which the vector's comment states)."""
import numpy as np
import pytest

from oracle import oracle as O
from pin_vectors import VECTORS

LOSSY = {"afl float32 negative zero counts as positive"}


@pytest.mark.parametrize("name,enc,dtype,x,want", VECTORS, ids=[v[0] for v in VECTORS])
def test_oracle_matches_hand_vector(name, enc, dtype, x, want):
    got = O.encode(enc, dtype, x)
    assert len(got) == len(want)
    for i, (g, w) in enumerate(zip(got, want)):
        assert g == w, f"output {i}: {g.hex()} != {w.hex()}"
    back = O.decode(enc, dtype, got)
    if name in LOSSY:
        assert back == np.abs(x).tobytes()
    else:
        assert back == x.tobytes()


def test_vectors_cover_every_pinned_codec():
    encs = {v[1] for v in VECTORS}
    assert encs >= {O.PATCH, O.OUTSIDE_PATCH, O.CONST, O.SCALE, O.FLOAT2INT, O.AFL, O.GFC}
