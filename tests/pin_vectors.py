"""Hand-computed known-answer vectors for the codecs whose outputs no
closed-form test of the oracle fixes on its own.

Each vector is (name, encoding, dtype, input, expected outputs): the expected
bytes are written out literally (or packed from literal numbers) from the
reference's rules as cited, worked by hand, never computed by the oracle's
code.  tests/test_oracle_pins.py checks the CPU oracle against them and
tests/test_gpu_pins.py the HIP library.  Each vector is chosen so that a
plausible mistake (a `<=` for a `<`, floor for truncation, rounding in double
instead of T, the wrong predictor, the wrong polarity) changes the bytes.
"""
import struct

import numpy as np

from oracle import oracle as O

INT_MIN, INT_MAX = -2**31, 2**31 - 1


def i32(*v):
    return np.array(v, dtype=np.int32)


def le32(*v):
    return struct.pack("<%di" % len(v), *v)


def le64(*v):
    return struct.pack("<%dq" % len(v), *v)


def h(s: str) -> bytes:
    return bytes.fromhex(s.replace(" ", ""))


VECTORS = []


def vec(name, enc, dtype, x, outs):
    VECTORS.append((name, enc, dtype, x, outs))


# ------------------------------------------------------------------ patch
# LowerPatchEncoding::GetOperator (compression/patch/lower_patch_encoding.cu:22-31):
# _min/_max are doubles, dist = T(max - min), thr = T(max - 0.1 * dist); the
# stencil is x < thr (stencil_operators.hpp:66-75).  out0 = packed stencil
# (util/stencil/stencil.cu:14-33: byte0 = n % 8, LSB first, padding = last
# row), out1 = x[s], out2 = x[!s].
#
# min 0, max 15: dist 15, 15 - 1.5 = 13.5 -> 13 (so 13 is NOT below it;
# rounding would give 14, `<=` would move 13).  s = 1 1 0 0 0 1 + padding 1 1.
vec("patch int32 fractional threshold", O.PATCH, O.D_INT, i32(0, 12, 13, 14, 15, 5),
    [h("06 e3"), le32(0, 12, 5), le32(13, 14, 15)])
# min -20, max -5: -5 - 1.5 = -6.5 truncates to -6 (floor would give -7).
# s = -20<-6, -5, -6, -7<-6 -> 1 0 0 1, padding 1.
vec("patch int32 negative threshold truncates toward zero", O.PATCH, O.D_INT, i32(-20, -5, -6, -7),
    [h("04 f9"), le32(-20, -7), le32(-5, -6)])
# float: thr = float(1.0 - 0.1 * 1.0) = float(0.9) = 0.89999998; x = 0.9f is
# that same float, so it is not below (in double 0.9f < 0.9 would hold).
vec("patch float32 threshold rounded to T", O.PATCH, O.D_FLOAT, np.array([0.0, 1.0, 0.9, 0.8], np.float32),
    [h("04 f9"), np.array([0.0, 0.8], np.float32).tobytes(), np.array([1.0, 0.9], np.float32).tobytes()])
# int16: dist = 100 - 0 = 100, thr = 90: 89 below, 90 not.
vec("patch int16 at the threshold", O.PATCH, O.D_SHORT, np.array([0, 90, 89, 100], np.int16),
    [h("04 05"), np.array([0, 89], np.int16).tobytes(), np.array([90, 100], np.int16).tobytes()])

# OutsidePatchEncoding::GetOperator (outside_patch_encoding.cu:22-32): low =
# T(min + 0.1 dist), high = T(max - 0.1 dist); s = !(x > high || x < low)
# (OutsideOperator, stencil_operators.hpp:14-25): the rows inside [low, high]
# go to out1.
# min 0, max 100: low 10, high 90: 0 out, 10 in, 11 in, 90 in, 91 out, 100 out, 50 in.
vec("outside patch int32 both bounds inclusive", O.OUTSIDE_PATCH, O.D_INT, i32(0, 10, 11, 90, 91, 100, 50),
    [h("07 ce"), le32(10, 11, 90, 50), le32(0, 91, 100)])
# min 0, max 15: low = T(1.5) = 1, high = T(13.5) = 13.
vec("outside patch int32 fractional bounds", O.OUTSIDE_PATCH, O.D_INT, i32(0, 1, 13, 14, 15),
    [h("05 06"), le32(1, 13), le32(0, 14, 15)])
# min -15, max 0: low = T(-13.5) = -13, high = T(-1.5) = -1 (truncation toward zero).
vec("outside patch int32 negative bounds", O.OUTSIDE_PATCH, O.D_INT, i32(-15, -14, -13, -1, 0),
    [h("05 0c"), le32(-13, -1), le32(-15, -14, 0)])

# ------------------------------------------------------------------ const
# ConstEncoding::Encode (compression/const/const_encoding.cu:19-25, 43-78):
# c = the most frequent value; stencil = x != c; metadata = c ++ packed
# stencil; data = x[x != c].
# 7 7 3 7 9 7 7 7 7 7: s = 0 0 1 0 1 0 0 0 | 0 0 + padding 0.
vec("const stencil marks the rows that differ", O.CONST, O.D_INT, i32(7, 7, 3, 7, 9, 7, 7, 7, 7, 7),
    [le32(7) + h("02 14 00"), le32(3, 9)])
# 5 5 5 1: s = 0 0 0 1, padding repeats the last row (1): 0b11111000.
vec("const stencil padding repeats the last row", O.CONST, O.D_INT, i32(5, 5, 5, 1),
    [le32(5) + h("04 f8"), le32(1)])
# all constant: no data rows, stencil all zero.
vec("const on a constant column", O.CONST, O.D_SHORT, np.full(9, -3, np.int16),
    [struct.pack("<h", -3) + h("01 00 00"), b""])

# ------------------------------------------------------------------ scale
# ScaleEncoding::Encode (compression/scale/scale_encoding.cu:30-60): m = min,
# but m = 0 when min < 0 and numeric_limits<T>::max() - max < |min|.  For int
# and long the subtraction is in T and wraps when max < 0; short and char
# promote to int; |INT_MIN| stays INT_MIN.  data = x - m (wrapping).
vec("scale keeps a negative min that cannot overflow", O.SCALE, O.D_INT, i32(-5, 10, 0),
    [le32(-5), le32(0, 15, 5)])
vec("scale overflow guard fires on a wide range", O.SCALE, O.D_INT, i32(-2000000000, 2000000000),
    [le32(0), le32(-2000000000, 2000000000)])
# INT_MAX - (-5) wraps to -2147483644 < 100: the guard fires for every
# all-negative int column.
vec("scale on an all-negative int32 column: the wrapped guard fires", O.SCALE, O.D_INT, i32(-100, -5, -50),
    [le32(0), le32(-100, -5, -50)])
vec("scale with INT_MIN: |INT_MIN| is negative, min kept", O.SCALE, O.D_INT, i32(INT_MIN, 0),
    [le32(INT_MIN), le32(0, INT_MIN)])
# int16 promotes: 32767 + 5 = 32772 >= 100, min kept.
vec("scale on an all-negative int16 column: no wrap", O.SCALE, O.D_SHORT, np.array([-100, -5], np.int16),
    [struct.pack("<h", -100), struct.pack("<2h", 0, 95)])
vec("scale on an all-negative int64 column: the wrapped guard fires", O.SCALE, O.D_TIME,
    np.array([-7, -3], np.int64), [le64(0), le64(-7, -3)])

# -------------------------------------------------------------- floatToInt
# FloatEncoding::Encode (compression/float/float_encoding.cu:17-50) with
# FloatingPointToIntegerOperator (util/transform/transform_operators.hpp:119-127):
# p = max precision; y = int(lrint(x * 10^p)) in double (round half to even);
# p = 10 and the raw bytes when p >= 10 or min * 10^p, max * 10^p (in T) leave
# the int range.  Metadata = int32 p.
#
# 100.01f = 100.01000213623046875 -> 10001.000213... -> 10001;
# 99.99f = 99.98999786376953125 -> 9998.9997... -> 9999 (truncation: 9998);
# 100.1f = 100.09999847412109375 -> 10009.9998... -> 10010.
vec("floatToInt float32 prices to cents by lrint in double", O.FLOAT2INT, O.D_FLOAT,
    np.array([100.01, 99.99, 100.1], np.float32), [le32(2), le32(10001, 9999, 10010)])
# -2.5 has precision 1: -25 exactly; 0.25 has precision 2 -> p = 2: -250, 25.
vec("floatToInt float32 negative values", O.FLOAT2INT, O.D_FLOAT, np.array([-2.5, 0.25], np.float32),
    [le32(2), le32(-250, 25)])
# 1e-12f never satisfies round(x*e)/e == x for e < 10^10: p = 10, raw bytes.
vec("floatToInt precision 10 keeps the raw bytes", O.FLOAT2INT, O.D_FLOAT, np.array([1e-12, 1.0], np.float32),
    [le32(10), np.array([1e-12, 1.0], np.float32).tobytes()])
# p = 1 (0.5), max * 10 = 3e9 >= INT_MAX: p = 10, raw bytes.
vec("floatToInt range overflow keeps the raw bytes", O.FLOAT2INT, O.D_FLOAT, np.array([3e8, 0.5], np.float32),
    [le32(10), np.array([3e8, 0.5], np.float32).tobytes()])
# p = 1, 3e7 * 10 fits: transformed.
vec("floatToInt range that fits", O.FLOAT2INT, O.D_FLOAT, np.array([3e7, 0.5], np.float32),
    [le32(1), le32(300000000, 5)])
vec("floatToInt float64 walk prices", O.FLOAT2INT, O.D_DOUBLE, np.array([100.25, 100.26, 99.97]),
    [le32(2), le32(10025, 10026, 9997)])
vec("floatToInt float64 out of range", O.FLOAT2INT, O.D_DOUBLE, np.array([2.5e8, 0.5]),
    [le32(10), np.array([2.5e8, 0.5]).tobytes()])
# integer columns: precision 0, scale factor 1, the bounds are strict.
vec("floatToInt int32 at INT_MIN falls back", O.FLOAT2INT, O.D_INT, i32(INT_MIN, 5),
    [le32(10), le32(INT_MIN, 5)])
vec("floatToInt int32 inside the bounds", O.FLOAT2INT, O.D_INT, i32(INT_MIN + 1, INT_MAX - 1),
    [le32(0), le32(INT_MIN + 1, INT_MAX - 1)])
vec("floatToInt int64 beyond the int range falls back", O.FLOAT2INT, O.D_TIME, np.array([2**40, 1], np.int64),
    [le32(10), np.array([2**40, 1], np.int64).tobytes()])
vec("floatToInt int64 inside the int range", O.FLOAT2INT, O.D_TIME, np.array([5, -3], np.int64),
    [le32(0), le32(5, -3)])

# -------------------------------------------------------------- AFL float
# AflEncoding::Encode<float> (compression/afl/afl_encoding.cu:115-167): split
# into mantissa / exponent / sign ints, AFL each of mantissa and exponent
# ([minBit, rest, 0, 0] + LSB-first words), data = both concatenated,
# metadata = size_t |mantissa words| + char sign (min >= 0: 1; max < 0: 2;
# else 0 followed by the packed sign stencil).
#
# 1.5 = 3fc00000 (m 0x400000, e 127), 2.0 = 40000000 (m 0, e 128),
# 3.25 = 40500000 (m 0x500000, e 128).  Mantissa: 23 bits, 69 bits in 3
# words, rest 27: words 00400000, 00000000, 00000014 (0x500000 >> 18).
# Exponent: 8 bits, 1 word, rest 8: 0080807f.
AFL_F_POS_DATA = h("17 1b 00 00  00 00 40 00 00 00 00 00 14 00 00 00  08 08 00 00  7f 80 80 00")
vec("afl float32 all positive", O.AFL, O.D_FLOAT, np.array([1.5, 2.0, 3.25], np.float32),
    [le64(12) + h("01"), AFL_F_POS_DATA])
# -1.5, -2.0, -3.25: the same mantissas and exponents, sign 2.
vec("afl float32 all negative", O.AFL, O.D_FLOAT, np.array([-1.5, -2.0, -3.25], np.float32),
    [le64(12) + h("02"), AFL_F_POS_DATA])
# -1.5, 2.0, 3.25: sign 0 + stencil of the sign bits 1 0 0 (padding 0).
vec("afl float32 mixed signs carry a sign stencil", O.AFL, O.D_FLOAT, np.array([-1.5, 2.0, 3.25], np.float32),
    [le64(12) + h("00 03 01"), AFL_F_POS_DATA])
# -0.0 >= 0 holds: a column of -0.0 and positives is "all positive" (sign 1),
# so the sign of zero is not kept (a lossy case of the reference).
# -0.0 = 80000000 (m 0, e 0), 1.0 = 3f800000 (m 0, e 127).  Mantissa: max 0,
# 1 bit, 2 bits in 1 word, rest 30.  Exponent: 7 bits, 14 bits, rest 18:
# word 00003f80.
vec("afl float32 negative zero counts as positive", O.AFL, O.D_FLOAT, np.array([-0.0, 1.0], np.float32),
    [le64(4) + h("01"), h("01 1e 00 00  00 00 00 00  07 12 00 00  80 3f 00 00")])

# -------------------------------------------------------------------- GFC
# GfcEncoding (compression/gfc/gfc_encoding.cpp:16-100), _gfcCompressKernelD
# (gfc_encoding_impl_double.cu:31-94), CompressDouble (:152-255): blocks =
# ceil(n / 32768), 16 chunks per block for double (32 for float), per =
# roundup32(max(32, ceil(N / chunks))).  Every 32-row subchunk is coded
# against 0 (first subchunk) or the chunk's row 31 (every later subchunk:
# `last = cData[chunkBegin + 31]`, not the previous subchunk's last row).
# Subchunk = 16 nibble bytes (lane 2j low, 2j+1 high; nibble = sign << 3 |
# code) then each lane's low residual bytes.  Double: 2 significant bytes are
# stored as 3; code = bytes - 1 from 3 bytes up.  Offsets: int32 per chunk,
# (start + 1) / 2 * 17 + chunk bytes (9 for float).
#
# 1536 doubles: 16 chunks of 96 rows (3 subchunks each).  Chunk 0:
#   rows 0..9: diffs 0, 7f, 1234, 123456, 12345678, 123456789a,
#              123456789abc, 123456789abcde, 4059000000000000, -5;
#   row 31 = 0x100; rows 32..63 = 0x100 except 32 = 0x101, 63 = 0x200;
#   rows 64..95 = 0x100 except 64 = 0x200 (against row 31: diff 0x100, code 2;
#   against row 63 it would be 0); every other row 0.
_GFC_D = np.zeros(1536, np.uint64)
_GFC_D[:10] = [0, 0x7F, 0x1234, 0x123456, 0x12345678, 0x123456789A, 0x123456789ABC, 0x123456789ABCDE,
               0x4059000000000000, 0xFFFFFFFFFFFFFFFB]
_GFC_D[31] = 0x100
_GFC_D[32:96] = 0x100
_GFC_D[32], _GFC_D[63], _GFC_D[64] = 0x101, 0x200, 0x200
GFC_D_CHUNK0 = (
    # subchunk 0: nibbles (codes 0 1 | 2 2 | 3 4 | 5 6 | 7 9 | ... | 0 2), residuals
    h("10 22 43 65 97 00 00 00 00 00 00 00 00 00 00 20")
    + h("7f  34 12 00  56 34 12  78 56 34 12  9a 78 56 34 12  bc 9a 78 56 34 12  de bc 9a 78 56 34 12"
        "00 00 00 00 00 00 59 40  05  00 01 00")
    # subchunk 1 (against row 31 = 0x100): row 32 diff 1, row 63 diff 0x100
    + h("01 00 00 00 00 00 00 00 00 00 00 00 00 00 00 20") + h("01  00 01 00")
    # subchunk 2 (against row 31): row 64 diff 0x100
    + h("02 00 00 00 00 00 00 00 00 00 00 00 00 00 00 00") + h("00 01 00"))
assert len(GFC_D_CHUNK0) == 96
vec("gfc float64 byte counts and the row-31 predictor", O.GFC, O.D_DOUBLE, _GFC_D.view(np.float64),
    [le32(1, 16, 1536), le32(96, *[816 * c + 48 for c in range(1, 16)]), GFC_D_CHUNK0 + bytes(15 * 48)])

# 3072 floats: 32 chunks of 96 rows.  Float codes = bytes (0..4).
_GFC_F = np.zeros(3072, np.uint32)
_GFC_F[:6] = [0, 0x7F, 0x1234, 0x123456, 0x12345678, 0xFFFFFFFB]
_GFC_F[31] = 0x100
_GFC_F[32:96] = 0x100
_GFC_F[32], _GFC_F[63], _GFC_F[64] = 0x101, 0x200, 0x200
GFC_F_CHUNK0 = (
    h("10 32 94 00 00 00 00 00 00 00 00 00 00 00 00 20")
    + h("7f  34 12  56 34 12  78 56 34 12  05  00 01")
    + h("01 00 00 00 00 00 00 00 00 00 00 00 00 00 00 20") + h("01  00 01")
    + h("02 00 00 00 00 00 00 00 00 00 00 00 00 00 00 00") + h("00 01"))
assert len(GFC_F_CHUNK0) == 66
vec("gfc float32 byte counts and the row-31 predictor", O.GFC, O.D_FLOAT, _GFC_F.view(np.float32),
    [le32(1, 32, 3072), le32(66, *[432 * c + 48 for c in range(1, 32)]), GFC_F_CHUNK0 + bytes(31 * 48)])
