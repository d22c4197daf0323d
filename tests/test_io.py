"""Time-series record files and header files (host code of libddj_hip.so,
no GPU): the reference's sample headers parse, binary and CSV files
round-trip every column type, padding and malformed input are handled, and
the `ddj` CLI writes the reference's sample data set."""
import os
import subprocess

import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g

REF_HEADERS = {  # sample_data/*.header of the reference, as text
    "info": "timestamp,time,0\nCORE VOLTAGE,float,6\nCPU TEMP,float,6\nGPU TEMP,float,6\n",
    "generated": ("TIME,time,0\nPATTERN_A_INT,int,0\nPATTERN_B_INT,int,0\nFLOAT_PREC_4,float,6\n"
                  "PATTERN_A_FLOAT,float,6\nPATTERN_B_FLOAT,float,6\n"),
    "nyse": "MsgSeqNum,int,0\nMsgType,short,0\nSymbol 1,char,0\nPriceNumerator,int,0\n",
}


def columns(n, seed=0):
    r = np.random.default_rng(seed)
    defn = g.FileDefinition(["t", "i", "f", "d", "s", "c"],
                            [g.D_TIME, g.D_INT, g.D_FLOAT, g.D_DOUBLE, g.D_SHORT, g.D_CHAR], [0, 0, 2, 4, 0, 0])
    cols = [1_400_000_000 + np.cumsum(r.integers(0, 3, n)).astype(np.int64),
            r.integers(-2**31, 2**31 - 1, n).astype(np.int32),
            (r.integers(-10**6, 10**6, n) / 100.0).astype(np.float32),
            r.integers(-10**9, 10**9, n) / 10**4,
            r.integers(-2**15, 2**15 - 1, n).astype(np.int16),
            r.integers(-128, 127, n).astype(np.int8)]
    return defn, cols


@pytest.mark.parametrize("name", sorted(REF_HEADERS))
def test_reference_headers_parse(tmp_path, name):
    p = tmp_path / f"{name}.header"
    p.write_text(REF_HEADERS[name])
    d = g.FileDefinition.read(str(p))
    lines = [ln.split(",") for ln in REF_HEADERS[name].strip().splitlines()]
    assert d.names == [ln[0] for ln in lines]
    assert [g.TYPE_NAMES[t] for t in d.dtypes] == [ln[1] for ln in lines]
    assert d.decimals == [int(ln[2]) for ln in lines]
    d.write(str(tmp_path / "again.header"))
    assert (tmp_path / "again.header").read_text() == REF_HEADERS[name]


@pytest.mark.parametrize("padding", [0, 3])
def test_binary_round_trip_and_layout(tmp_path, padding):
    defn, cols = columns(1000, seed=padding)
    path = str(tmp_path / "x.inf")
    g.write_series(defn, path, cols, g.FORMAT_BIN, padding)
    raw = open(path, "rb").read()
    rec = sum(c.itemsize for c in cols) + padding
    assert len(raw) == 1000 * rec
    # record 7 by hand: every column's bytes in order, then zero padding
    want = b"".join(c[7:8].tobytes() for c in cols) + bytes(padding)
    assert raw[7 * rec:8 * rec] == want
    back = g.read_series(defn, path, g.FORMAT_BIN, padding, batch=333)
    for a, b in zip(back, cols):
        assert a.tobytes() == b.tobytes()


def test_csv_round_trip_and_text(tmp_path):
    defn, cols = columns(500, seed=9)
    path = str(tmp_path / "x.csv")
    g.write_series(defn, path, cols, g.FORMAT_CSV)
    first = open(path).readline().strip().split(",")
    assert first[0] == str(cols[0][0]) and first[1] == str(cols[1][0])
    assert first[2] == "%.2f" % cols[2][0] and first[3] == "%.4f" % cols[3][0]
    back = g.read_series(defn, path, g.FORMAT_CSV, batch=128)
    for k, (a, b) in enumerate(zip(back, cols)):
        assert a.tobytes() == b.tobytes(), k  # every value is exact at its decimals


def test_malformed_input_is_an_error(tmp_path):
    defn = g.FileDefinition(["a", "b"], [g.D_INT, g.D_INT])
    bad = tmp_path / "bad.csv"
    bad.write_text("1,2\n3\n")
    with pytest.raises(g.DdjError):
        g.read_series(defn, str(bad), g.FORMAT_CSV)
    bad.write_text("1,2,3\n")
    with pytest.raises(g.DdjError):
        g.read_series(defn, str(bad), g.FORMAT_CSV)
    trunc = tmp_path / "t.inf"
    trunc.write_bytes(bytes(12))  # one and a half 8-byte records
    with pytest.raises(g.DdjError):
        g.read_series(defn, str(trunc), g.FORMAT_BIN)
    h = tmp_path / "h.header"
    h.write_text("x,int64,0\n")
    with pytest.raises(g.DdjError):
        g.FileDefinition.read(str(h))


def test_cli_generates_the_sample_data_set(tmp_path):
    r = subprocess.run([g.CLI_PATH, "-g", "2000", "-o", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = g.FileDefinition.read(str(tmp_path / "generated.header"))
    assert d.names == ["TIME", "PATTERN_A_INT", "PATTERN_B_INT", "FLOAT_PREC_4", "PATTERN_A_FLOAT",
                       "PATTERN_B_FLOAT"]
    b = g.read_series(d, str(tmp_path / "generated.inf"))
    c = g.read_series(d, str(tmp_path / "generated.csv"), g.FORMAT_CSV)
    assert all(x.size == 2000 for x in b)
    for k in range(3):  # integer columns are identical in both files
        assert np.array_equal(b[k], c[k])
    assert np.all(np.diff(b[0]) >= 0)  # time never goes back


def test_cli_usage_errors():
    r = subprocess.run([g.CLI_PATH, "-c"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
