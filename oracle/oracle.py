"""CPU oracle for the column-compression path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module.  It is the checker: the product library
(``libddj_hip.so``) never calls it, and nothing here is measured as the
product.

Two layers:

* codecs, statistics, stencil and FL bit-packing: ``libddj_oracle.so``
  (``ddj_oracle.cpp``), a sequential C++ restatement of the reference's
  kernels, reached through ctypes;
* the cascade, the serialized tree format and the planner: restated below in
  plain Python from the reference's host code (citations are to the
  reference's ``src/`` tree).
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libddj_oracle.so")

# EncodingType (compression/encoding_type.hpp:17-30)
DELTA, AFL, SCALE, RLE, DICT, UNIQUE, PATCH, NONE, FLOAT2INT, CONST, GFC = range(11)
OUTSIDE_PATCH = 100
ENC_NAMES = {DELTA: "delta", AFL: "afl", SCALE: "scale", RLE: "rle", DICT: "dict", UNIQUE: "unique",
             PATCH: "patch", NONE: "none", FLOAT2INT: "floatToInt", CONST: "constData", GFC: "gfc"}
ENC_BY_NAME = {v: k for k, v in ENC_NAMES.items()}

# DataType codes (the reference's values are unknown; see ddj_oracle.h)
D_INT, D_TIME, D_FLOAT, D_DOUBLE, D_SHORT, D_CHAR = range(6)
NP_DTYPE = {D_INT: np.int32, D_TIME: np.int64, D_FLOAT: np.float32, D_DOUBLE: np.float64,
            D_SHORT: np.int16, D_CHAR: np.int8}
DTYPE_SIZE = {D_INT: 4, D_TIME: 8, D_FLOAT: 4, D_DOUBLE: 8, D_SHORT: 2, D_CHAR: 1}
HEADER = struct.Struct("<hhi")  # EncodingMetadataHeader (compression/encoding_metadata_header.hpp:13-17)


class OracleError(RuntimeError):
    pass


class _Buf(ctypes.Structure):
    _fields_ = [("p", ctypes.POINTER(ctypes.c_uint8)), ("n", ctypes.c_size_t)]


class Stats(ctypes.Structure):
    """DataStatistics (util/statistics/cuda_array_statistics.hpp:24-34)."""
    _fields_ = [("min", ctypes.c_double), ("max", ctypes.c_double), ("minBitCnt", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("sorted", ctypes.c_int32), ("rlMetric", ctypes.c_float),
                ("mean", ctypes.c_double), ("size", ctypes.c_uint64)]


def build() -> str:
    """Compile libddj_oracle.so (oracle/Makefile) and return its path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_encode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                 ctypes.POINTER(_Buf), ctypes.POINTER(ctypes.c_int)]
        L.orc_decode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Buf), ctypes.c_int, ctypes.POINTER(_Buf)]
        L.orc_predict_sizes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        L.orc_stats_compute.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(Stats)]
        L.orc_most_frequent.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_stencil_pack.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(_Buf)]
        L.orc_stencil_unpack.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(_Buf)]
        L.orc_fl_pack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                  ctypes.c_size_t]
        L.orc_fl_unpack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_fl_pack_lanes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_void_p, ctypes.c_size_t]
        L.orc_fl_unpack_lanes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_size_t, ctypes.c_void_p]
        L.orc_fl_fetch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_encoding_info.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
        L.orc_bitlen.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_float_precision.argtypes = [ctypes.c_float]
        L.orc_free.argtypes = [ctypes.POINTER(_Buf)]
        _lib = L
    return _lib


def _take(b: _Buf) -> bytes:
    out = ctypes.string_at(b.p, b.n) if b.n else b""
    lib().orc_free(ctypes.byref(b))
    return out


def _check(rc: int, what: str):
    if rc != 0:
        raise OracleError(f"{what} failed with status {rc}")


def as_bytes(x) -> bytes:
    if isinstance(x, (bytes, bytearray)):
        return bytes(x)
    return np.ascontiguousarray(x).tobytes()


def encoding_info(enc: int, dtype: int) -> Tuple[int, List[int]]:
    n = ctypes.c_int()
    rt = (ctypes.c_int * 2)()
    _check(lib().orc_encoding_info(enc, dtype, ctypes.byref(n), rt), "encoding_info")
    return n.value, [rt[0], rt[1]][: max(n.value, 0)]


def encode(enc: int, dtype: int, data) -> List[bytes]:
    """Encoding::Encode -> [metadata, out1, ...]."""
    raw = as_bytes(data)
    outs = (_Buf * 4)()
    n = ctypes.c_int()
    buf = ctypes.create_string_buffer(raw, len(raw)) if raw else None
    _check(lib().orc_encode(enc, dtype, buf, len(raw), outs, ctypes.byref(n)), f"encode {enc}/{dtype}")
    return [_take(outs[i]) for i in range(n.value)]


def decode(enc: int, dtype: int, parts: Sequence[bytes]) -> bytes:
    """Encoding::Decode([metadata, child outputs...]) -> raw bytes."""
    keep = [ctypes.create_string_buffer(p, max(len(p), 1)) for p in parts]
    ins = (_Buf * len(parts))()
    for i, (p, k) in enumerate(zip(parts, keep)):
        ins[i].p = ctypes.cast(k, ctypes.POINTER(ctypes.c_uint8))
        ins[i].n = len(p)
    out = _Buf()
    _check(lib().orc_decode(enc, dtype, ins, len(parts), ctypes.byref(out)), f"decode {enc}/{dtype}")
    return _take(out)


def predict_sizes(enc: int, dtype: int, data) -> Tuple[int, int]:
    raw = as_bytes(data)
    m, d = ctypes.c_size_t(), ctypes.c_size_t()
    buf = ctypes.create_string_buffer(raw, max(len(raw), 1))
    _check(lib().orc_predict_sizes(enc, dtype, buf, len(raw), ctypes.byref(m), ctypes.byref(d)), "predict_sizes")
    return m.value, d.value


def stats(dtype: int, data) -> Stats:
    raw = as_bytes(data)
    st = Stats()
    buf = ctypes.create_string_buffer(raw, max(len(raw), 1))
    _check(lib().orc_stats_compute(dtype, buf, len(raw), ctypes.byref(st)), "stats")
    return st


def most_frequent(dtype: int, data, k: int) -> Tuple[np.ndarray, int]:
    raw = as_bytes(data)
    out = np.zeros(k, dtype=NP_DTYPE[dtype])
    nd = ctypes.c_int()
    buf = ctypes.create_string_buffer(raw, max(len(raw), 1))
    _check(lib().orc_most_frequent(dtype, buf, len(raw), k, out.ctypes.data, ctypes.byref(nd)), "most_frequent")
    return out, nd.value


def stencil_pack(s) -> bytes:
    s = np.ascontiguousarray(s, dtype=np.int32)
    out = _Buf()
    _check(lib().orc_stencil_pack(s.ctypes.data, s.size, ctypes.byref(out)), "stencil_pack")
    return _take(out)


def stencil_unpack(b: bytes) -> np.ndarray:
    buf = ctypes.create_string_buffer(b, max(len(b), 1))
    out = _Buf()
    _check(lib().orc_stencil_unpack(buf, len(b), ctypes.byref(out)), "stencil_unpack")
    return np.frombuffer(_take(out), dtype=np.int32)


def fl_pack(dtype: int, bits: int, x: np.ndarray) -> np.ndarray:
    """run_afl_compress_cpu<T,1> restated: ceil(n*b/W) words of T."""
    x = np.ascontiguousarray(x)
    W = 8 * x.itemsize
    words = (bits * x.size + W - 1) // W
    out = np.zeros(words, dtype=x.dtype)
    _check(lib().orc_fl_pack(dtype, bits, x.ctypes.data, x.size, out.ctypes.data, words), "fl_pack")
    return out


def fl_unpack(dtype: int, bits: int, words: np.ndarray, n: int) -> np.ndarray:
    words = np.ascontiguousarray(words)
    out = np.zeros(n, dtype=words.dtype)
    _check(lib().orc_fl_unpack(dtype, bits, words.ctypes.data, n, out.ctypes.data), "fl_unpack")
    return out


def fl_lanes_words(itemsize: int, bits: int, n: int) -> int:
    """Words of T in the 32-lane stream of n values: whole blocks of 32 lanes x W values."""
    W = 8 * itemsize
    return (n + 32 * W - 1) // (32 * W) * 32 * bits


def fl_pack_lanes(dtype: int, bits: int, x: np.ndarray) -> np.ndarray:
    """run_afl_compress_cpu<T,32> restated over whole 32-lane blocks."""
    x = np.ascontiguousarray(x)
    words = fl_lanes_words(x.itemsize, bits, x.size)
    out = np.zeros(words, dtype=x.dtype)
    _check(lib().orc_fl_pack_lanes(dtype, bits, x.ctypes.data, x.size, out.ctypes.data, words), "fl_pack_lanes")
    return out


def fl_unpack_lanes(dtype: int, bits: int, words: np.ndarray, n: int) -> np.ndarray:
    words = np.ascontiguousarray(words)
    out = np.zeros(n, dtype=words.dtype)
    _check(lib().orc_fl_unpack_lanes(dtype, bits, words.ctypes.data, words.size, n, out.ctypes.data),
           "fl_unpack_lanes")
    return out


def fl_fetch(dtype: int, bits: int, lanes: int, words: np.ndarray, n: int, rows: np.ndarray) -> np.ndarray:
    """afl_decompress_base_value_gpu at each row (compression/afl/afl_gpu.cu:197-223)."""
    words = np.ascontiguousarray(words)
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    out = np.zeros(rows.size, dtype=words.dtype)
    _check(lib().orc_fl_fetch(dtype, bits, lanes, words.ctypes.data, n, rows.ctypes.data, rows.size,
                              out.ctypes.data), "fl_fetch")
    return out


def float_precision(x: float) -> int:
    return lib().orc_float_precision(x)


def bitlen(dtype: int, mn, mx) -> int:
    a = np.array([mn], dtype=NP_DTYPE[dtype])
    b = np.array([mx], dtype=NP_DTYPE[dtype])
    return lib().orc_bitlen(dtype, a.ctypes.data, b.ctypes.data)


def alt_bitlen(v: int) -> int:
    return lib().orc_alt_bitlen(v)


def compression_ratio(inp: int, out: int) -> float:
    """Encoding::GetCompressionRatio (compression/encoding.hpp:78-84)."""
    if out <= 0 or out > inp:
        return 1.0
    return inp / out


# --------------------------------------------------------------------- tree
@dataclass
class Node:
    """CompressionNode (tree/compression_node.hpp:35-115), reduced to what the
    serialized format and the planner read."""
    enc: int
    dtype: int
    children: List["Node"] = field(default_factory=list)
    ratio: float = 1.0

    def copy(self) -> "Node":
        return Node(self.enc, self.dtype, [c.copy() for c in self.children], self.ratio)

    def fix(self):
        """CompressionNode::Fix (compression_node.cpp:209-223)."""
        if not self.children:
            n, _ = encoding_info(self.enc, self.dtype)
            for _ in range(n):
                self.children.append(Node(NONE, self.dtype, [], 1.0))
        else:
            for c in self.children:
                c.fix()

    def preorder(self) -> List[int]:
        out = [self.enc]
        for c in self.children:
            out += c.preorder()
        return out

    def to_string(self) -> str:
        """CompressionNode::Print (compression_node.cpp:180-186)."""
        s = f"{ENC_NAMES[self.enc]}[{self.ratio:g}],"
        for c in self.children:
            s += c.to_string()
        return s


def tree_from_path(path: Sequence[int], dtype: int) -> Node:
    """PathGenerator::GenerateTree / CreateTree (optimizer/path_generator.cpp:70-91):
    pre-order path, children typed by GetReturnTypes."""
    pos = [0]

    def make(dt: int) -> Node:
        et = path[pos[0]]
        node = Node(et, dt)
        if et == NONE:
            return node
        n, rtypes = encoding_info(et, dt)
        for i in range(n):
            pos[0] += 1
            node.children.append(make(rtypes[i]))
        return node

    return make(dtype)


def tree_compress(node: Node, data: bytes) -> List[bytes]:
    """CompressionNode::Compress (compression_node.cpp:103-138): header +
    metadata, then each child's records; a leaf appends its first output."""
    outs = encode(node.enc, node.dtype, data)
    parts = [HEADER.pack(node.enc, node.dtype, len(outs[0])) + outs[0]]
    if not node.children:
        parts.append(outs[1])
    else:
        for i, child in enumerate(node.children):
            parts += tree_compress(child, outs[i + 1])
    node.ratio = compression_ratio(len(data), sum(len(p) for p in parts))
    return parts


def compress(node: Node, data) -> bytes:
    """CompressionTree::Compress (tree/compression_tree.cpp:98-103)."""
    return b"".join(tree_compress(node, as_bytes(data)))


def parse(blob: bytes, off: int = 0) -> Tuple[Node, int, list]:
    """DecompressNodes (compression_tree.cpp:105-146): pre-order headers;
    a none node's data length is the size_t in its metadata."""
    if off + HEADER.size > len(blob):
        raise OracleError("truncated header")
    enc, dt, mlen = HEADER.unpack_from(blob, off)
    off += HEADER.size
    if mlen < 0 or off + mlen > len(blob):
        raise OracleError("bad metadata length")
    meta = blob[off:off + mlen]
    off += mlen
    node = Node(enc, dt)
    if enc == NONE:
        (dlen,) = struct.unpack_from("<Q", meta, 0)
        if off + dlen > len(blob):
            raise OracleError("bad leaf length")
        payload = [meta, blob[off:off + dlen]]
        off += dlen
        return node, off, payload
    n, _ = encoding_info(enc, dt)
    kids = []
    for _ in range(n):
        child, off, sub = parse(blob, off)
        node.children.append(child)
        kids.append(sub)
    return node, off, [meta, kids]


def _decode_parsed(node: Node, payload) -> bytes:
    """CompressionNode::Decompress (compression_node.cpp:140-161)."""
    if node.enc == NONE:
        return decode(NONE, node.dtype, payload)
    meta, kids = payload
    parts = [meta] + [_decode_parsed(c, k) for c, k in zip(node.children, kids)]
    return decode(node.enc, node.dtype, parts)


def decompress(blob: bytes) -> bytes:
    """CompressionTree::Decompress (compression_tree.cpp:148-155)."""
    node, _, payload = parse(blob)
    return _decode_parsed(node, payload)


def predict_size(node: Node, data: bytes) -> int:
    """CompressionNode::PredictCompressionSize (compression_node.cpp:163-177)."""
    m, d = predict_sizes(node.enc, node.dtype, data)
    size = m + HEADER.size
    if not node.children:
        return size + d
    outs = encode(node.enc, node.dtype, data)
    _, rtypes = encoding_info(node.enc, node.dtype)
    for i, child in enumerate(node.children):
        size += predict_size(child, outs[i + 1])
    return size


# ----------------------------------------------------------------- planner
MAX_LEVEL = 3  # optimizer/path_generator.cpp:15


def generate_paths() -> List[List[int]]:
    """PathGenerator::GeneratePaths (path_generator.cpp:20-68): 6 x 5 paths."""
    first = [[], [FLOAT2INT], [DELTA], [SCALE], [DELTA, FLOAT2INT], [SCALE, FLOAT2INT]]
    second = [[DICT, AFL, NONE, AFL, NONE], [CONST, AFL, NONE], [RLE, AFL, NONE, AFL, NONE],
              [AFL, NONE], [PATCH, AFL, NONE, AFL, NONE]]
    return [a + b for a in first for b in second]


def get_continuations(et: int, dt: int, st: Stats, level: int) -> List[int]:
    """PathGenerator::GetContinuations (path_generator.cpp:93-146)."""
    out: List[int] = []
    if level != 0 and (et == NONE or st.size <= 0):
        return out
    out.append(NONE)
    if et in (AFL, GFC):
        return out
    if level < MAX_LEVEL:
        if et != DELTA:
            out.append(DELTA)
        if et != SCALE and st.min != 0:
            out.append(SCALE)
        if et != FLOAT2INT and dt == D_FLOAT:
            out.append(FLOAT2INT)
        if et != PATCH and et != DICT:
            out.append(PATCH)
        if et != DICT and et != PATCH:
            out.append(DICT)
        if st.sorted or st.rlMetric >= 1.5:
            if et != CONST:
                out.append(CONST)
            if et != RLE:
                out.append(RLE)
    out.append(GFC if dt in (D_FLOAT, D_DOUBLE) else AFL)
    return out


def full_statistics_update(data: bytes, et: int, dt: int, level: int) -> List[Tuple[Node, int]]:
    """CompressionOptimizer::FullStatisticsUpdate (optimizer/compression_optimizer.cpp:88-140)
    with CrossTrees (:38-78).  Returns (tree, exact size) pairs in the
    reference's enumeration order.  Note the reference keeps ``part1`` across
    loop iterations; it is only read after being assigned or when empty."""
    n = len(data) // DTYPE_SIZE[dt]
    st = stats(dt, data[: n * DTYPE_SIZE[dt]])
    result: List[Tuple[Node, int]] = []
    part1: List[Tuple[Node, int]] = []
    for c in get_continuations(et, dt, st, level):
        compr = encode(c, dt, data)
        psize = sum(len(x) for x in compr) + 8
        parent = Node(c, dt, [], compression_ratio(len(data), psize))
        nres, rtypes = encoding_info(c, dt)
        if nres == 1:
            kids = full_statistics_update(compr[1], c, rtypes[0], level + 1)
            part1 = []
            for k, ks in kids:
                out_size = ks + len(compr[0]) + 8
                t = Node(c, dt, [k.copy()], compression_ratio(len(data), out_size))
                part1.append((t, out_size))
        elif nres == 2:
            left = full_statistics_update(compr[1], c, rtypes[0], level + 1)
            right = full_statistics_update(compr[2], c, rtypes[1], level + 1)
            part1 = []
            for l, ls in left:
                for r, rs in right:
                    out_size = ls + rs + len(compr[0]) + 8
                    t = Node(c, dt, [l.copy(), r.copy()], compression_ratio(len(data), out_size))
                    part1.append((t, out_size))
        if not part1:
            part1 = [(parent, psize)]
        result += part1
    return result


def sample_size(nbytes: int, dt: int) -> int:
    """GetSampleDataForFullUpdateSize (compression_optimizer.cpp:148-154)."""
    n = nbytes // DTYPE_SIZE[dt]
    return nbytes if n <= 10000 else 10000 * DTYPE_SIZE[dt]


def plan_exhaustive(data, dt: int) -> Tuple[Node, int, int]:
    """Phase 1 of CompressionOptimizer::CompressData (:156-182): candidates on
    the first 10,000 rows, Fix() each, keep the first with the largest root
    ratio.  Returns (tree, its exact sample size, number of candidates)."""
    raw = as_bytes(data)
    sample = raw[: sample_size(len(raw), dt)]
    cands = full_statistics_update(sample, NONE, dt, 0)
    best, best_size, best_ratio = None, 0, -1.0
    for t, s in cands:
        t.fix()
        if t.ratio > best_ratio:
            best, best_size, best_ratio = t, s, t.ratio
    return best, best_size, len(cands)


def optimize_tree(data, dt: int) -> Node:
    """CompressionOptimizer::OptimizeTree (compression_optimizer.cpp:16-36):
    the smallest predicted size over the 30 generated paths, else none."""
    raw = as_bytes(data)
    best = Node(NONE, dt)
    min_size = len(raw)
    for path in generate_paths():
        t = tree_from_path(path, dt)
        s = predict_size(t, raw)
        if s < min_size:
            min_size, best = s, t
    return best


# --------------------------------------------------- edge statistics (phase 2)
def edges(root: Node, max_height: int = 5) -> List[Tuple[Node, Node, int]]:
    """CompressionTree::GetEdges (compression_tree.cpp:201-230): BFS numbering
    over a full binary tree, missing children replaced by a placeholder."""
    out = []
    maxno = (1 << (max_height + 1)) - 2
    fake = Node(NONE, D_INT)
    fifo = [root]
    no = 0
    while fifo and no < maxno:
        el = fifo.pop(0)
        ch = el.children
        if len(ch) >= 1:
            out.append((el, ch[0], no))
            fifo.append(ch[0])
        else:
            fifo.append(fake)
        no += 1
        if len(ch) >= 2:
            out.append((el, ch[1], no))
            fifo.append(ch[1])
        else:
            fifo.append(fake)
        no += 1
    return out


class EdgeStats:
    """CompressionStatistics (tree/compression_statistics.cpp:12-84): per edge
    slot, a map (from, to) -> running ratio.  Iteration over a slot follows
    insertion order here (the reference iterates an unordered_map)."""

    def __init__(self, height: int = 5):
        self.height = height
        self.stat = [dict() for _ in range((1 << (height + 1)) - 2)]

    def update(self, no: int, edge: Tuple[int, int], ratio: float):
        cur = self.stat[no].get(edge, 0.0)
        if cur < 1.0:
            cur = 1.0
        if ratio < 1.0:
            ratio = 1.0
        self.stat[no][edge] = (cur + ratio) / 2.0

    def set(self, no: int, edge: Tuple[int, int], value: float):
        self.stat[no][edge] = value

    def get(self, no: int, edge: Tuple[int, int]) -> float:
        return self.stat[no].setdefault(edge, 0.0)

    def get_any(self, no: int):
        best = (None, 1.0)
        for k, v in self.stat[no].items():
            best = (k, v)
        return best

    def get_best(self, no: int, old: Tuple[int, int], same_beginning: bool = False):
        best_edge, best_value = (old[0], NONE), 1.0
        for k, v in self.stat[no].items():
            if same_beginning and k[0] != old[0]:
                continue
            if k[1] != old[1] and v > best_value:
                best_edge, best_value = k, v
        return best_edge, best_value

    def copy(self) -> "EdgeStats":
        c = EdgeStats(self.height)
        c.stat = [dict(d) for d in self.stat]
        return c


def update_statistics(root: Node, st: EdgeStats):
    """CompressionTree::UpdateStatistics (compression_tree.cpp:232-240)."""
    for a, b, no in edges(root, st.height):
        st.update(no, (a.enc, b.enc), (a.ratio + b.ratio) / 2.0)


def decodable(root: Node) -> bool:
    """Every leaf is `none` (a tree the decoder can walk)."""
    if not root.children:
        return root.enc == NONE
    return all(decodable(c) for c in root.children)


class Optimizer:
    """CompressionOptimizer::CompressData (optimizer/compression_optimizer.cpp:156-193)
    with OptimalTree (optimizer/optimal_tree.cpp:14-93) and the optimizer's
    edge statistics (tree/compression_statistics.cpp)."""

    MAX_EDGE_NO = (1 << 6) - 2

    def __init__(self):
        self.tree: Optional[Node] = None
        self.stats = EdgeStats(5)
        self.old_stats = EdgeStats(5)
        self.aliased = False
        self.old_ratio = 0.0
        self.parts = 0

    def compress_data(self, data, dt: int) -> bytes:
        raw = as_bytes(data)
        if self.tree is None or self.parts % 10 == 9:
            sample = raw[: sample_size(len(raw), dt)]
            cands = full_statistics_update(sample, NONE, dt, 0)
            best, bratio = None, -1.0
            for t, _ in cands:
                t.fix()
                update_statistics(t, self.stats)
                if t.ratio > bratio:
                    best, bratio = t, t.ratio
            self.tree = best
            self.old_ratio = best.ratio
            self.old_stats = self.stats.copy()
            self.aliased = False
        blob = compress(self.tree, raw)
        update_statistics(self.tree, self.stats)
        self.try_correct_tree()
        self.parts += 1
        return blob

    def replace(self, node: Node, edge_no: int):
        """OptimalTree::Replace (optimal_tree.cpp:14-37)."""
        if edge_no > self.MAX_EDGE_NO:
            return
        n, rtypes = encoding_info(node.enc, node.dtype)
        for i in range(n):
            t = node.enc
            (frm, to), value = self.stats.get_best(edge_no + i, (t, t))
            if value > 1:
                child = Node(to, rtypes[i])
                node.children.append(child)
                self.replace(child, 2 * (edge_no + i) + 2)
            else:
                node.children.append(Node(NONE, rtypes[i]))

    def try_correct_tree(self) -> bool:
        """OptimalTree::TryCorrectTree (optimal_tree.cpp:39-93)."""
        old_ratio, new_ratio = self.old_ratio, self.tree.ratio
        if old_ratio <= new_ratio and old_ratio > 1:
            return False
        olds = self.stats if self.aliased else self.old_stats
        for frm, to, no in edges(self.tree, 5):
            typ = (frm.enc, to.enc)
            old_value = olds.get(no, typ)
            new_value = self.stats.get(no, typ)
            best_edge, best_value = self.stats.get_best(no, typ)
            if new_value < old_value or best_value > new_value:
                same_edge, same_value = self.stats.get_best(no, typ, True)
                if same_value > new_value:
                    to.children, to.ratio, to.enc = [], 1.0, same_edge[1]
                    self.replace(to, 2 * (no + 1))
                else:
                    frm.children, frm.ratio, frm.enc = [], 1.0, best_edge[0]
                    self.replace(frm, no - 1 if no % 2 else no)
                self.old_ratio = self.tree.ratio
                self.aliased = True
                return True
        return False
