"""MI355X column-compression engine: Python mirror of the reference's host API.

The compute path is ``libddj_hip.so`` (hand-written HIP kernels for gfx950
behind the C ABI declared in ``include/ddj.h``); this module only binds it
with ctypes.  The class and method names follow the reference
(dzitkowskik/GpuTimeSeriesCompressionOptimizer, src/):

* :class:`Encoding`        -- compression/encoding.hpp (Encode / Decode /
  GetNumberOfResults / GetReturnTypes / GetMetadataSize / GetCompressedSize)
* :class:`CompressionTree` -- tree/compression_tree.hpp (Compress /
  Decompress / GetPredictedSizeAfterCompression)
* :class:`CompressionOptimizer` -- optimizer/compression_optimizer.hpp
  (OptimizeTree / CompressData) plus the exhaustive phase-1 planner
* :class:`ParallelTSCompressor` -- parallel_ts_compressor.hpp

There is no fallback: if the library is missing every call raises.
"""
from __future__ import annotations

import ctypes
import os
import weakref
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libddj_hip.so")

# EncodingType (compression/encoding_type.hpp:17-30)
DELTA, AFL, SCALE, RLE, DICT, UNIQUE, PATCH, NONE, FLOAT_TO_INT, CONST, GFC = range(11)
OUTSIDE_PATCH = 100
ENCODING_NAMES = {DELTA: "delta", AFL: "afl", SCALE: "scale", RLE: "rle", DICT: "dict", UNIQUE: "unique",
                  PATCH: "patch", NONE: "none", FLOAT_TO_INT: "floatToInt", CONST: "constData", GFC: "gfc",
                  OUTSIDE_PATCH: "patch(outside)"}

# DataType codes carried in tree headers (see include/ddj.h)
D_INT, D_TIME, D_FLOAT, D_DOUBLE, D_SHORT, D_CHAR = range(6)
NP_DTYPE = {D_INT: np.int32, D_TIME: np.int64, D_FLOAT: np.float32, D_DOUBLE: np.float64, D_SHORT: np.int16,
            D_CHAR: np.int8}
DTYPE_OF_NP = {np.dtype(v): k for k, v in NP_DTYPE.items()}

PLAN_EXHAUSTIVE, PLAN_OPTIMIZE30, PLAN_LEARNED = 0, 1, 2

STATUS = {0: "OK", 1: "EINVAL", 2: "EUNSUPPORTED", 3: "ECORRUPT", 4: "ENOMEM", 5: "EHIP", 6: "ETIMEOUT"}


class DdjError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class _Buf(ctypes.Structure):
    _fields_ = [("dev", ctypes.c_void_p), ("bytes", ctypes.c_size_t)]


class Stats(ctypes.Structure):
    """ddj_stats_t: DataStatistics (util/statistics/cuda_array_statistics.hpp:24-34) + planner features."""
    _fields_ = [("min", ctypes.c_double), ("max", ctypes.c_double), ("minBitCnt", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("sorted", ctypes.c_int32), ("rlMetric", ctypes.c_float),
                ("mean", ctypes.c_double), ("size", ctypes.c_uint64), ("runs", ctypes.c_uint64),
                ("distinct_est", ctypes.c_uint64), ("bit_hist", ctypes.c_uint32 * 65),
                ("delta_bits_max", ctypes.c_uint32)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libddj_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m gputimeseriescompressionoptimizer_amd.build`")
    L = ctypes.CDLL(path)
    vp, sz, i32, u32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
    P = ctypes.POINTER
    sig = {
        "ddj_ctx_create": (i32, [P(vp), i32]),
        "ddj_ctx_destroy": (None, [vp]),
        "ddj_ctx_sync": (i32, [vp]),
        "ddj_last_error": (ctypes.c_char_p, [vp]),
        "ddj_ctx_trim": (None, [vp]),
        "ddj_ctx_stream": (vp, [vp]),
        "ddj_buf_alloc": (i32, [vp, sz, P(_Buf)]),
        "ddj_buf_free": (None, [vp, P(_Buf)]),
        "ddj_buf_detach": (i32, [vp, P(_Buf)]),
        "ddj_memcpy_h2d": (i32, [vp, vp, vp, sz]),
        "ddj_memcpy_d2h": (i32, [vp, vp, vp, sz]),
        "ddj_memcpy_d2d": (i32, [vp, vp, vp, sz]),
        "ddj_encoding_info": (i32, [i32, i32, P(u32), P(i32)]),
        "ddj_encode": (i32, [vp, i32, i32, vp, sz, P(_Buf)]),
        "ddj_decode": (i32, [vp, i32, i32, P(_Buf), u32, P(_Buf)]),
        "ddj_predict_sizes": (i32, [vp, i32, i32, vp, sz, P(sz), P(sz)]),
        "ddj_stats": (i32, [vp, i32, vp, sz, P(Stats)]),
        "ddj_fl_pack": (i32, [vp, i32, i32, vp, sz, vp]),
        "ddj_fl_unpack": (i32, [vp, i32, i32, vp, sz, sz, vp]),
        "ddj_fl_fetch": (i32, [vp, i32, i32, i32, vp, sz, sz, vp, sz, vp]),
        "ddj_fl_lanes_bytes": (sz, [i32, i32, sz]),
        "ddj_fl_pack_lanes": (i32, [vp, i32, i32, vp, sz, vp]),
        "ddj_fl_unpack_lanes": (i32, [vp, i32, i32, vp, sz, sz, vp]),
        "ddj_tree_compress": (i32, [vp, P(i32), sz, i32, vp, sz, P(_Buf), P(ctypes.c_double)]),
        "ddj_tree_decompress": (i32, [vp, vp, sz, P(_Buf), P(i32)]),
        "ddj_tree_decompress_noalloc": (i32, [vp, vp, sz, vp, sz, P(sz), P(i32)]),
        "ddj_tree_decompress_rows": (i32, [vp, vp, sz, sz, sz, P(_Buf), P(i32)]),
        "ddj_tree_predict_size": (i32, [vp, P(i32), sz, i32, vp, sz, P(sz)]),
        "ddj_plan": (i32, [vp, i32, i32, vp, sz, P(i32), sz, P(sz), P(ctypes.c_double), P(u64)]),
        "ddj_optimizer_create": (i32, [vp, P(vp)]),
        "ddj_optimizer_destroy": (None, [vp]),
        "ddj_optimizer_compress": (i32, [vp, i32, vp, sz, P(_Buf)]),
        "ddj_optimizer_tree": (i32, [vp, P(i32), sz, P(sz)]),
        "ddj_profile_enable": (i32, [vp, i32]),
        "ddj_profile_kernel": (i32, [vp, ctypes.c_char_p, P(ctypes.c_double), P(u64), P(ctypes.c_double)]),
        "ddj_profile_reset": (i32, [vp]),
        "ddj_ts_create": (i32, [vp, u32, P(i32), sz, P(vp)]),
        "ddj_ts_destroy": (None, [vp]),
        "ddj_ts_compress": (i32, [vp, P(vp), sz, ctypes.c_char_p, i32]),
        "ddj_ts_reader_open": (i32, [vp, ctypes.c_char_p, u32, P(i32), P(vp)]),
        "ddj_ts_reader_next": (i32, [vp, P(vp), sz, P(sz)]),
        "ddj_ts_reader_close": (None, [vp]),
        "ddj_io_last_error": (ctypes.c_char_p, []),
        "ddj_file_def_read": (i32, [ctypes.c_char_p, P(vp)]),
        "ddj_file_def_create": (i32, [u32, P(ctypes.c_char_p), P(i32), P(i32), P(vp)]),
        "ddj_file_def_write": (i32, [vp, ctypes.c_char_p]),
        "ddj_file_def_columns": (u32, [vp]),
        "ddj_file_def_column": (i32, [vp, u32, P(ctypes.c_char_p), P(i32), P(i32)]),
        "ddj_file_def_free": (None, [vp]),
        "ddj_series_open": (i32, [vp, ctypes.c_char_p, i32, i32, i32, P(vp)]),
        "ddj_series_read": (i32, [vp, sz, P(vp), P(sz)]),
        "ddj_series_write": (i32, [vp, P(vp), sz]),
        "ddj_series_close": (None, [vp]),
        "ddj_batch_compress": (i32, [vp, vp, u32, u32, P(_Buf)]),
        "ddj_batch_compress_config": (i32, [vp, vp, u32, vp, P(_Buf)]),
        "ddj_batch_autotune": (i32, [vp, vp, u32, vp, P(ctypes.c_double)]),
        "ddj_planner_candidates": (i32, [P(i32), sz, P(sz)]),
        "ddj_planner_candidates_class": (i32, [i32, P(i32), sz, P(sz)]),
        "ddj_planner_create_class": (i32, [vp, i32, u64, P(vp)]),
        "ddj_planner_features": (i32, [vp, i32, vp, sz, P(ctypes.c_float)]),
        "ddj_planner_labels": (i32, [vp, i32, vp, sz, P(ctypes.c_float)]),
        "ddj_planner_create": (i32, [vp, u64, P(vp)]),
        "ddj_planner_destroy": (None, [vp]),
        "ddj_planner_params": (i32, [vp, vp, vp]),
        "ddj_planner_gradient": (i32, [vp, vp, vp, sz, vp, P(ctypes.c_float)]),
        "ddj_planner_train": (i32, [vp, vp, vp, sz, i32, i32, ctypes.c_float, u64, P(ctypes.c_float)]),
        "ddj_planner_predict": (i32, [vp, vp, sz, vp]),
        "ddj_planner_plan": (i32, [vp, i32, vp, sz, ctypes.c_float, P(i32), sz, P(sz), P(i32),
                                   P(ctypes.c_float)]),
        "ddj_planner_plan_checked": (i32, [vp, i32, vp, sz, ctypes.c_float, i32, P(i32), sz, P(sz), P(i32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def encoding_info(enc: int, dtype: int) -> Tuple[int, List[int]]:
    """Encoding::GetNumberOfResults and GetReturnTypes."""
    L = load_library()
    n = ctypes.c_uint32()
    rt = (ctypes.c_int * 2)()
    st = L.ddj_encoding_info(enc, dtype, ctypes.byref(n), rt)
    if st:
        raise DdjError(st, f"encoding_info({enc}, {dtype})")
    return n.value, [rt[0], rt[1]][: n.value]


class DeviceBuffer:
    """Device memory owned by a :class:`Context` (freed with the object)."""

    def __init__(self, ctx: "Context", dev: int, nbytes: int, owned: bool = True):
        self.ctx = ctx
        self.dev = dev or 0
        self.nbytes = nbytes
        self._owned = owned

    def __len__(self):
        return self.nbytes

    def to_bytes(self) -> bytes:
        return self.ctx.download(self)

    def to_numpy(self, dtype) -> np.ndarray:
        return np.frombuffer(self.to_bytes(), dtype=dtype)

    def view(self, offset: int, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self.ctx, self.dev + offset if self.dev else 0, nbytes, owned=False)

    def free(self):
        if self._owned and self.dev and self.ctx._h:
            b = _Buf(self.dev, self.nbytes)
            self.ctx._L.ddj_buf_free(self.ctx._h, ctypes.byref(b))
        self.dev = 0
        self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _as_c(self) -> _Buf:
        return _Buf(self.dev or None, self.nbytes)


class Context:
    """A GPU, a HIP stream and a caching device allocator (ddj_ctx)."""

    def __init__(self, device: int = 0):
        self._L = load_library()
        h = ctypes.c_void_p()
        st = self._L.ddj_ctx_create(ctypes.byref(h), device)
        if st:
            raise DdjError(st, f"ddj_ctx_create(device={device})")
        self._h = h
        self._deps = weakref.WeakSet()  # objects holding native state allocated from this context

    def check(self, st: int, what: str):
        if st:
            msg = self._L.ddj_last_error(self._h)
            raise DdjError(st, f"{what}: {msg.decode() if msg else ''}")

    def _track(self, obj) -> None:
        self._deps.add(obj)

    def close(self):
        """Releases every optimizer, planner and column compressor made from
        this context, then the context itself."""
        if getattr(self, "_h", None):
            for d in list(self._deps):
                d._release()
            self._L.ddj_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self.check(self._L.ddj_ctx_sync(self._h), "sync")

    def profile(self, on: bool = True):
        """Record HIP events around every kernel launch of this context."""
        self.check(self._L.ddj_profile_enable(self._h, 1 if on else 0), "profile")

    def profile_reset(self):
        self.check(self._L.ddj_profile_reset(self._h), "profile_reset")

    def kernel_time(self, name: str) -> Tuple[float, int, float]:
        """(total milliseconds, launches, algorithmic bytes) of one kernel since the last reset."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        b = ctypes.c_double()
        self.check(self._L.ddj_profile_kernel(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n),
                                              ctypes.byref(b)), "kernel_time")
        return ms.value, n.value, b.value

    KERNELS = ["k_adam", "k_add_scalar", "k_code_edges", "k_copy_pairs", "k_delta_decode", "k_delta_encode",
               "k_delta_sums", "k_fill_zero", "k_fl_delta_sums", "k_fl_edges", "k_fl_fetch", "k_fl_pack",
               "k_fl_pack_lanes", "k_fl_unpack", "k_fl_unpack_delta_scan", "k_fl_unpack_lanes", "k_float_compose",
               "k_float_split", "k_float_to_int", "k_gather_copy", "k_gfc_chunk_bytes", "k_gfc_decode", "k_gfc_encode",
               "k_gfc_sizes", "k_gfc_walk", "k_hist_count", "k_hist_init", "k_hist_topk", "k_int_to_float",
               "k_length_sums", "k_merge", "k_merge_fl", "k_minmax", "k_minmax_delta", "k_minmax_fold",
               "k_minmax_prec", "k_mlp_backward", "k_mlp_forward", "k_order_keys", "k_parse_tree", "k_radix_count",
               "k_radix_scatter", "k_rle_count", "k_rle_encode", "k_rle_expand", "k_rle_lengths", "k_rle_tile_runs",
               "k_scan_block_apply", "k_scan_block_sums", "k_scan_lengths", "k_scan_tiles", "k_split", "k_split_count",
               "k_split_fl", "k_stats", "k_stencil_count", "k_to_host", "k_unique_count", "k_unique_decode",
               "k_unique_encode", "k_unique_write"]

    def kernel_table(self):
        """{kernel: (ms, launches, algorithmic bytes)} for every kernel that ran."""
        out = {}
        for k in self.KERNELS:
            t = self.kernel_time(k)
            if t[1]:
                out[k] = t
        return out

    def stream(self) -> int:
        return self._L.ddj_ctx_stream(self._h) or 0

    def alloc(self, nbytes: int) -> DeviceBuffer:
        b = _Buf()
        self.check(self._L.ddj_buf_alloc(self._h, nbytes, ctypes.byref(b)), "alloc")
        return DeviceBuffer(self, b.dev, nbytes)

    def upload(self, data) -> DeviceBuffer:
        if isinstance(data, np.ndarray):
            data = np.ascontiguousarray(data)
            raw, n = data.ctypes.data, data.nbytes
            keep = data
        else:
            keep = ctypes.create_string_buffer(bytes(data), max(len(data), 1))
            raw, n = ctypes.addressof(keep), len(data)
        buf = self.alloc(n)
        if n:
            self.check(self._L.ddj_memcpy_h2d(self._h, buf.dev, raw, n), "upload")
        del keep
        return buf

    def download(self, buf: DeviceBuffer) -> bytes:
        if buf.nbytes == 0:
            return b""
        out = ctypes.create_string_buffer(buf.nbytes)
        self.check(self._L.ddj_memcpy_d2h(self._h, out, buf.dev, buf.nbytes), "download")
        return out.raw

    def _wrap(self, b: _Buf) -> DeviceBuffer:
        return DeviceBuffer(self, b.dev or 0, b.bytes)

    def stats(self, buf: DeviceBuffer, dtype: int) -> Stats:
        """CudaArrayStatistics::GenerateStatistics."""
        s = Stats()
        self.check(self._L.ddj_stats(self._h, dtype, buf.dev or None, buf.nbytes, ctypes.byref(s)), "stats")
        return s

    def fl_pack(self, buf: DeviceBuffer, dtype: int, bits: int, n: int) -> DeviceBuffer:
        W = 8 * np.dtype(NP_DTYPE[dtype]).itemsize
        nbytes = (bits * n + W - 1) // W * (W // 8)
        out = self.alloc(nbytes)
        self.check(self._L.ddj_fl_pack(self._h, dtype, bits, buf.dev or None, n, out.dev or None), "fl_pack")
        return out

    def fl_unpack(self, buf: DeviceBuffer, dtype: int, bits: int, n: int) -> DeviceBuffer:
        out = self.alloc(n * np.dtype(NP_DTYPE[dtype]).itemsize)
        self.check(self._L.ddj_fl_unpack(self._h, dtype, bits, buf.dev or None, buf.nbytes, n, out.dev or None),
                   "fl_unpack")
        return out


    def fl_fetch(self, stream: DeviceBuffer, dtype: int, bits: int, n: int, rows: DeviceBuffer,
                 lanes: int = 1) -> DeviceBuffer:
        """Values at the uint64 row indices `rows` of an n-value FL stream
        (afl_decompress_value_gpu, compression/afl/afl_gpu.cu:113-121)."""
        count = rows.nbytes // 8
        out = self.alloc(count * np.dtype(NP_DTYPE[dtype]).itemsize)
        self.check(self._L.ddj_fl_fetch(self._h, dtype, bits, lanes, stream.dev or None, stream.nbytes, n,
                                        rows.dev or None, count, out.dev or None), "fl_fetch")
        return out

    def fl_pack_lanes(self, buf: DeviceBuffer, dtype: int, bits: int, n: int) -> DeviceBuffer:
        """The 32-lane interleaved FL stream (run_afl_compress_gpu<T,32>)."""
        out = self.alloc(self._L.ddj_fl_lanes_bytes(dtype, bits, n))
        self.check(self._L.ddj_fl_pack_lanes(self._h, dtype, bits, buf.dev or None, n, out.dev or None),
                   "fl_pack_lanes")
        return out

    def fl_unpack_lanes(self, buf: DeviceBuffer, dtype: int, bits: int, n: int) -> DeviceBuffer:
        out = self.alloc(n * np.dtype(NP_DTYPE[dtype]).itemsize)
        self.check(self._L.ddj_fl_unpack_lanes(self._h, dtype, bits, buf.dev or None, buf.nbytes, n,
                                               out.dev or None), "fl_unpack_lanes")
        return out


class Encoding:
    """One reference encoding (compression/encoding.hpp:18-113)."""

    def __init__(self, ctx: Context, encoding_type: int):
        self.ctx = ctx
        self.encoding_type = encoding_type

    def GetNumberOfResults(self, dtype: int = D_INT) -> int:
        return encoding_info(self.encoding_type, dtype)[0]

    def GetReturnTypes(self, dtype: int) -> List[int]:
        return encoding_info(self.encoding_type, dtype)[1]

    def Encode(self, data: DeviceBuffer, dtype: int) -> List[DeviceBuffer]:
        n, _ = encoding_info(self.encoding_type, dtype)
        k = 1 + max(n, 1)  # `none` returns its data copy as outs[1] (none_encoding.cu:13-30)
        outs = (_Buf * k)()
        self.ctx.check(self.ctx._L.ddj_encode(self.ctx._h, self.encoding_type, dtype, data.dev or None, data.nbytes,
                                              outs), f"encode {ENCODING_NAMES.get(self.encoding_type)}")
        return [self.ctx._wrap(outs[i]) for i in range(k)]

    def Decode(self, parts: Sequence[DeviceBuffer], dtype: int) -> DeviceBuffer:
        ins = (_Buf * len(parts))(*[p._as_c() for p in parts])
        out = _Buf()
        self.ctx.check(self.ctx._L.ddj_decode(self.ctx._h, self.encoding_type, dtype, ins, len(parts),
                                              ctypes.byref(out)), f"decode {ENCODING_NAMES.get(self.encoding_type)}")
        return self.ctx._wrap(out)

    def _sizes(self, data: DeviceBuffer, dtype: int) -> Tuple[int, int]:
        m, d = ctypes.c_size_t(), ctypes.c_size_t()
        self.ctx.check(self.ctx._L.ddj_predict_sizes(self.ctx._h, self.encoding_type, dtype, data.dev or None,
                                                     data.nbytes, ctypes.byref(m), ctypes.byref(d)), "predict")
        return m.value, d.value

    def GetMetadataSize(self, data: DeviceBuffer, dtype: int) -> int:
        return self._sizes(data, dtype)[0]

    def GetCompressedSize(self, data: DeviceBuffer, dtype: int) -> int:
        return self._sizes(data, dtype)[1]

    @staticmethod
    def GetCompressionRatio(input_size: int, output_size: int) -> float:
        """Encoding::GetCompressionRatio (compression/encoding.hpp:78-84)."""
        if output_size <= 0 or output_size > input_size:
            return 1.0
        return input_size / output_size


def _path_arg(path: Sequence[int]):
    arr = (ctypes.c_int * len(path))(*path)
    return arr, len(path)


class CompressionTree:
    """A cascade given as a pre-order path of encodings
    (PathGenerator::GenerateTree, optimizer/path_generator.cpp:70-91)."""

    def __init__(self, ctx: Context, path: Optional[Sequence[int]] = None, dtype: int = D_INT):
        self.ctx = ctx
        self.path = list(path) if path is not None else None
        self.dtype = dtype
        self.node_ratios: List[float] = []

    def Compress(self, data: DeviceBuffer) -> DeviceBuffer:
        arr, n = _path_arg(self.path)
        ratios = (ctypes.c_double * n)()
        out = _Buf()
        self.ctx.check(self.ctx._L.ddj_tree_compress(self.ctx._h, arr, n, self.dtype, data.dev or None, data.nbytes,
                                                     ctypes.byref(out), ratios), "tree compress")
        self.node_ratios = list(ratios)
        return self.ctx._wrap(out)

    def GetCompressionRatio(self) -> float:
        return self.node_ratios[0] if self.node_ratios else 1.0

    def Decompress(self, blob: DeviceBuffer) -> DeviceBuffer:
        out = _Buf()
        dt = ctypes.c_int()
        self.ctx.check(self.ctx._L.ddj_tree_decompress(self.ctx._h, blob.dev or None, blob.nbytes, ctypes.byref(out),
                                                       ctypes.byref(dt)), "tree decompress")
        self.dtype = dt.value
        return self.ctx._wrap(out)

    def DecompressRows(self, blob: DeviceBuffer, begin: int, count: int) -> DeviceBuffer:
        """Rows [begin, begin + count) of the blob's column (partial decode)."""
        out = _Buf()
        dt = ctypes.c_int()
        self.ctx.check(self.ctx._L.ddj_tree_decompress_rows(self.ctx._h, blob.dev or None, blob.nbytes, begin, count,
                                                            ctypes.byref(out), ctypes.byref(dt)), "tree rows")
        self.dtype = dt.value
        return self.ctx._wrap(out)

    def GetPredictedSizeAfterCompression(self, data: DeviceBuffer, dtype: Optional[int] = None) -> int:
        arr, n = _path_arg(self.path)
        s = ctypes.c_size_t()
        self.ctx.check(self.ctx._L.ddj_tree_predict_size(self.ctx._h, arr, n, self.dtype if dtype is None else dtype,
                                                         data.dev or None, data.nbytes, ctypes.byref(s)), "predict")
        return s.value


class CompressionOptimizer:
    """CompressionOptimizer (optimizer/compression_optimizer.hpp:28-96)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx._L.ddj_optimizer_create(ctx._h, ctypes.byref(h)), "optimizer create")
        self._h = h
        ctx._track(self)

    def _release(self):
        if self._h and self.ctx._h:
            self.ctx._L.ddj_optimizer_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def _plan(self, mode: int, data: DeviceBuffer, dtype: int) -> Tuple[List[int], float, int]:
        path = (ctypes.c_int * 4096)()
        n = ctypes.c_size_t()
        r = ctypes.c_double()
        cand = ctypes.c_uint64()
        self.ctx.check(self.ctx._L.ddj_plan(self.ctx._h, mode, dtype, data.dev or None, data.nbytes, path, 4096,
                                            ctypes.byref(n), ctypes.byref(r), ctypes.byref(cand)), "plan")
        return list(path[: n.value]), r.value, cand.value

    def OptimizeTree(self, data: DeviceBuffer, dtype: int) -> CompressionTree:
        """The 30-path brute force (compression_optimizer.cpp:16-36)."""
        path, _, _ = self._plan(PLAN_OPTIMIZE30, data, dtype)
        return CompressionTree(self.ctx, path, dtype)

    def PlanExhaustive(self, data: DeviceBuffer, dtype: int) -> Tuple[CompressionTree, float, int]:
        """Phase 1 of CompressData: FullStatisticsUpdate on the first 10,000 rows."""
        path, ratio, cand = self._plan(PLAN_EXHAUSTIVE, data, dtype)
        return CompressionTree(self.ctx, path, dtype), ratio, cand

    def CompressData(self, part: DeviceBuffer, dtype: int) -> DeviceBuffer:
        out = _Buf()
        self.ctx.check(self.ctx._L.ddj_optimizer_compress(self._h, dtype, part.dev or None, part.nbytes,
                                                          ctypes.byref(out)), "CompressData")
        return self.ctx._wrap(out)

    def GetOptimalTree(self) -> List[int]:
        path = (ctypes.c_int * 4096)()
        n = ctypes.c_size_t()
        self.ctx.check(self.ctx._L.ddj_optimizer_tree(self._h, path, 4096, ctypes.byref(n)), "optimizer tree")
        return list(path[: n.value])


class ParallelTSCompressor:
    """ParallelTSCompressor (parallel_ts_compressor.cpp:17-102) over host
    columns: rows cut into batches of `batch_size` (SetBatchSize; the
    reference's default is 1e6), one CompressionOptimizer per column, the
    file holding [size_t bytes][blob] per column per batch."""

    def __init__(self, ctx: Context, dtypes: Sequence[int], batch_size: int = 1_000_000):
        self.ctx = ctx
        self.dtypes = list(dtypes)
        self.batch_size = int(batch_size)
        self._h = None
        self._new()

    def _new(self):
        self._free()
        arr = (ctypes.c_int * len(self.dtypes))(*self.dtypes)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx._L.ddj_ts_create(self.ctx._h, len(self.dtypes), arr, self.batch_size,
                                                 ctypes.byref(h)), "ts create")
        self._h = h
        self.ctx._track(self)

    def _free(self):
        if self._h and self.ctx._h:
            self.ctx._L.ddj_ts_destroy(self._h)
        self._h = None

    def _release(self):
        self._free()

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

    def SetBatchSize(self, size: int):
        self.batch_size = int(size)
        self._new()

    def Compress(self, columns: Sequence[np.ndarray], path: str, append: bool = False):
        cols = [np.ascontiguousarray(c, dtype=NP_DTYPE[d]) for c, d in zip(columns, self.dtypes)]
        if len(cols) != len(self.dtypes) or len({c.size for c in cols}) != 1:
            raise ValueError("one array per column, all of the same length")
        ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
        self.ctx.check(self.ctx._L.ddj_ts_compress(self._h, ptrs, cols[0].size, os.fsencode(path), int(append)),
                       "ts compress")

    def Decompress(self, path: str) -> List[np.ndarray]:
        arr = (ctypes.c_int * len(self.dtypes))(*self.dtypes)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx._L.ddj_ts_reader_open(self.ctx._h, os.fsencode(path), len(self.dtypes), arr,
                                                      ctypes.byref(h)), "ts open")
        parts = [[] for _ in self.dtypes]
        try:
            bufs = [np.empty(self.batch_size, NP_DTYPE[d]) for d in self.dtypes]
            ptrs = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
            rows = ctypes.c_size_t()
            while True:
                self.ctx.check(self.ctx._L.ddj_ts_reader_next(h, ptrs, self.batch_size, ctypes.byref(rows)),
                               "ts next")
                if rows.value == 0:
                    break
                for k, b in enumerate(bufs):
                    parts[k].append(b[: rows.value].copy())
        finally:
            self.ctx._L.ddj_ts_reader_close(h)
        return [np.concatenate(p) if p else np.empty(0, NP_DTYPE[d]) for p, d in zip(parts, self.dtypes)]


FORMAT_BIN, FORMAT_CSV = 0, 1
TYPE_NAMES = {D_INT: "int", D_TIME: "time", D_FLOAT: "float", D_DOUBLE: "double", D_SHORT: "short", D_CHAR: "char"}


def _io_check(st: int, what: str):
    if st != 0:
        raise DdjError(st, f"{what}: {(load_library().ddj_io_last_error() or b'').decode()}")


class FileDefinition:
    """The columns of a time-series file: name, DataType, decimals (the
    header files of the reference's sample_data/, read by
    TimeSeriesReader::ReadFileDefinition, main.cpp:92-94)."""

    def __init__(self, names: Sequence[str], dtypes: Sequence[int], decimals: Optional[Sequence[int]] = None):
        self.names = list(names)
        self.dtypes = list(dtypes)
        self.decimals = list(decimals) if decimals is not None else [0] * len(self.dtypes)

    @staticmethod
    def read(header_path: str) -> "FileDefinition":
        L = load_library()
        h = ctypes.c_void_p()
        _io_check(L.ddj_file_def_read(os.fsencode(header_path), ctypes.byref(h)), "read header")
        try:
            names, dts, decs = [], [], []
            for i in range(L.ddj_file_def_columns(h)):
                nm, dt, dc = ctypes.c_char_p(), ctypes.c_int(), ctypes.c_int()
                _io_check(L.ddj_file_def_column(h, i, ctypes.byref(nm), ctypes.byref(dt), ctypes.byref(dc)), "column")
                names.append(nm.value.decode())
                dts.append(dt.value)
                decs.append(dc.value)
        finally:
            L.ddj_file_def_free(h)
        return FileDefinition(names, dts, decs)

    def _handle(self):
        L = load_library()
        n = len(self.dtypes)
        names = (ctypes.c_char_p * n)(*[x.encode() for x in self.names])
        h = ctypes.c_void_p()
        _io_check(L.ddj_file_def_create(n, names, (ctypes.c_int * n)(*self.dtypes), (ctypes.c_int * n)(*self.decimals),
                                        ctypes.byref(h)), "file definition")
        return h

    def write(self, header_path: str):
        L = load_library()
        h = self._handle()
        try:
            _io_check(L.ddj_file_def_write(h, os.fsencode(header_path)), "write header")
        finally:
            L.ddj_file_def_free(h)


def read_series(definition: FileDefinition, path: str, fmt: int = FORMAT_BIN, padding: int = 0,
                batch: int = 1 << 20) -> List[np.ndarray]:
    """Every record of a binary or CSV time-series file, one array per column."""
    L = load_library()
    d = definition._handle()
    f = ctypes.c_void_p()
    try:
        _io_check(L.ddj_series_open(d, os.fsencode(path), fmt, padding, 0, ctypes.byref(f)), "open")
    finally:
        L.ddj_file_def_free(d)
    parts = [[] for _ in definition.dtypes]
    try:
        bufs = [np.empty(batch, NP_DTYPE[t]) for t in definition.dtypes]
        ptrs = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
        rows = ctypes.c_size_t()
        while True:
            _io_check(L.ddj_series_read(f, batch, ptrs, ctypes.byref(rows)), "read")
            for k, b in enumerate(bufs):
                parts[k].append(b[: rows.value].copy())
            if rows.value < batch:
                break
    finally:
        L.ddj_series_close(f)
    return [np.concatenate(p) for p in parts]


def write_series(definition: FileDefinition, path: str, columns: Sequence[np.ndarray], fmt: int = FORMAT_BIN,
                 padding: int = 0):
    L = load_library()
    cols = [np.ascontiguousarray(c, dtype=NP_DTYPE[t]) for c, t in zip(columns, definition.dtypes)]
    if len(cols) != len(definition.dtypes) or len({c.size for c in cols}) != 1:
        raise ValueError("one array per column, all of the same length")
    d = definition._handle()
    f = ctypes.c_void_p()
    try:
        _io_check(L.ddj_series_open(d, os.fsencode(path), fmt, padding, 1, ctypes.byref(f)), "open")
    finally:
        L.ddj_file_def_free(d)
    try:
        ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
        _io_check(L.ddj_series_write(f, ptrs, cols[0].size), "write")
    finally:
        L.ddj_series_close(f)


CLI_PATH = os.path.join(HERE, "ddj")


MLP_FEATURES, MLP_HIDDEN, MLP_OUTPUTS, MLP_OUTPUTS_PADDED = 32, 128, 48, 64
MLP_PARAM_COUNT = 32 * 128 + 128 + 128 * 128 + 128 + 128 * 64 + 64


PLANNER_INT, PLANNER_FLOAT = 0, 1


def planner_class(dtype: int) -> int:
    """The learned planner's column class of a DataType."""
    return PLANNER_FLOAT if dtype in (D_FLOAT, D_DOUBLE) else PLANNER_INT


def planner_candidates(cls: int = PLANNER_INT) -> List[List[int]]:
    """The 48 candidate trees of a learned-planner class (pre-order paths)."""
    L = load_library()
    offs = (ctypes.c_size_t * (MLP_OUTPUTS + 1))()
    L.ddj_planner_candidates_class(cls, None, 0, offs)
    paths = (ctypes.c_int * offs[MLP_OUTPUTS])()
    st = L.ddj_planner_candidates_class(cls, paths, offs[MLP_OUTPUTS], offs)
    if st:
        raise DdjError(st, "planner candidates")
    return [list(paths[offs[i]:offs[i + 1]]) for i in range(MLP_OUTPUTS)]


def split_params(flat: np.ndarray):
    """Views W1 [32][128], b1, W2 [128][128], b2, W3 [128][64], b3 [64] of a flat parameter vector."""
    o, out = 0, []
    for shape in ((32, 128), (128,), (128, 128), (128,), (128, 64), (64,)):
        n = int(np.prod(shape))
        out.append(flat[o:o + n].reshape(shape))
        o += n
    return out


class LearnedPlanner:
    """The learned planner (no counterpart in the reference): features of a
    block -> predicted log2(compressed / input bytes) of 48 candidate trees,
    trained with the library's own matrix-core forward/backward kernels."""

    def __init__(self, ctx: Context, seed: int = 0, cls: int = PLANNER_INT):
        self.ctx = ctx
        self.cls = cls
        h = ctypes.c_void_p()
        ctx.check(ctx._L.ddj_planner_create_class(ctx._h, cls, seed, ctypes.byref(h)), "planner create")
        self._h = h
        ctx._track(self)

    def _release(self):
        if self._h and self.ctx._h:
            self.ctx._L.ddj_planner_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def Features(self, data: DeviceBuffer, dtype: int) -> np.ndarray:
        out = np.zeros(MLP_FEATURES, np.float32)
        self.ctx.check(self.ctx._L.ddj_planner_features(self.ctx._h, dtype, data.dev or None, data.nbytes,
                                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                       "features")
        return out

    def Labels(self, data: DeviceBuffer, dtype: int) -> np.ndarray:
        out = np.zeros(MLP_OUTPUTS, np.float32)
        self.ctx.check(self.ctx._L.ddj_planner_labels(self.ctx._h, dtype, data.dev or None, data.nbytes,
                                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "labels")
        return out

    def Params(self) -> np.ndarray:
        out = np.zeros(MLP_PARAM_COUNT, np.float32)
        self.ctx.check(self.ctx._L.ddj_planner_params(self._h, out.ctypes.data, None), "params")
        return out

    def SetParams(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == MLP_PARAM_COUNT
        self.ctx.check(self.ctx._L.ddj_planner_params(self._h, None, flat.ctypes.data), "params")

    def Gradient(self, X: np.ndarray, T: np.ndarray) -> Tuple[np.ndarray, float]:
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(MLP_PARAM_COUNT, np.float32)
        loss = ctypes.c_float()
        self.ctx.check(self.ctx._L.ddj_planner_gradient(self._h, X.ctypes.data, T.ctypes.data, X.shape[0],
                                                        g.ctypes.data, ctypes.byref(loss)), "gradient")
        return g, loss.value

    def Train(self, X: np.ndarray, T: np.ndarray, epochs: int, batch: int = 256, lr: float = 1e-3,
              seed: int = 0) -> float:
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = ctypes.c_float()
        self.ctx.check(self.ctx._L.ddj_planner_train(self._h, X.ctypes.data, T.ctypes.data, X.shape[0], epochs, batch,
                                                     lr, seed, ctypes.byref(loss)), "train")
        return loss.value

    def Predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], MLP_OUTPUTS), np.float32)
        self.ctx.check(self.ctx._L.ddj_planner_predict(self._h, X.ctypes.data, X.shape[0], Y.ctypes.data),
                       "predict")
        return Y

    def Plan(self, data: DeviceBuffer, dtype: int, min_margin: float = 0.0) -> Tuple[List[int], bool, float]:
        path = (ctypes.c_int * 4096)()
        n = ctypes.c_size_t()
        fb = ctypes.c_int()
        m = ctypes.c_float()
        self.ctx.check(self.ctx._L.ddj_planner_plan(self._h, dtype, data.dev or None, data.nbytes, min_margin, path,
                                                    4096, ctypes.byref(n), ctypes.byref(fb), ctypes.byref(m)),
                       "learned plan")
        return list(path[: n.value]), bool(fb.value), m.value

    def PlanChecked(self, data: DeviceBuffer, dtype: int, within: float = 0.3,
                    max_checked: int = 4) -> Tuple[List[int], int]:
        """The smallest prediction, or among the candidates predicted within
        `within` (log2) of it (at most max_checked) the smallest exact size:
        (pre-order path, candidates compressed)."""
        path = (ctypes.c_int * 4096)()
        n = ctypes.c_size_t()
        k = ctypes.c_int()
        self.ctx.check(self.ctx._L.ddj_planner_plan_checked(self._h, dtype, data.dev or None, data.nbytes, within,
                                                            max_checked, path, 4096, ctypes.byref(n),
                                                            ctypes.byref(k)), "learned plan (checked)")
        return list(path[: n.value]), k.value

    def Save(self, path: str):
        np.save(path, self.Params(), allow_pickle=False)

    def Load(self, path: str):
        self.SetParams(np.load(path, allow_pickle=False))


class _Column(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("d_in", ctypes.c_void_p), ("in_bytes", ctypes.c_size_t),
                ("preorder", ctypes.POINTER(ctypes.c_int)), ("path_len", ctypes.c_size_t)]


class _BatchConfig(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_uint32), ("range_rows_log2", ctypes.c_uint32), ("wg_per_cu", ctypes.c_uint32)]


def _columns(columns):
    keep = [(ctypes.c_int * len(p))(*p) for _, _, p in columns]
    arr = (_Column * max(len(columns), 1))()
    for i, ((buf, dt, path), pa) in enumerate(zip(columns, keep)):
        arr[i] = _Column(dt, buf.dev or None, buf.nbytes, pa, len(path))
    return arr, keep


def batch_compress(ctx: Context, columns: Sequence[Tuple[DeviceBuffer, int, Sequence[int]]],
                   workers: int = 4, config: Optional[dict] = None) -> List[DeviceBuffer]:
    """Compresses (buffer, dtype, tree path) columns together
    (ddj_batch_compress, or ddj_batch_compress_config with `config`:
    {"workers", "range_rows_log2", "wg_per_cu"}, zero meaning default)."""
    arr, keep = _columns(columns)
    outs = (_Buf * max(len(columns), 1))()
    if config is None:
        ctx.check(ctx._L.ddj_batch_compress(ctx._h, arr, len(columns), workers, outs), "batch compress")
    else:
        cfg = _BatchConfig(config.get("workers", 0), config.get("range_rows_log2", 0), config.get("wg_per_cu", 0))
        ctx.check(ctx._L.ddj_batch_compress_config(ctx._h, arr, len(columns), ctypes.byref(cfg), outs),
                  "batch compress")
    return [ctx._wrap(outs[i]) for i in range(len(columns))]


def batch_autotune(ctx: Context, columns: Sequence[Tuple[DeviceBuffer, int, Sequence[int]]]) -> Tuple[dict, float]:
    """The fastest batch configuration for these columns (ddj_batch_autotune)
    and its time in ms."""
    arr, keep = _columns(columns)
    cfg = _BatchConfig()
    ms = ctypes.c_double()
    ctx.check(ctx._L.ddj_batch_autotune(ctx._h, arr, len(columns), ctypes.byref(cfg), ctypes.byref(ms)),
              "batch autotune")
    return {"workers": cfg.workers, "range_rows_log2": cfg.range_rows_log2, "wg_per_cu": cfg.wg_per_cu}, ms.value
