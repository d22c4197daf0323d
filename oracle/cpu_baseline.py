"""CPU baselines -- MEASUREMENT INFRASTRUCTURE ONLY (bench.py's cpu_baseline
leg and tests/ use it; the product never does).

* :func:`fl_loops` times the reference's only CPU path, its FL host loops
  (run_afl_compress_cpu<T,1> / run_afl_decompress_cpu<T,1>,
  compression/afl/afl_gpu.cu:6-65, restated in fl_loops.h and built by the
  Makefile with g++ -O3 for x86-64-v3 and -v4), on one host thread and on
  several (groups of 32 values split between threads).
* :func:`oracle_trees` times the oracle (the reference's codecs restated,
  one thread per column part) compressing the same trees as the GPU bench.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")


def host_cores() -> dict:
    """Core counts of this host: os.cpu_count, the scheduler affinity, nproc
    and std::thread::hardware_concurrency."""
    out = {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip())
    except (OSError, ValueError, subprocess.CalledProcessError):
        out["nproc"] = None
    out["hardware_concurrency"] = int(_fl_lib().flcpu_hardware_concurrency())
    return out


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln for ln in f if ln.startswith("flags")), "").split()
    except OSError:
        return False
    return all(k in flags for k in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))


_lib = None
_level = None


def _fl_lib():
    global _lib, _level
    if _lib is None:
        _level = "x86-64-v4" if _has_avx512() else "x86-64-v3"
        path = os.path.join(BUILD, "libddj_flcpu_v4.so" if _level == "x86-64-v4" else "libddj_flcpu_v3.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(path)
        L.flcpu_hardware_concurrency.restype = ctypes.c_uint
        L.flcpu_bench.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                  ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def fl_loops(x: np.ndarray, bits: int, threads: int, reps: int = 10):
    """Median seconds of FL pack and unpack of `x` to `bits` bits on `threads`
    host threads (after two warm-ups), the packed stream and the unpacked
    values."""
    L = _fl_lib()
    x = np.ascontiguousarray(x)
    W = 8 * x.itemsize
    words = (bits * x.size + W - 1) // W
    packed = np.zeros(words + 1, x.dtype)
    out = np.zeros_like(x)
    ps, us = ctypes.c_double(), ctypes.c_double()
    rc = L.flcpu_bench(x.itemsize, bits, x.ctypes.data, x.size, packed.ctypes.data, words, out.ctypes.data,
                       threads, reps, ctypes.byref(ps), ctypes.byref(us))
    if rc != 0:
        raise RuntimeError(f"flcpu_bench failed ({rc})")
    return ps.value, us.value, packed[:words], out


def isa_level() -> str:
    _fl_lib()
    return _level


def oracle_trees(columns, budget_s: float, threads: int):
    """The oracle compressing each (tree path, column sample) pair, repeated
    for about budget_s seconds; with threads > 1 each column sample is cut
    into `threads` parts compressed concurrently (ctypes releases the GIL
    inside the oracle).  Returns (input bytes per second, passes, seconds)."""
    from oracle import oracle as O

    parts = []
    for path, x in columns:
        tree = O.tree_from_path(path, O.D_INT)
        for piece in np.array_split(x, threads):
            parts.append((tree, np.ascontiguousarray(piece)))
    nbytes = sum(p.nbytes for _, p in parts)
    reps, t0 = 0, time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        while True:
            list(ex.map(lambda tp: O.compress(tp[0].copy(), tp[1]), parts))
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget_s:
                break
    return nbytes * reps / el, reps, el
