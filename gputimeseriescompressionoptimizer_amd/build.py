"""Build libddj_hip.so in-tree with hipcc for gfx950, and the `ddj` CLI.

Every .hip / .cpp under csrc/ is compiled to an object in _build/ (in
parallel) and linked into gputimeseriescompressionoptimizer_amd/libddj_hip.so;
cli/main.cpp (host code over the C ABI only) is linked against it into
gputimeseriescompressionoptimizer_amd/ddj.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libddj_hip.so")
CLI = os.path.join(HERE, "ddj")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("DDJ_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-result",
    f"-I{os.path.join(HERE, '..', 'include')}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, src + ".o")
    cmd = [HIPCC] + COMMON
    cmd += ["-x", "hip"] + DEVICE  # hipcc treats .cpp as HIP as well; one target for all
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sources()
    workers = min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


def build_cli() -> str:
    """The command-line compressor: host C++ linked against libddj_hip.so (rpath $ORIGIN)."""
    tmp = CLI + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(HERE, '..', 'include')}",
           os.path.join(HERE, "cli", "main.cpp"), "-o", tmp, f"-L{HERE}", "-l:libddj_hip.so",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
