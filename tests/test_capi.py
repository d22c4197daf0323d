"""The C ABI library loads and exports every entry point include/ddj.h declares
(no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ddj.h")
LIB = os.path.join(ROOT, "gputimeseriescompressionoptimizer_amd", "libddj_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ddj_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["ddj_encode", "ddj_decode", "ddj_tree_compress", "ddj_tree_decompress", "ddj_plan", "ddj_stats"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libddj_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_encoding_info_without_gpu():
    import gputimeseriescompressionoptimizer_amd as g

    # pure host function: arity and return types of every encoding
    assert g.encoding_info(g.RLE, g.D_DOUBLE) == (2, [g.D_INT, g.D_DOUBLE])
    assert g.encoding_info(g.FLOAT_TO_INT, g.D_FLOAT) == (1, [g.D_INT])
    assert g.encoding_info(g.NONE, g.D_CHAR) == (0, [])
    assert g.encoding_info(g.DICT, g.D_SHORT) == (2, [g.D_SHORT, g.D_SHORT])
    assert g.encoding_info(g.GFC, g.D_FLOAT) == (2, [g.D_FLOAT, g.D_FLOAT])
    with pytest.raises(g.DdjError):
        g.encoding_info(42, g.D_INT)


def _build_cpp(out):
    """Compiles tests/cpp/roundtrip.cpp against include/ddj.hpp and the library."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "gputimeseriescompressionoptimizer_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{os.path.join(root, 'include')}",
           os.path.join(root, "tests", "cpp", "roundtrip.cpp"), f"-L{libdir}", "-lddj_hip", f"-Wl,-rpath,{libdir}",
           "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_cpp_mirror_compiles_and_links(tmp_path):
    """include/ddj.hpp (the reference's class names over the C ABI) builds
    with a plain host compiler and links against the library."""
    assert os.path.exists(_build_cpp(str(tmp_path / "roundtrip")))


def test_profiled_kernel_names_are_listed():
    """Every kernel the library times (KernelTimer names in csrc/) is one the
    Python profiler asks for, so benchmark tables miss none."""
    import gputimeseriescompressionoptimizer_amd as g

    csrc = os.path.join(ROOT, "gputimeseriescompressionoptimizer_amd", "csrc")
    timed = set()
    for f in os.listdir(csrc):
        if f.endswith((".hip", ".cpp", ".h")):
            timed |= set(re.findall(r'KernelTimer kt_\(c, "([a-z_0-9]+)"\)', open(os.path.join(csrc, f)).read()))
    assert timed, "no timed kernels found"
    assert timed <= set(g.Context.KERNELS), sorted(timed - set(g.Context.KERNELS))


def test_headers_compile_warning_free(tmp_path):
    """include/ddj.h (C and C++) and include/ddj.hpp compile with -Wall -Werror."""
    import subprocess

    inc = os.path.join(ROOT, "include")
    src = tmp_path / "t.cpp"
    src.write_text('#include "ddj.h"\n#include "ddj.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", f"-I{inc}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    csrc = tmp_path / "t.c"
    csrc.write_text('#include "ddj.h"\nint main(void) { return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", f"-I{inc}", str(csrc)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
