"""The C++ mirror classes (include/ddj.hpp) on the GPU: a host program
built with g++ round-trips columns through Encoding, CompressionTree,
CompressionOptimizer and ParallelTSCompressor."""
import subprocess

import pytest

from test_capi import _build_cpp

pytestmark = pytest.mark.gpu


def test_cpp_roundtrip(tmp_path):
    exe = _build_cpp(str(tmp_path / "roundtrip"))
    r = subprocess.run([exe, str(tmp_path / "cols.ddj")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cpp roundtrip ok" in r.stdout
