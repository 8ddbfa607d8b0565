"""Host check of the request server's arena allocator (seal-cuda_amd/csrc/range_alloc.h):
tests/host/range_alloc_check.cpp drives it through a random take / give sequence against a
granule-by-granule model (no overlap, first fit, coalescing, failure only when nothing
fits).  Host code only: no GPU needed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "seal-cuda_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not on PATH")
def test_range_allocator_against_a_model(tmp_path):
    exe = tmp_path / "range_alloc_check"
    src = os.path.join(HERE, "host", "range_alloc_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, src, "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
