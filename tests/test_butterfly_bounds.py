"""Host check of the NTT's lazy butterflies and quotient-estimate Shoup product.

Builds tests/host/butterfly_check.cpp with hipcc (host code only; no GPU needed)
against the same header the gfx950 kernels include, and runs it: congruences against
exact 128-bit arithmetic and the lazy ranges ([0, 8q) forward, [0, 4q) inverse,
[0, 4q) for the twiddle product, and for q < 2^60 the wide forward schedule's
[0, 16q) -> [0, 12q) and [0, 12q) -> [0, 16q)) for primes of 17 to 61 bits with
random and extreme operands.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "seal-cuda_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_butterfly_ranges_and_congruences(tmp_path):
    exe = tmp_path / "butterfly_check"
    src = os.path.join(HERE, "host", "butterfly_check.cpp")
    subprocess.run(["hipcc", "-std=c++17", "-O2", "-I", CSRC, src, "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
