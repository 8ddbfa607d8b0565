"""The troy_cuda-style C++ API (include/troyhip.hpp) end to end on the GPU.

Runs tests/cpp/facade_test (built with the library by __graft_entry__.build() /
make -C seal-cuda_amd): BFV and CKKS pipelines through KeyGenerator, Encryptor,
Evaluator, Decryptor and the encoders, each result checked against the plaintext
arithmetic it encrypts, and the reference's exception types for invalid calls.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "facade_test")


def test_facade_binary_is_built():
    assert os.path.exists(EXE), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_facade_end_to_end():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("facade ok")
