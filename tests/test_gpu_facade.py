"""The troy_cuda-style C++ API (include/troyhip.hpp) end to end on the GPU.

Runs tests/cpp/facade_test and tests/cpp/troyn_kat_test (built with the library by
__graft_entry__.build() / make -C seal-cuda_amd): BFV and CKKS pipelines through KeyGenerator, Encryptor,
Evaluator, Decryptor and the encoders, each result checked against the plaintext
arithmetic it encrypts, and the reference's exception types for invalid calls.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "facade_test")
KAT = os.path.join(HERE, "cpp", "troyn_kat_test")


def test_facade_binaries_are_built():
    for exe in (EXE, KAT):
        assert os.path.exists(exe), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_facade_end_to_end():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("facade ok")


@pytest.mark.gpu
def test_reference_tests_through_troyn_alias():
    """tests/cpp/troyn_kat_test.cpp: the reference's BFVRelinearize decrypt KATs
    (test/evaluator_cuda.cu:2459-2534) compiled with namespace troyn = troyhip, plus
    hexadecimal plaintexts, plaintext files, parms_id hash blocks, cpu() conversions,
    KSwitchKeys, per-generator key material and the pooled steady state."""
    out = subprocess.run([KAT], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("troyn kat ok")
