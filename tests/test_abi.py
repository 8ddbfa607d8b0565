"""CPU-side checks of the C ABI: the library loads and exports every entry point
declared in include/sealhip.h, and host-only entry points behave (no GPU needed)."""
import ctypes
import os
import re

import pytest

import checkers as ck
import sealhip

HEADER = os.path.join(ck.ROOT, "include", "sealhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sealhip.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 6
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_coeff_modulus_create_matches_oracle():
    for n, bits in [(8192, [43, 43, 44, 44, 44]), (65536, [60] * 25), (32768, [60] + [40] * 14 + [60]),
                    (4096, [30, 30, 40]), (2, [3, 5, 4, 5])]:
        assert sealhip.coeff_modulus_create(n, bits) == ck.coeff_modulus_create(n, bits)
    assert sealhip.coeff_modulus_create(8192, [40, 50], 65537) == ck.coeff_modulus_create(8192, [40, 50], 65537)


def test_coeff_modulus_create_errors():
    with pytest.raises(sealhip.SealHipError) as e:
        sealhip.coeff_modulus_create(1000, [30])
    assert e.value.status == -1
    with pytest.raises(sealhip.SealHipError):
        sealhip.coeff_modulus_create(1024, [61])
