"""The C++ API's parms_id (BLAKE2b-256 of the parameter words, include/troyhip.hpp) on
the host: BLAKE2b against RFC 7693's "abc" vector and the empty-input digests, and
parms_id against the reference library's own EncryptionParameters::parmsID for every
level of BFV and CKKS chains -- the key that makes saved ciphertexts portable between
the two."""
import os
import shutil
import subprocess

import pytest

import checkers as ck

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(HERE, "..", "include")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if shutil.which("hipcc") is None:
        pytest.skip("hipcc not on PATH")
    out = tmp_path_factory.mktemp("ph") / "parms_hash_check"
    subprocess.run(["hipcc", "-std=c++17", "-O1", "-I", INCLUDE, os.path.join(HERE, "host", "parms_hash_check.cpp"),
                    "-o", str(out)], check=True, capture_output=True)
    return str(out)


def test_blake2b_known_answers(exe):
    out = subprocess.run([exe, "blake2b", "abc"], capture_output=True, text=True, check=True).stdout.split()
    # RFC 7693 appendix A (BLAKE2b-512 of "abc") and the published BLAKE2b-256 of "abc"
    assert out[0] == ("ba80a53f981c4d0d6a2797b69f12f6e94c212f14685ac4b74b12bb6fdbffa2d1"
                      "7d87c5392aab792dc252d5de4533cc9518d38aa8dbf1925ab92386edd4009923")
    assert out[1] == "bddd813c634239723171ef3fee98579b94964e3bb1cb3e427262c8c068d52319"


@pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")
@pytest.mark.ref
@pytest.mark.parametrize("scheme,logn,bits,t", [
    (1, 12, [40, 40, 50], 65537),
    (2, 13, [60, 40, 40, 40, 60], 0),
    (2, 15, [60] + [40] * 14 + [60], 0),
])
def test_parms_id_matches_reference(exe, scheme, logn, bits, t):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    sess = ck.RefSession(scheme, n, moduli, t)
    for L in range(1, len(moduli) + 1):
        args = [exe, "parms", str(scheme), str(n), str(t)] + [str(q) for q in moduli[:L]]
        got = [int(x) for x in subprocess.run(args, capture_output=True, text=True, check=True).stdout.split()]
        assert got == sess.parms_id(L), f"level with {L} primes"


HEX_CASES = ["7FFx^3 + 1x^1 + 3", "1x^10 + 2", "abcx^2 + f", "0", "0x^5 + 1", "1x^40 + 8x^30 + 18x^20 + 20x^10 + 10",
             "FFFFFFFFFFFFFFFFx^1", "10000000000000000x^1", "000000000000000001x^2", "1x^2 + 1x^3", "1x^1 +1",
             "x^2", "1x^2 + ", "1x2", "1 + 1", "1x^ + 1", "-1", "1x^1", "1x^0", "", "1x^2 + 0x^1 + 0",
             "3x^2147483646", "1 ", " 1"]


def hexpoly(exe, s):
    out = subprocess.run([exe, "hexpoly", s], capture_output=True, text=True, check=True).stdout.split("\n")
    if out[0] == "invalid":
        return None
    k = int(out[0])
    return [int(x) for x in out[1:1 + k]], out[1 + k]


@pytest.mark.parametrize("s", ["7FFx^3 + 1x^1 + 3", "abcx^2 + f", "1x^10 + 2"])
def test_hex_plaintext_format(exe, s):
    """Hexadecimal plaintext strings: coefficients by power, upper-case output."""
    coeffs, text = hexpoly(exe, s)
    want = {"7FFx^3 + 1x^1 + 3": [3, 1, 0, 0x7FF], "abcx^2 + f": [15, 0, 0xABC],
            "1x^10 + 2": [2] + [0] * 9 + [1]}[s]
    assert coeffs == want and text == s.upper().replace("X", "x")


@pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")
@pytest.mark.ref
@pytest.mark.parametrize("s", [c for c in HEX_CASES if "2147483646" not in c])
def test_hex_plaintext_matches_reference(exe, s):
    """Plaintext(hex) and to_string() against the reference's own parser (src/plaintext.cpp:91-200):
    the same coefficients and string for valid input, a rejection for invalid input."""
    import ctypes
    import numpy as np
    cap = 64
    coeffs = np.zeros(cap, dtype=np.uint64)
    count = ctypes.c_uint64(cap)
    buf = ctypes.create_string_buffer(4096)
    rc = ck.ref().ref_plaintext_hex(s.encode(), ck.ptr(coeffs), ctypes.byref(count), buf, 4096)
    got = hexpoly(exe, s)
    if rc == -1:
        assert got is None, f"{s!r}: the reference rejects it"
        return
    assert rc == 0 and got is not None, f"{s!r}: the reference accepts it"
    want = [int(x) for x in coeffs[:count.value]]
    # an all-zero string leaves the reference plaintext at its previous (empty) size
    assert got[0] == want, s
    assert got[1] == buf.value.decode(), s
