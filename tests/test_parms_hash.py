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
