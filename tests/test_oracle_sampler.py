"""The oracle's sampler streams, pinned without the GPU.

* ChaCha20's block function against RFC 8439 section 2.3.2 (the test vector's state
  words 12-15 are this build's 64-bit counter and nonce words) and against OpenSSL's
  ChaCha20 keystream (libcrypto, an independent implementation; its 16-byte IV is the
  same four state words, little endian) for random keys, counters and nonces;
* the keyed stream's word layout (word i = block i / 8, words 2k and 2k + 1);
* the splitmix64 stream against its closed form (Python integers);
* the distributions: ternary in {-1, 0, 1} with frequencies 1/3, centred binomial with
  |s| <= 21 and variance 21/2, uniform words below q, one integer in every limb.
"""
import ctypes
import ctypes.util
import struct

import numpy as np
import pytest

import checkers as ck

KEY = bytes(range(32))


def test_rfc8439_block_vector():
    # RFC 8439 2.3.2: key 00..1f, block count 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00
    out = ck.chacha20_block(KEY, 1 | (0x09000000 << 32), 0x4a000000)
    want = [0xe4e7f110, 0x15593bd1, 0x1fdd0f50, 0xc47120a3, 0xc7f4d1c7, 0x0368c033, 0x9aaa2204, 0x4e6cd4c3,
            0x466482d2, 0x09aa9f07, 0x05d7c214, 0xa2028bd9, 0xd19c12b5, 0xb94e16de, 0xe883d0cb, 0x4e3c50a2]
    assert [int(x) for x in out] == want


def _openssl():
    name = ctypes.util.find_library("crypto")
    if not name:
        pytest.skip("libcrypto not found")
    lib = ctypes.CDLL(name)
    vp = ctypes.c_void_p
    lib.EVP_CIPHER_CTX_new.restype = vp
    lib.EVP_chacha20.restype = vp
    lib.EVP_EncryptInit_ex.argtypes = [vp, vp, vp, ctypes.c_char_p, ctypes.c_char_p]
    lib.EVP_EncryptUpdate.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                      ctypes.c_int]
    lib.EVP_CIPHER_CTX_free.argtypes = [vp]
    return lib


def openssl_keystream(lib, key, counter32, nonce96, nbytes):
    ctx = lib.EVP_CIPHER_CTX_new()
    try:
        iv = struct.pack("<I", counter32) + nonce96
        assert lib.EVP_EncryptInit_ex(ctx, lib.EVP_chacha20(), None, key, iv) == 1
        out = ctypes.create_string_buffer(nbytes)
        n = ctypes.c_int(0)
        assert lib.EVP_EncryptUpdate(ctx, out, ctypes.byref(n), bytes(nbytes), nbytes) == 1
        assert n.value == nbytes
        return out.raw
    finally:
        lib.EVP_CIPHER_CTX_free(ctx)


def test_keystream_matches_openssl():
    lib = _openssl()
    rng = np.random.default_rng(5)
    for _ in range(20):
        key = rng.bytes(32)
        counter = int(rng.integers(0, 1 << 30))  # low counter word, no carry over 4 blocks
        counter_hi = int(rng.integers(0, 1 << 32))
        nonce = int(rng.integers(0, 1 << 63))
        # state words 12..15 = counter lo, counter hi, nonce lo, nonce hi
        nonce96 = struct.pack("<IQ", counter_hi, nonce)
        want = openssl_keystream(lib, key, counter, nonce96, 4 * 64)
        got = b"".join(ck.chacha20_block(key, counter + b + (counter_hi << 32), nonce).astype("<u4").tobytes()
                       for b in range(4))
        assert got == want


def test_keyed_stream_word_layout():
    logn, n = 6, 64
    q = ck.coeff_modulus_create(1 << logn, [40])
    seed = 0x1234_5678_9abc_def0
    key = bytes(range(100, 132))
    # uniform output i = (word 2i << 64 | word 2i+1) mod q, word j = block j / 8, words 2k, 2k+1
    u = ck.sample_poly(2, key, seed, logn, q)[0]

    def word(j):
        b = ck.chacha20_block(key, j // 8, seed)
        k = j % 8
        return int(b[2 * k]) | (int(b[2 * k + 1]) << 32)

    for i in (0, 1, 3, 4, 17, n - 1):
        assert int(u[i]) == ((word(2 * i) << 64) | word(2 * i + 1)) % q[0]


def test_splitmix_stream_closed_form():
    logn, n = 5, 32
    q = ck.coeff_modulus_create(1 << logn, [50])
    seed = 77
    M = (1 << 64) - 1

    def word(i):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    u = ck.sample_poly(2, None, seed, logn, q)[0]
    assert [int(x) for x in u] == [((word(2 * i) << 64) | word(2 * i + 1)) % q[0] for i in range(n)]
    t = ck.sample_poly(0, None, seed, logn, q)[0]
    assert [(int(x) - q[0] if int(x) > q[0] // 2 else int(x)) for x in t] == \
        [((word(c) * 3) >> 64) - 1 for c in range(n)]


@pytest.mark.parametrize("key", [None, KEY])
def test_distributions(key):
    logn = 14
    moduli = ck.coeff_modulus_create(1 << logn, [50, 50, 60])
    t = ck.sample_poly(0, key, 3, logn, moduli)
    c = [np.where(t[i] > moduli[i] // 2, t[i].astype(np.int64) - np.int64(moduli[i]), t[i].astype(np.int64))
         for i in range(3)]
    assert set(np.unique(c[0])) <= {-1, 0, 1}
    assert all(np.array_equal(c[0], x) for x in c[1:])
    for v in (-1, 0, 1):
        assert abs(np.mean(c[0] == v) - 1 / 3) < 0.02
    e = ck.sample_poly(1, key, 4, logn, moduli)
    s = np.where(e[0] > moduli[0] // 2, e[0].astype(np.int64) - np.int64(moduli[0]), e[0].astype(np.int64))
    assert np.abs(s).max() <= 21 and abs(s.var() - 10.5) < 0.5 and abs(s.mean()) < 0.1
    u = ck.sample_poly(2, key, 5, logn, moduli)
    for i, q in enumerate(moduli):
        assert u[i].max() < q
        assert abs(np.mean(u[i].astype(np.float64)) / q - 0.5) < 0.02
    # a different key or seed gives an unrelated stream
    assert not np.array_equal(ck.sample_poly(2, key, 6, logn, moduli), u)
