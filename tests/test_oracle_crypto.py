"""Pins the oracle's encryption, decryption, key-switching key generation and BatchEncoder.

Decryption and the BatchEncoder are deterministic, so they are compared word for word
with the reference's own Decryptor / BatchEncoder (oracle/_ref) on arbitrary words.
Encryption and key generation take explicit randomness; they are pinned by what
decryption must return (the message, or the key structure P * s' + e) and by the
reference Decryptor decrypting the oracle's ciphertexts.
"""
import numpy as np
import pytest

import checkers as ck

BFV, CKKS = 1, 2
ref_only = pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")


def small_poly(rng, n, bound):
    return rng.integers(-bound, bound + 1, size=n)


def keygen(rng, moduli, logn):
    """secret key (ternary, NTT form at key level) and public key (symmetric zero)."""
    n = 1 << logn
    s = small_poly(rng, n, 1)
    sk = ck.ntt_rns(ck.rns_of_small(s, moduli), moduli, logn)
    a = ck.uniform_residues(rng, (len(moduli), n), moduli)
    e = ck.rns_of_small(small_poly(rng, n, 5), moduli)
    pk = ck.encrypt_zero_symmetric(CKKS, sk, a, e, moduli, logn)  # key level, NTT form
    return s, sk, pk


def bfv_encrypt(rng, pk, plain, moduli, t, logn):
    """Encryptor::encrypt (src/encryptor.cpp:88-233): zero at the level above the first
    data level (= key level), divide-and-round by the special prime, then + Delta m."""
    n = 1 << logn
    K = len(moduli)
    u = ck.rns_of_small(small_poly(rng, n, 1), moduli)
    e0 = ck.rns_of_small(small_poly(rng, n, 5), moduli)
    e1 = ck.rns_of_small(small_poly(rng, n, 5), moduli)
    z = ck.encrypt_zero_asymmetric(BFV, pk, u, e0, e1, moduli, logn)
    ct = ck.rescale(BFV, z, moduli, logn)
    ct[0] = ck.bfv_add_plain_scaled(ct[0], plain, moduli[:K - 1], t, logn)
    return ct


@pytest.mark.parametrize("logn,bits,t", [(10, [40, 40, 50], 65537), (12, [43, 43, 44, 44, 44], 65537)])
def test_bfv_round_trip_with_oracle_keys(logn, bits, t):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    rng = np.random.default_rng(logn)
    s, sk, pk = keygen(rng, moduli, logn)
    values = rng.integers(0, t, size=n, dtype=np.uint64)
    plain = ck.batch_encode(values, t, logn)
    ct = bfv_encrypt(rng, pk, plain, moduli, t, logn)
    dec = ck.bfv_decrypt(ct, sk[:len(moduli) - 1], moduli, t, logn)
    assert np.array_equal(dec, plain)
    assert np.array_equal(ck.batch_decode(dec, t, logn), values)


def test_kswitch_key_structure():
    """Decrypting digit j of a key-switching key gives P * s' in limb j (- noise)."""
    logn = 8
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [40, 40, 40, 50])
    K = len(moduli)
    rng = np.random.default_rng(3)
    s, sk, pk = keygen(rng, moduli, logn)
    s2 = ck.rns_of_small(small_poly(rng, n, 1), moduli)
    new_key = ck.ntt_rns(s2, moduli, logn)
    a = ck.uniform_residues(rng, (K - 1, K, n), moduli)
    e = np.stack([ck.rns_of_small(small_poly(rng, n, 4), moduli) for _ in range(K - 1)])
    key = ck.generate_kswitch_key(sk, new_key, a, e, moduli, logn)
    P = moduli[-1]
    for j in range(K - 1):
        phase = ck.ntt_rns(ck.decrypt_phase(CKKS, key[j], sk, moduli, logn), moduli, logn, inverse=True)
        for i, q in enumerate(moduli):
            # c0 = -(a s + e) (+ P s' in limb j), c1 = a: phase = P s' - e
            want = (-e[j][i].astype(object) + (P * s2[i].astype(object) if i == j else 0)) % q
            assert np.array_equal(phase[i].astype(object), want)


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("scheme,logn,bits,L", [
    (BFV, 10, [40, 40, 50], 2),
    (BFV, 12, [43, 43, 44, 44, 44], 4),
    (BFV, 12, [43, 43, 44, 44, 44], 2),
    (CKKS, 10, [60, 40, 40, 60], 3),
])
def test_decrypt_matches_reference_on_arbitrary_words(scheme, logn, bits, L):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    t = 65537
    sess = ck.RefSession(scheme, n, moduli, t if scheme == BFV else 0)
    rng = np.random.default_rng(logn * L)
    sk = ck.ntt_rns(ck.rns_of_small(small_poly(rng, n, 1), moduli), moduli, logn)
    ct = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    ref = sess.decrypt(ct, sk, L)
    if scheme == BFV:
        assert np.array_equal(ck.bfv_decrypt(ct, sk[:L], moduli, t, logn), ref)
    else:
        assert np.array_equal(ck.decrypt_phase(CKKS, ct, sk[:L], moduli, logn), ref)


@ref_only
@pytest.mark.ref
def test_reference_decrypts_oracle_encryption():
    logn, t = 12, 65537
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [43, 43, 44, 44, 44])
    sess = ck.RefSession(BFV, n, moduli, t)
    rng = np.random.default_rng(9)
    s, sk, pk = keygen(rng, moduli, logn)
    values = rng.integers(0, t, size=n, dtype=np.uint64)
    plain = sess.batch_encode(values)
    ct = bfv_encrypt(rng, pk, plain, moduli, t, logn)
    assert np.array_equal(sess.decrypt(ct, sk, len(moduli) - 1), plain)


@ref_only
@pytest.mark.ref
def test_reference_encryption_shim_round_trips():
    """The reference's own Encryptor behind ref_encrypt (timed by the benchmark's CPU
    baseline) encrypts under an injected public key: BFV decrypts to the plaintext
    exactly, CKKS to the plaintext plus a small error (|e| < 2^20 per coefficient)."""
    logn, t = 12, 65537
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [43, 43, 44, 44, 44])
    rng = np.random.default_rng(17)
    s, sk, pk = keygen(rng, moduli, logn)
    L = len(moduli) - 1
    bfv = ck.RefSession(BFV, n, moduli, t)
    plain = bfv.batch_encode(rng.integers(0, t, size=n, dtype=np.uint64))
    ct = bfv.encrypt(pk, plain, L)
    assert np.array_equal(bfv.decrypt(ct, sk, L), plain)
    ckks = ck.RefSession(CKKS, n, moduli)
    for level in (L, 2):
        pl = ck.uniform_residues(rng, (level, n), moduli[:level])
        dec = ckks.decrypt(ckks.encrypt(pk, pl, level), sk, level)
        diff = ck.ntt_rns(ck.mod_sub(dec, pl, moduli), moduli[:level], logn, inverse=True)
        for i, q in enumerate(moduli[:level]):
            d = diff[i].astype(object)
            assert all(min(int(x), q - int(x)) < (1 << 20) for x in d)


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("logn,t", [(10, 65537), (13, 65537), (12, 786433)])
def test_batch_encoder_matches_reference(logn, t):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [40, 40, 50]) if logn < 13 else ck.coeff_modulus_create(n, [43, 43, 44])
    sess = ck.RefSession(BFV, n, moduli, t)
    rng = np.random.default_rng(logn)
    for count in (n, n // 3, 0):
        values = rng.integers(0, t, size=count, dtype=np.uint64)
        assert np.array_equal(ck.batch_encode(values, t, logn), sess.batch_encode(values))
    plain = rng.integers(0, t, size=n, dtype=np.uint64)
    assert np.array_equal(ck.batch_decode(plain, t, logn), sess.batch_decode(plain))
