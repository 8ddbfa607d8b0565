"""Pins the oracle's CKKS encoder against the reference CKKSEncoder (oracle/_ref) and
against the mathematics of the canonical embedding.

Tolerances (double precision, so not bit-exact; hazard of scale/n rounding):
  * encode: plaintext words equal the reference's except rare coefficients whose
    rounding differs by one;
  * decode: within 2^-30 relative (absolute for |v| < 1) of the reference;
  * decode(encode(v)) = v within 2^-20 at scale 2^40.
"""
import numpy as np
import pytest

import checkers as ck

CKKS = 2
ref_only = pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")


def rand_slots(rng, count):
    return rng.uniform(-1, 1, count) + 1j * rng.uniform(-1, 1, count)


@pytest.mark.parametrize("logn,bits,L,scale", [(10, [60, 40, 40, 60], 3, 2.0 ** 40), (12, [60, 50, 50, 60], 2, 2.0 ** 50)])
def test_round_trip(logn, bits, L, scale):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    rng = np.random.default_rng(logn)
    v = rand_slots(rng, n // 2)
    p = ck.ckks_encode(v, moduli[:L], scale, logn)
    back = ck.ckks_decode(p, moduli[:L], scale, logn)
    assert np.max(np.abs(back - v)) < 2.0 ** -20


def test_encode_is_canonical_embedding():
    """The plaintext's coefficient polynomial m(X) satisfies m(zeta^(3^i)) = scale * v_i."""
    logn = 5
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [50, 50])
    rng = np.random.default_rng(2)
    v = rand_slots(rng, n // 2)
    scale = 2.0 ** 30
    p = ck.ckks_encode(v, moduli, scale, logn)
    coeffs = ck.ntt_rns(p, moduli, logn, inverse=True)[0].astype(object)
    q = moduli[0]
    m = np.array([int(c) - q if int(c) > q // 2 else int(c) for c in coeffs], dtype=np.float64)
    for i in range(n // 2):
        k = pow(3, i, 2 * n)
        z = np.exp(1j * np.pi * k / n)
        val = np.sum(m * z ** np.arange(n)) / scale
        assert abs(val - v[i]) < 1e-6


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("logn,bits,L,scale,count", [
    (10, [60, 40, 40, 60], 3, 2.0 ** 40, 512),
    (12, [60, 40, 40, 40, 60], 4, 2.0 ** 40, 1000),
    (13, [60, 50, 60], 2, 2.0 ** 50, 4096),
    (11, [60, 60, 60, 60], 3, 2.0 ** 80, 1024),   # coefficients above 64 bits
])
def test_matches_reference(logn, bits, L, scale, count):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    sess = ck.RefSession(CKKS, n, moduli)
    rng = np.random.default_rng(logn + L)
    v = rand_slots(rng, count)
    ours = ck.ckks_encode(v, moduli[:L], scale, logn)
    theirs = sess.ckks_encode(v, L, scale)
    coef_o = ck.ntt_rns(ours, moduli[:L], logn, inverse=True).astype(object)
    coef_t = ck.ntt_rns(theirs, moduli[:L], logn, inverse=True).astype(object)
    diff = np.mod(coef_o - coef_t, moduli[0])
    diff = np.minimum(diff, moduli[0] - diff)
    assert diff.max() <= 1 and np.mean(diff != 0) < 1e-2
    d_o = ck.ckks_decode(theirs, moduli[:L], scale, logn)
    d_t = sess.ckks_decode(theirs, L, scale)
    assert np.max(np.abs(d_o - d_t) / np.maximum(1, np.abs(d_t))) < 2.0 ** -30
