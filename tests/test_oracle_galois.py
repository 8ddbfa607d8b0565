"""Pins the oracle's Galois key switch (applyGaloisInplace) against the reference's own
Evaluator on injected words, and checks the automorphism maps against the
mathematics: sigma_elt is a ring automorphism (it commutes with the NTT-domain
product) and composing the maps of elt and elt^-1 is the identity."""
import numpy as np
import pytest

import checkers as ck
from checkers import ptr

BFV, CKKS = 1, 2
ref_only = pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")


def test_automorphism_inverse_and_multiplicative():
    logn = 6
    n = 1 << logn
    q = ck.get_primes(2 * n, 50, 1)[0]
    rng = np.random.default_rng(1)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    b = rng.integers(0, q, size=n, dtype=np.uint64)
    o = ck.oracle()
    for step in (1, 3, -1, 0):
        elt = ck.galois_elt(logn, step)
        inv = pow(elt, -1, 2 * n)
        s, back = np.zeros(n, dtype=np.uint64), np.zeros(n, dtype=np.uint64)
        o.or_apply_galois(logn, elt, q, ptr(a), ptr(s))
        o.or_apply_galois(logn, inv, q, ptr(s), ptr(back))
        assert np.array_equal(back, a)
        # sigma(a * b) == sigma(a) * sigma(b) in the NTT domain
        fa, fb = ck.ntt(a, q, logn), ck.ntt(b, q, logn)
        prod = (fa.astype(object) * fb.astype(object) % q).astype(np.uint64)
        sp, sa, sb = (np.zeros(n, dtype=np.uint64) for _ in range(3))
        o.or_apply_galois_ntt(logn, elt, ptr(prod), ptr(sp))
        o.or_apply_galois_ntt(logn, elt, ptr(fa), ptr(sa))
        o.or_apply_galois_ntt(logn, elt, ptr(fb), ptr(sb))
        assert np.array_equal(sp, (sa.astype(object) * sb.astype(object) % q).astype(np.uint64))


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("scheme,logn,bits,L,steps", [
    (CKKS, 10, [60, 40, 40, 60], 3, [1, -1, 5, 0]),
    (CKKS, 12, [50, 40, 40, 40, 50], 2, [7, -64]),
    (BFV, 10, [40, 40, 50], 2, [1, 3, 0]),
    (BFV, 12, [43, 43, 44, 44, 44], 4, [-1, 100]),
])
def test_galois_keyswitch_matches_reference(scheme, logn, bits, L, steps):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    sess = ck.RefSession(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(logn + L)
    for step in steps:
        elt = ck.galois_elt(logn, step)
        ct = ck.uniform_residues(rng, (2, L, n), moduli[:L])
        key = ck.random_key(rng, moduli, logn)
        assert np.array_equal(ck.apply_galois_keyswitch(scheme, ct, elt, key, moduli, L, logn),
                              sess.apply_galois(ct, elt, key, L))
