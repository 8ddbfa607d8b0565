"""Pins the oracle's tensor product, key switching and rescale.

Against the reference's own Evaluator (oracle/_ref) on identical injected words, and
against closed forms that a dropped term or wrong index would break: the tensor
product equals a schoolbook negacyclic product, rescale equals CRT-compose /
divide-and-round / decompose with Python integers, and key switching is linear in
the target and maps a zero key to an unchanged ciphertext.
"""
import numpy as np
import pytest

import checkers as ck

BFV, CKKS = 1, 2
ref_only = pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")


def params(logn, bits):
    n = 1 << logn
    return n, ck.coeff_modulus_create(n, bits)


def test_tensor_is_negacyclic_product():
    logn = 5
    n, moduli = params(logn, [40, 40, 50])
    L = 2
    rng = np.random.default_rng(1)
    a = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    out = ck.ckks_multiply(a, b, moduli[:L], logn)
    coef = lambda x: ck.ntt_rns(x, moduli[:L], logn, inverse=True)
    A, B, O = coef(a), coef(b), coef(out)
    for i, q in enumerate(moduli[:L]):
        def pm(x, y):
            r = [0] * n
            for s in range(n):
                for t in range(n):
                    v = int(x[s]) * int(y[t])
                    if s + t >= n:
                        r[s + t - n] -= v
                    else:
                        r[s + t] += v
            return [v % q for v in r]
        c1 = [(u + v) % q for u, v in zip(pm(A[0, i], B[1, i]), pm(A[1, i], B[0, i]))]
        assert [int(x) for x in O[0, i]] == pm(A[0, i], B[0, i])
        assert [int(x) for x in O[1, i]] == c1
        assert [int(x) for x in O[2, i]] == pm(A[1, i], B[1, i])


@pytest.mark.parametrize("scheme", [BFV, CKKS])
def test_rescale_is_crt_divide_and_round(scheme):
    logn = 4
    n, moduli = params(logn, [30, 30, 40, 50])
    L = 3
    rng = np.random.default_rng(2)
    ct = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    out = ck.rescale(scheme, ct, moduli, logn)
    coef_in = ck.ntt_rns(ct, moduli[:L], logn, inverse=True) if scheme == CKKS else ct
    coef_out = ck.ntt_rns(out, moduli[:L - 1], logn, inverse=True) if scheme == CKKS else out
    Q = 1
    for q in moduli[:L]:
        Q *= q
    qL = moduli[L - 1]
    for p in range(2):
        for c in range(n):
            x = 0
            for i, q in enumerate(moduli[:L]):
                Qi = Q // q
                x += int(coef_in[p, i, c]) * Qi * pow(Qi, -1, q)
            x %= Q
            # SEAL rounds as floor((x + floor(qL/2)) / qL) on the representative in [0, Q)
            y = (x + qL // 2) // qL
            for i, q in enumerate(moduli[:L - 1]):
                assert int(coef_out[p, i, c]) == y % q


@pytest.mark.parametrize("scheme", [BFV, CKKS])
def test_switch_key_linearity_and_zero_key(scheme):
    logn = 5
    n, moduli = params(logn, [40, 40, 40, 50])
    K, L = len(moduli), 3
    rng = np.random.default_rng(3)
    key = ck.random_key(rng, moduli, logn)
    base = ck.uniform_residues(rng, (3, L, n), moduli[:L])
    zero_key = np.zeros_like(key)
    # zero key: ModDown of zero plus rounding leaves (c0, c1) unchanged
    assert np.array_equal(ck.relinearize(scheme, base, zero_key, moduli, L, logn), base[:2])
    # the update is linear in the key up to the ModDown rounding (at most 1 per
    # coefficient): f(k1) + f(k2) - f(k1 + k2) - (c0, c1) in {-1, 0, 1}
    key2 = ck.random_key(rng, moduli, logn)
    ksum = key.copy()
    for i, q in enumerate(moduli):
        ksum[:, :, i] = (key[:, :, i] + key2[:, :, i]) % np.uint64(q)
    fx = ck.relinearize(scheme, base, key, moduli, L, logn)
    fy = ck.relinearize(scheme, base, key2, moduli, L, logn)
    fs = ck.relinearize(scheme, base, ksum, moduli, L, logn)
    assert not np.array_equal(fx, base[:2])
    conv = (lambda v: ck.ntt_rns(v, moduli[:L], logn, inverse=True)) if scheme == CKKS else (lambda v: v)
    d = conv(fx).astype(object) + conv(fy).astype(object) - conv(fs).astype(object) - conv(base[:2]).astype(object)
    for i, q in enumerate(moduli[:L]):
        r = np.mod(d[:, i], q)
        assert np.all((r <= 1) | (r >= q - 1)), "key switching is not linear up to rounding"


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("scheme,logn,bits,L", [
    (CKKS, 10, [60, 40, 40, 60], 3),
    (CKKS, 11, [60, 40, 40, 40, 60], 2),
    (BFV, 10, [40, 40, 50], 2),
    (BFV, 12, [43, 43, 44, 44, 44], 4),
])
def test_evaluator_matches_reference(scheme, logn, bits, L):
    n, moduli = params(logn, bits)
    plain = 65537 if scheme == BFV else 0
    if scheme == BFV:
        moduli = ck.coeff_modulus_create(n, bits)
    sess = ck.RefSession(scheme, n, moduli, plain)
    rng = np.random.default_rng(logn + L)
    a = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    ct3 = ck.uniform_residues(rng, (3, L, n), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    if scheme == CKKS:
        assert np.array_equal(ck.ckks_multiply(a, b, moduli[:L], logn), sess.multiply(a, b, L))
    assert np.array_equal(ck.relinearize(scheme, ct3, key, moduli, L, logn), sess.relinearize(ct3, key, L))
    if L >= 2:
        assert np.array_equal(ck.rescale(scheme, a, moduli, logn), sess.rescale(a, L))


def test_behz_bases_size_rule():
    # RNSTool::initialize: |B| = |q| unless 32 + bits(t) + bits(q) >= 61 (|q| + 1)
    n = 1 << 12
    q = ck.coeff_modulus_create(n, [60, 60])
    msk, gamma, B = ck.behz_bases(q, 65537, 12)
    assert len(B) == 2
    q3 = ck.coeff_modulus_create(n, [30, 30, 30])
    assert len(ck.behz_bases(q3, 65537, 12)[2]) == 3
    # 61-bit primes = 1 mod 2N, descending: m_sk > gamma > B_0 > B_1
    primes = ck.get_primes(2 * n, 61, 4)
    assert [msk, gamma] + B == primes


def test_bfv_multiply_approximates_exact_scaled_product():
    """BEHZ output = round(t/q * (a (x) b)) mod q up to a small base-conversion error."""
    logn = 3
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [30, 30])
    t = 17
    Q = moduli[0] * moduli[1]
    rng = np.random.default_rng(4)
    a = ck.uniform_residues(rng, (2, 2, n), moduli)
    b = ck.uniform_residues(rng, (2, 2, n), moduli)
    out = ck.bfv_multiply(a, b, moduli, t, logn)

    def compose(x):  # [L][n] -> centred integers in (-Q/2, Q/2]; BEHZ's Montgomery
        r = []       # lift (smMrq) yields the centred representative
        for c in range(n):
            v = 0
            for i, q in enumerate(moduli):
                Qi = Q // q
                v += int(x[i, c]) * Qi * pow(Qi, -1, q)
            v %= Q
            r.append(v - Q if v > Q // 2 else v)
        return r

    A = [compose(a[p]) for p in range(2)]
    Bq = [compose(b[p]) for p in range(2)]

    def pm(x, y):
        r = [0] * n
        for s in range(n):
            for u in range(n):
                if s + u >= n:
                    r[s + u - n] -= x[s] * y[u]
                else:
                    r[s + u] += x[s] * y[u]
        return r

    exact = [pm(A[0], Bq[0]), [u + v for u, v in zip(pm(A[0], Bq[1]), pm(A[1], Bq[0]))], pm(A[1], Bq[1])]
    for d in range(3):
        got = compose(out[d])
        for c in range(n):
            want = (t * exact[d][c]) // Q
            diff = (got[c] - want) % Q
            diff = min(diff, Q - diff)
            assert diff <= 8, (d, c, diff)


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("logn,bits,L,t", [
    (10, [40, 40, 50], 2, 65537),
    (12, [43, 43, 44, 44, 44], 4, 65537),
    (11, [60, 60, 60], 2, 1024),
    (10, [30, 30, 30, 40], 3, 786433),
])
def test_bfv_multiply_matches_reference(logn, bits, L, t):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    sess = ck.RefSession(BFV, n, moduli, t)
    rng = np.random.default_rng(logn + t)
    a = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    assert np.array_equal(ck.bfv_multiply(a, b, moduli[:L], t, logn), sess.multiply(a, b, L))


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("scheme,logn,bits,L,t", [
    (CKKS, 10, [60, 40, 40, 60], 3, 0),
    (BFV, 10, [40, 40, 50], 2, 65537),
    (BFV, 12, [43, 43, 44, 44, 44], 4, 65537),
    (BFV, 11, [30, 30, 40], 2, (1 << 35) - 31),   # t above every q_i: no fast plain lift
])
def test_elementwise_matches_reference(scheme, logn, bits, L, t):
    """add / sub / negate (closed form) and the plaintext operations (oracle) equal the
    reference Evaluator's words on the same inputs."""
    n, moduli = params(logn, bits)
    if t and t > min(moduli[:L]):
        # t must be prime-free relative to q only for batching; any t works for plain ops
        pass
    sess = ck.RefSession(scheme, n, moduli, t)
    rng = np.random.default_rng(logn * 7 + L)
    a = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    assert np.array_equal(ck.mod_add(a, b, moduli), sess.ct_binary(0, a, b, L))
    assert np.array_equal(ck.mod_sub(a, b, moduli), sess.ct_binary(1, a, b, L))
    assert np.array_equal(ck.mod_neg(a, moduli), sess.ct_binary(2, a, None, L))
    if scheme == CKKS:
        p = ck.uniform_residues(rng, (L, n), moduli[:L])
        want = a.copy()
        want[0] = ck.mod_add(a[0], p, moduli)
        assert np.array_equal(want, sess.plain_op(0, a, p, L))
        want[0] = ck.mod_sub(a[0], p, moduli)
        assert np.array_equal(want, sess.plain_op(1, a, p, L))
        assert np.array_equal(ck.multiply_plain_ntt(a, p, moduli, logn), sess.plain_op(2, a, p, L))
    else:
        m = rng.integers(0, t, size=n, dtype=np.uint64)
        want = a.copy()
        want[0] = ck.bfv_add_plain_scaled(a[0], m, moduli, t, logn)
        assert np.array_equal(want, sess.plain_op(0, a, m, L))
        want[0] = ck.bfv_sub_plain_scaled(a[0], m, moduli, t, logn)
        assert np.array_equal(want, sess.plain_op(1, a, m, L))
        assert np.array_equal(ck.bfv_multiply_plain(a, m, moduli, t, logn), sess.plain_op(2, a, m, L))
        # a monomial below ceil(t/2) takes the CPU's monomial path without a lift: same words
        mono = np.zeros(n, dtype=np.uint64)
        mono[5] = (t - 1) // 2
        assert np.array_equal(ck.bfv_multiply_plain(a, mono, moduli, t, logn), sess.plain_op(2, a, mono, L))


@pytest.mark.parametrize("logn,bits,L,t", [(8, [40, 40, 50], 2, 65537), (9, [30, 30, 40], 2, (1 << 35) - 31)])
def test_bfv_multiply_plain_is_negacyclic_product_with_lift(logn, bits, L, t):
    """The oracle's BFV plain multiply equals the schoolbook negacyclic product of each
    limb with the centred plaintext (m - t for m >= ceil(t/2)), computed with Python
    integers -- a dropped lift, a wrong threshold or a cyclic wrap fails it."""
    n, moduli = params(logn, bits)
    rng = np.random.default_rng(11)
    a = ck.uniform_residues(rng, (1, L, n), moduli[:L])
    m = rng.integers(0, t, size=n, dtype=np.uint64)
    m[0] = (t + 1) // 2   # exactly at the threshold: lifted
    m[1] = (t - 1) // 2   # just below: not lifted
    got = ck.bfv_multiply_plain(a, m, moduli, t, logn)
    centred = [int(v) - t if int(v) >= (t + 1) // 2 else int(v) for v in m]
    for i in range(L):
        q = moduli[i]
        x = [int(v) for v in a[0, i]]
        acc = [0] * n
        for u in range(n):
            if x[u] == 0:
                continue
            for v in range(n):
                if centred[v] == 0:
                    continue
                k = u + v
                if k < n:
                    acc[k] += x[u] * centred[v]
                else:
                    acc[k - n] -= x[u] * centred[v]
        assert [c % q for c in acc] == [int(v) for v in got[0, i]]
