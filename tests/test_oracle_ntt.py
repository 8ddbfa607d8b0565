"""Pins the oracle's arithmetic, prime selection, NTT and Galois maps.

Known answers come from the reference's own tests (test/utils/ntt.cpp,
test/utils/galois.cpp, test/modulus.cpp); the rest is checked against what the
mathematics fixes (a brute-force negacyclic DFT, the convolution theorem, round
trips) and, where it was built, against the reference CPU library itself.
"""
import ctypes

import numpy as np
import pytest

import checkers as ck
from checkers import ptr

Q_KAT = 0xFFFFFFFFFFC0001


def tables(logn, q):
    n = 1 << logn
    arrs = [np.zeros(n, dtype=np.uint64) for _ in range(4)]
    inv_n = np.zeros(2, dtype=np.uint64)
    assert ck.oracle().or_ntt_tables_export(logn, q, *[ptr(a) for a in arrs], ptr(inv_n)) == 0
    return arrs + [inv_n]


def test_primitive_root_kat():
    # test/utils/ntt.cpp:52-73 (NTTPrimitiveRootsTest)
    rp, rq, irp, irq, _ = tables(1, Q_KAT)
    assert int(rp[0]) == 1 and int(rp[1]) == 288794978602139552
    assert int(irp[1]) == int(ck.oracle().or_invmod(288794978602139552, Q_KAT))
    rp, *_ = tables(2, Q_KAT)
    assert [int(x) for x in rp] == [1, 288794978602139552, 178930308976060547, 748001537669050592]


def test_ntt_kat():
    # test/utils/ntt.cpp:75-99 (NegacyclicNTTTest)
    assert list(ck.ntt([0, 0], Q_KAT, 1)) == [0, 0]
    assert list(ck.ntt([1, 0], Q_KAT, 1)) == [1, 1]
    assert list(ck.ntt([1, 1], Q_KAT, 1)) == [288794978602139553, 864126526004445282]


def test_shoup_and_barrett_closed_forms():
    o = ck.oracle()
    rng = np.random.default_rng(5)
    for q in [Q_KAT, 0x7E00001, 17, (1 << 60) - 93]:
        for w in rng.integers(0, q, size=20, dtype=np.uint64):
            assert int(o.or_shoup(int(w), q)) == (int(w) << 64) // q
        ratio = (ctypes.c_uint64 * 3)()
        o.or_barrett_ratio(ctypes.c_uint64(q), ratio)
        full = (1 << 128) // q
        assert ratio[0] == full & (2**64 - 1) and ratio[1] == full >> 64
        assert ratio[2] == (1 << 128) % q


def test_mulmod_inv_pow_brute():
    o = ck.oracle()
    rng = np.random.default_rng(1)
    for q in [Q_KAT, 65537, (1 << 61) - 1]:
        for _ in range(50):
            a, b = (int(x) for x in rng.integers(0, q, size=2, dtype=np.uint64))
            assert int(o.or_mulmod(a, b, q)) == a * b % q
            e = int(rng.integers(0, 1 << 62))
            assert int(o.or_powmod(a, e, q)) == pow(a, e, q)
            if a:
                assert int(o.or_invmod(a, q)) == pow(a, -1, q)
    assert int(o.or_invmod(6, 9)) == 0  # not invertible


def test_primes_kat_and_closed_form():
    # test/utils/numth.cpp (GetPrimes): primes = 1 mod factor, descending, in the bit range
    ps = ck.get_primes(2 * 8192, 60, 4)
    assert ps == sorted(ps, reverse=True)
    for p in ps:
        assert p % (2 * 8192) == 1 and p.bit_length() == 60
        assert all(pow(a, p - 1, p) == 1 for a in (2, 3, 5, 7))
    # every skipped candidate between the found primes is composite
    cands = range(ps[0], ps[-1] - 1, -2 * 8192)
    found = [c for c in cands if ck.oracle().or_is_prime(c)]
    assert found == ps
    # test/modulus.cpp CoeffModTest.CustomTest: Create(2, {3}) == {5}; Create(2,{3,4}) == {5,13}
    assert ck.coeff_modulus_create(2, [3]) == [5]
    assert ck.coeff_modulus_create(2, [3, 4]) == [5, 13]
    assert ck.coeff_modulus_create(2, [3, 5, 4, 5]) == [5, 17, 13, 29]
    # test/modulus.cpp:249-262: bit sizes kept, all = 1 mod 64, all distinct
    cm = ck.coeff_modulus_create(32, [30, 40, 30, 30, 40])
    assert [c.bit_length() for c in cm] == [30, 40, 30, 30, 40]
    assert all(c % 64 == 1 and ck.oracle().or_is_prime(c) for c in cm) and len(set(cm)) == 5
    # equal bit sizes are handed out in ascending order of appearance
    assert cm[0] < cm[2] < cm[3] and cm[1] < cm[4]
    # BFVDefault(8192) (utils/globals.cpp:39) moduli are NTT primes for N = 8192
    for q in [0x7FFFFFD8001, 0x7FFFFFC8001, 0xFFFFFFFC001, 0xFFFFFF6C001, 0xFFFFFEBC001]:
        assert ck.oracle().or_is_prime(q) and q % (2 * 8192) == 1


def brute_negacyclic_dft(a, q, logn):
    """Evaluate a(X) at psi^(2 bitrev(k)+1): the bit-reversed negacyclic NTT."""
    n = 1 << logn
    r = ctypes.c_uint64()
    assert ck.oracle().or_minimal_primitive_root(2 * n, q, ctypes.byref(r)) == 0
    psi = r.value
    out = []
    for k in range(n):
        e = 2 * int(ck.oracle().or_reverse_bits(k, logn)) + 1
        x = pow(psi, e, q)
        acc = 0
        for c in reversed([int(v) for v in a]):
            acc = (acc * x + c) % q
        out.append(acc)
    return out


@pytest.mark.parametrize("logn,bits", [(3, 30), (5, 60), (6, 50), (8, 60)])
def test_ntt_matches_brute_force_dft(logn, bits):
    n = 1 << logn
    q = ck.get_primes(2 * n, bits, 1)[0]
    rng = np.random.default_rng(logn)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    got = [int(x) for x in ck.ntt(a, q, logn)]
    assert got == brute_negacyclic_dft(a, q, logn)
    # the root is the minimal primitive 2n-th root: no smaller element has order 2n
    r = ctypes.c_uint64()
    ck.oracle().or_minimal_primitive_root(2 * n, q, ctypes.byref(r))
    for c in range(2, min(int(r.value), 5000)):
        assert pow(c, n, q) != q - 1


def negacyclic_schoolbook(a, b, q):
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            v = int(a[i]) * int(b[j])
            if k >= n:
                out[k - n] = (out[k - n] - v) % q
            else:
                out[k] = (out[k] + v) % q
    return out


def test_convolution_theorem_and_round_trip():
    logn = 6
    n = 1 << logn
    q = ck.get_primes(2 * n, 60, 1)[0]
    rng = np.random.default_rng(2)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    b = rng.integers(0, q, size=n, dtype=np.uint64)
    fa, fb = ck.ntt(a, q, logn), ck.ntt(b, q, logn)
    prod = np.array([int(x) * int(y) % q for x, y in zip(fa, fb)], dtype=np.uint64)
    back = ck.ntt(prod, q, logn, inverse=True)
    assert [int(x) for x in back] == negacyclic_schoolbook(a, b, q)
    assert np.array_equal(ck.ntt(fa, q, logn, inverse=True), a)


def test_galois_kats():
    o = ck.oracle()
    # test/utils/galois.cpp:26-60
    assert [o.or_galois_elt_from_step(3, s) for s in (0, 1, -3, 2, -2, 3, -1)] == [15, 3, 3, 9, 9, 11, 11]
    inp = np.arange(8, dtype=np.uint64)
    out = np.zeros(8, dtype=np.uint64)
    o.or_apply_galois(3, 3, 17, ptr(inp), ptr(out))
    assert list(out) == [0, 14, 6, 1, 13, 7, 2, 12]
    o.or_apply_galois_ntt(3, 3, ptr(inp), ptr(out))
    assert list(out) == [4, 5, 7, 6, 1, 0, 2, 3]


def test_galois_ntt_commutes_with_ntt():
    # sigma applied in coefficient form then NTT == NTT then the NTT-domain permutation
    logn = 7
    n = 1 << logn
    q = ck.get_primes(2 * n, 50, 1)[0]
    rng = np.random.default_rng(3)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    o = ck.oracle()
    for step in (1, 2, -1, 5, 0):
        elt = o.or_galois_elt_from_step(logn, step)
        s = np.zeros(n, dtype=np.uint64)
        o.or_apply_galois(logn, elt, q, ptr(a), ptr(s))
        lhs = ck.ntt(s, q, logn)
        fa = ck.ntt(a, q, logn)
        rhs = np.zeros(n, dtype=np.uint64)
        o.or_apply_galois_ntt(logn, elt, ptr(fa), ptr(rhs))
        assert np.array_equal(lhs, rhs)


# ------------------------------------------------------------------ against the reference library

ref_only = pytest.mark.skipif(ck.ref() is None, reason="reference CPU library oracle/_ref not built here")


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("n,bits", [
    (8192, [43, 43, 44, 44, 44]),
    (65536, [60] * 25),
    (32768, [60] + [40] * 14 + [60]),
    (65536, [60] + [50] * 23 + [60]),
    (1024, [30, 20, 30, 20]),
])
def test_coeff_modulus_matches_reference(n, bits):
    ours = ck.coeff_modulus_create(n, bits)
    out = np.zeros(len(bits), dtype=np.uint64)
    b = (ctypes.c_int * len(bits))(*bits)
    assert ck.ref().ref_coeff_modulus_create(n, b, len(bits), ptr(out)) == 0
    assert ours == [int(x) for x in out]


@ref_only
@pytest.mark.ref
@pytest.mark.parametrize("logn,bits", [(10, 60), (13, 44), (12, 30), (16, 60), (15, 40)])
def test_ntt_tables_and_transforms_match_reference(logn, bits):
    n = 1 << logn
    q = ck.get_primes(2 * n, bits, 2)[1]
    ours = tables(logn, q)
    theirs = [np.zeros(n, dtype=np.uint64) for _ in range(4)] + [np.zeros(2, dtype=np.uint64)]
    assert ck.ref().ref_ntt_tables(logn, q, *[ptr(a) for a in theirs]) == 0
    for a, b in zip(ours, theirs):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(logn)
    a = rng.integers(0, q, size=3 * n, dtype=np.uint64)
    for inverse in (0, 1):
        r = a.copy()
        ck.ref().ref_ntt(logn, q, ptr(r), 3, inverse)
        assert np.array_equal(ck.ntt(a, q, logn, inverse=bool(inverse)), r)


@ref_only
@pytest.mark.ref
def test_galois_matches_reference():
    logn = 10
    n = 1 << logn
    q = ck.get_primes(2 * n, 50, 1)[0]
    rng = np.random.default_rng(9)
    a = rng.integers(0, q, size=n, dtype=np.uint64)
    o = ck.oracle()
    for step in (0, 1, -1, 7, -100, 255):
        elt = o.or_galois_elt_from_step(logn, step)
        assert elt == ck.ref().ref_galois_elt_from_step(logn, step)
        for ntt in (0, 1):
            r = np.zeros(n, dtype=np.uint64)
            s = np.zeros(n, dtype=np.uint64)
            ck.ref().ref_apply_galois(logn, elt, q, ptr(a), ptr(r), ntt)
            if ntt:
                o.or_apply_galois_ntt(logn, elt, ptr(a), ptr(s))
            else:
                o.or_apply_galois(logn, elt, q, ptr(a), ptr(s))
            assert np.array_equal(r, s)
