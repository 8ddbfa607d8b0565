"""GPU parity of the coefficient-wise evaluator operations.

add / sub / negate and the plaintext operations (CKKS add/sub/multiply by an NTT-form
plaintext; BFV add/sub of Delta m and the lifted plain multiply; sums of NTT-form plain
products, one or several outputs at once) equal the oracle word for word on batches of
ciphertexts, in place and out of place, including a plain modulus larger than every
data prime (no fast plain lift).
"""
import numpy as np
import pytest

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


@pytest.mark.parametrize("scheme,logn,bits,L,count", [
    (CKKS, 12, [60, 40, 40, 60], 3, 3),
    (BFV, 13, [43, 43, 44, 44, 44], 4, 2),
    (CKKS, 16, [60] * 5, 4, 2),
    (CKKS, 3, [30, 30, 30], 3, 3),    # N = 8: limbs shorter than a wave's words
])
def test_add_sub_negate(scheme, logn, bits, L, count):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(logn)
    a = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    da, db = to_dev(a), to_dev(b)
    out = empty_dev((count, 2, L, n))
    polys = count * 2
    ctx.add(da.data_ptr(), db.data_ptr(), out.data_ptr(), polys, L, stream())
    assert np.array_equal(from_dev(out), ck.mod_add(a, b, moduli))
    ctx.sub(da.data_ptr(), db.data_ptr(), out.data_ptr(), polys, L, stream())
    assert np.array_equal(from_dev(out), ck.mod_sub(a, b, moduli))
    ctx.negate(da.data_ptr(), out.data_ptr(), polys, L, stream())
    assert np.array_equal(from_dev(out), ck.mod_neg(a, moduli))
    # in place: a += b
    ctx.add(da.data_ptr(), db.data_ptr(), da.data_ptr(), polys, L, stream())
    assert np.array_equal(from_dev(da), ck.mod_add(a, b, moduli))


def test_elementwise_at_odd_word_offsets():
    """Operands at an odd word offset (8-byte aligned only) take the one-word-per-thread
    kernels and give the words of the 16-byte-aligned ones."""
    logn, L, polys = 11, 2, 3
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [50, 50, 60])
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(4)
    a = ck.uniform_residues(rng, (polys, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (polys, L, n), moduli[:L])
    plain = ck.uniform_residues(rng, (L, n), moduli[:L])
    words = polys * L * n
    buf = empty_dev((4 * words + 3 * L * n + 8,))
    da, db, dout, dpl = buf[1:1 + words], buf[3 + words:3 + 2 * words], buf[5 + 2 * words:5 + 3 * words], \
        buf[7 + 3 * words:7 + 3 * words + L * n]
    da.copy_(to_dev(a).view(-1))
    db.copy_(to_dev(b).view(-1))
    dpl.copy_(to_dev(plain).view(-1))
    for op, want in ((ctx.add, ck.mod_add(a, b, moduli)), (ctx.sub, ck.mod_sub(a, b, moduli))):
        op(da.data_ptr(), db.data_ptr(), dout.data_ptr(), polys, L, stream())
        np.testing.assert_array_equal(from_dev(dout).reshape(polys, L, n), want)
    ctx.negate(da.data_ptr(), dout.data_ptr(), polys, L, stream())
    np.testing.assert_array_equal(from_dev(dout).reshape(polys, L, n), ck.mod_neg(a, moduli))
    ctx.multiply_plain_ntt(da.data_ptr(), dpl.data_ptr(), dout.data_ptr(), polys, L, stream())
    np.testing.assert_array_equal(from_dev(dout).reshape(polys, L, n), ck.multiply_plain_ntt(a, plain, moduli, logn))


@pytest.mark.parametrize("logn,bits,L,count", [(12, [60, 40, 40, 60], 3, 3), (15, [60] + [40] * 6 + [60], 7, 2)])
def test_ckks_plain_ops(logn, bits, L, count):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(logn + 1)
    a = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    p = ck.uniform_residues(rng, (L, n), moduli[:L])
    da, dp = to_dev(a), to_dev(p)
    out = empty_dev((count, 2, L, n))
    for subtract in (False, True):
        ctx.add_plain(da.data_ptr(), dp.data_ptr(), out.data_ptr(), count, 2, L, subtract, stream())
        want = a.copy()
        for c in range(count):
            want[c, 0] = (ck.mod_sub if subtract else ck.mod_add)(a[c, 0], p, moduli)
        assert np.array_equal(from_dev(out), want)
    ctx.multiply_plain_ntt(da.data_ptr(), dp.data_ptr(), out.data_ptr(), count * 2, L, stream())
    want = ck.multiply_plain_ntt(a.reshape(count * 2, L, n), p, moduli, logn).reshape(count, 2, L, n)
    assert np.array_equal(from_dev(out), want)


@pytest.mark.parametrize("logn,bits,L,t,count", [
    (13, [43, 43, 44, 44, 44], 4, 65537, 2),
    (12, [30, 30, 40], 2, (1 << 35) - 31, 3),
    (16, [60] * 5, 4, 65537, 1),
])
def test_bfv_plain_ops(logn, bits, L, t, count):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(BFV, n, moduli, t)
    rng = np.random.default_rng(logn + 2)
    a = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    m = rng.integers(0, t, size=n, dtype=np.uint64)
    m[0], m[1] = (t + 1) // 2, (t - 1) // 2
    da, dm = to_dev(a), to_dev(m)
    out = empty_dev((count, 2, L, n))
    for subtract in (False, True):
        ctx.bfv_add_plain(da.data_ptr(), dm.data_ptr(), out.data_ptr(), count, 2, L, subtract, stream())
        want = a.copy()
        f = ck.bfv_sub_plain_scaled if subtract else ck.bfv_add_plain_scaled
        for c in range(count):
            want[c, 0] = f(a[c, 0], m, moduli, t, logn)
        assert np.array_equal(from_dev(out), want)
    wb = ctx.bfv_multiply_plain_workspace_size(L)
    w = ws(wb)
    ctx.bfv_multiply_plain(da.data_ptr(), dm.data_ptr(), out.data_ptr(), count * 2, L, w.data_ptr(), wb, stream())
    want = ck.bfv_multiply_plain(a.reshape(count * 2, L, n), m, moduli, t, logn).reshape(count, 2, L, n)
    assert np.array_equal(from_dev(out), want)
    # in place
    ctx.bfv_multiply_plain(da.data_ptr(), dm.data_ptr(), da.data_ptr(), count * 2, L, w.data_ptr(), wb, stream())
    assert np.array_equal(from_dev(da), want)


def test_plain_ops_reject_bad_arguments():
    n = 1 << 10
    moduli = ck.coeff_modulus_create(n, [30, 30, 30])
    ctx = hip().Context(CKKS, n, moduli)
    d = empty_dev((2, 2, n))
    with pytest.raises(hip().SealHipError):
        ctx.add(d.data_ptr(), d.data_ptr(), d.data_ptr(), 2, 4, stream())   # more limbs than primes
    with pytest.raises(hip().SealHipError):
        ctx.bfv_add_plain(d.data_ptr(), d.data_ptr(), d.data_ptr(), 1, 2, 2, False, stream())   # CKKS context


@pytest.mark.parametrize("scheme,terms,outputs,polys,logn", [
    (CKKS, 70, 11, 4, 10), (BFV, 5, 2, 4, 10),
    (CKKS, 3, 3, 6, 10),    # a partial group of polynomials (4 + 2)
    (CKKS, 65, 1, 9, 11),   # 4 + 4 + 1, across the fold
    (BFV, 2, 5, 2, 10),     # one polynomial per thread
    (CKKS, 4, 2, 5, 4),     # N = 16: rows shorter than a wave
])
def test_multiply_plain_accumulate_many(scheme, terms, outputs, polys, logn):
    """sum_j plains[o][j] * cts[j] for several outputs o at once equals the oracle's plain
    products and sums (70 terms cross the 64-term fold; 11 outputs take an 8-wide and a
    partial 4-wide pass; polynomials in groups of four with partial groups), and each
    output equals the single-output call."""
    L = 3
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [60, 50, 40, 60])
    ctx = hip().Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(terms + outputs)
    cts = ck.uniform_residues(rng, (terms, polys, L, n), moduli[:L])
    plains = ck.uniform_residues(rng, (outputs, terms, L, n), moduli[:L])
    dc, dp = to_dev(cts), to_dev(plains)
    out = empty_dev((outputs, polys, L, n))
    ctx.multiply_plain_accumulate_many(dc.data_ptr(), dp.data_ptr(), terms, outputs, polys, L, out.data_ptr(), stream())
    got = from_dev(out)
    one = empty_dev((polys, L, n))
    for o in range(outputs):
        want = None
        for j in range(terms):
            t = ck.multiply_plain_ntt(cts[j], plains[o, j], moduli[:L], logn)
            want = t if want is None else ck.mod_add(want, t, moduli)
        np.testing.assert_array_equal(got[o], want)
        ctx.multiply_plain_accumulate(dc.data_ptr(), dp[o].data_ptr(), terms, polys, L, one.data_ptr(), stream())
        np.testing.assert_array_equal(from_dev(one), want)
