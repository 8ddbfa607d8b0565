"""GPU CKKS encoder parity and the CKKS chain end to end.

Encoder: plaintext words equal the oracle's except rare coefficients whose rounding
differs by one (double-precision DWT, the reference's own tolerance case); decode
within 2^-30 relative of the oracle on the same plaintext.
Chain (the CKKS chain workload's parameters: N = 2^15, 60 + 14 x 40 + 60 bits, scale
2^40): encode -> encrypt -> 3 x (multiply -> relinearize -> rescale) -> decrypt ->
decode equals the oracle's decode of the same words within 2^-20 absolute, every
intermediate ciphertext word equals the oracle's on the same inputs, and the result
is the product of the inputs up to CKKS noise.
"""
import math

import numpy as np
import pytest
import torch

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
CKKS = 2


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def to_dev_f64(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def slots(rng, count):
    return rng.uniform(-1, 1, count) + 1j * rng.uniform(-1, 1, count)


def interleave(v):
    return np.stack([np.real(v), np.imag(v)], axis=1).reshape(-1)


@pytest.mark.parametrize("logn,bits,L,scale,count", [
    (10, [60, 40, 40, 60], 3, 2.0 ** 40, 512),
    (13, [60, 40, 40, 40, 60], 4, 2.0 ** 40, 3000),
    (15, [60] + [40] * 14 + [60], 15, 2.0 ** 40, 1 << 14),
    (12, [60, 60, 60, 60], 3, 2.0 ** 80, 2048),
])
def test_encode_decode_parity(logn, bits, L, scale, count):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(logn)
    v = slots(rng, count)
    dv = to_dev_f64(interleave(v))
    plain = empty_dev((L, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    ctx.ckks_encode(dv.data_ptr(), count, scale, L, plain.data_ptr(), eb.data_ptr(), ew, stream())
    got = from_dev(plain)
    want = ck.ckks_encode(v, moduli[:L], scale, logn)
    cg = ck.ntt_rns(got, moduli[:L], logn, inverse=True).astype(object)
    cw = ck.ntt_rns(want, moduli[:L], logn, inverse=True).astype(object)
    q0 = moduli[0]
    diff = np.mod(cg - cw, q0)
    diff = np.minimum(diff, q0 - diff)
    # an integer rounded from a double is fixed only to the double's precision: at
    # scale 2^40 that means +-1 in rare near-half cases; at 2^80 the coefficients'
    # ulp is ~2^27, so allow 2^-45 of the coefficient magnitude (~ scale * |v|)
    tol = max(1.0, scale * 4 * 2.0 ** -45)
    assert float(diff.max()) <= tol
    if tol == 1.0:
        assert np.mean(diff != 0) < 1e-2
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    dw = ctx.ckks_decode_workspace_size(L)
    db = ws(dw)
    dwant = to_dev(want)
    ctx.ckks_decode(dwant.data_ptr(), L, scale, out.data_ptr(), db.data_ptr(), dw, stream())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    dec = o[0::2] + 1j * o[1::2]
    ref = ck.ckks_decode(want, moduli[:L], scale, logn)
    assert np.max(np.abs(dec - ref) / np.maximum(1, np.abs(ref))) < 2.0 ** -30
    assert np.max(np.abs(dec[:count] - v)) < 2.0 ** -20


def test_encode_rejects_too_large():
    n = 1 << 10
    moduli = ck.coeff_modulus_create(n, [30, 30])
    ctx = hip().Context(CKKS, n, moduli)
    dv = to_dev_f64(interleave(np.array([1e12 + 0j])))
    plain = empty_dev((2, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    with pytest.raises(hip().SealHipError):
        ctx.ckks_encode(dv.data_ptr(), 1, 2.0 ** 40, 2, plain.data_ptr(), eb.data_ptr(), ew, stream())


@pytest.mark.parametrize("bad", [np.inf, -np.inf, np.nan])
def test_encode_rejects_non_finite(bad):
    n = 1 << 10
    moduli = ck.coeff_modulus_create(n, [60, 40, 60])
    ctx = hip().Context(CKKS, n, moduli)
    dv = to_dev_f64(interleave(np.array([0.5 + 0j, complex(bad, 0.0), 0.25j])))
    plain = empty_dev((2, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    with pytest.raises(hip().SealHipError, match="not finite"):
        ctx.ckks_encode(dv.data_ptr(), 3, 2.0 ** 40, 2, plain.data_ptr(), eb.data_ptr(), ew, stream())


def test_ckks_chain_end_to_end():
    logn = 15
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [60] + [40] * 14 + [60])
    ctx = hip().Context(CKKS, n, moduli)
    K = len(moduli)
    L = K - 1
    scale = 2.0 ** 40
    sk, pk, rk = empty_dev((K, n)), empty_dev((2, K, n)), empty_dev((K - 1, 2, K, n))
    ctx.keygen_secret(1, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 2, pk.data_ptr(), stream())
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    ctx.keygen_relin(sk.data_ptr(), 3, rk.data_ptr(), kb.data_ptr(), kw, stream())
    rng = np.random.default_rng(4)
    x, y = slots(rng, n // 2), slots(rng, n // 2)
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    cts = []
    for vec, seed in ((x, 10), (y, 11)):
        dv = to_dev_f64(interleave(vec))
        p = empty_dev((L, n))
        ctx.ckks_encode(dv.data_ptr(), n // 2, scale, L, p.data_ptr(), eb.data_ptr(), ew, stream())
        ct = empty_dev((1, 2, L, n))
        cw = ctx.encrypt_workspace_size(L)
        cb = ws(cw)
        ctx.encrypt(pk.data_ptr(), p.data_ptr(), L, seed, ct.data_ptr(), cb.data_ptr(), cw, stream())
        cts.append(ct)
    acc, other = cts
    want = x.copy()
    cur_scale = scale
    key_words = from_dev(rk)
    for step in range(3):
        prod = empty_dev((1, 3, L, n))
        ctx.ckks_multiply(acc.data_ptr(), other.data_ptr(), prod.data_ptr(), 1, L, stream())
        a_w, b_w = from_dev(acc)[0], from_dev(other)[0]
        kw2 = ctx.keyswitch_workspace_size(1, L)
        kb2 = ws(kw2)
        ctx.relinearize(prod.data_ptr(), rk.data_ptr(), 1, L, kb2.data_ptr(), kw2, stream())
        nxt = empty_dev((1, 2, L - 1, n))
        rw = ctx.rescale_workspace_size(1, 2, L)
        rb = ws(rw)
        relin = prod[:, :2].contiguous()
        ctx.rescale_to_next(relin.data_ptr(), nxt.data_ptr(), 1, 2, L, rb.data_ptr(), rw, stream())
        # words equal the oracle's on the same inputs
        o3 = ck.ckks_multiply(a_w, b_w, moduli[:L], logn)
        o2 = ck.relinearize(CKKS, o3, key_words, moduli, L, logn)
        np.testing.assert_array_equal(from_dev(nxt)[0], ck.rescale(CKKS, o2, moduli, logn))
        cur_scale = cur_scale * scale / moduli[L - 1]
        want = want * y
        L -= 1
        acc = nxt
        # the multiplier drops to the new level by rescaling an encryption of y at the new scale
        if step < 2:
            p = empty_dev((L, n))
            dv = to_dev_f64(interleave(y))
            ctx.ckks_encode(dv.data_ptr(), n // 2, scale, L, p.data_ptr(), eb.data_ptr(), ew, stream())
            other = empty_dev((1, 2, L, n))
            cw = ctx.encrypt_workspace_size(L)
            cb = ws(cw)
            ctx.encrypt(pk.data_ptr(), p.data_ptr(), L, 20 + step, other.data_ptr(), cb.data_ptr(), cw, stream())
    dec = empty_dev((L, n))
    ctx.decrypt(acc.data_ptr(), 2, sk.data_ptr(), L, dec.data_ptr(), 0, 0, stream())
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    dw = ctx.ckks_decode_workspace_size(L)
    db = ws(dw)
    ctx.ckks_decode(dec.data_ptr(), L, cur_scale, out.data_ptr(), db.data_ptr(), dw, stream())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    got = o[0::2] + 1j * o[1::2]
    # the GPU decode equals the oracle's decode of the same plaintext words within 2^-20
    ref = ck.ckks_decode(from_dev(dec), moduli[:L], cur_scale, logn)
    assert np.max(np.abs(got - ref)) < 2.0 ** -20
    # and the chain computed x * y^3 (CKKS noise at scale 2^40 after three rescales)
    assert np.max(np.abs(got - want)) < 2.0 ** -10


def test_encode_deferred_check_and_graph_capture():
    """shc_ckks_encode_deferred: the same plaintext words as shc_ckks_encode, the range
    check as a device status word (0; -1 with a zero plaintext where shc_ckks_encode
    raises), and a captured graph of it replays to the eager words."""
    logn, n = 12, 1 << 12
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 60])
    L = 3
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(77)
    v = slots(rng, n // 2)
    dv = to_dev_f64(interleave(v))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    want = empty_dev((L, n))
    ctx.ckks_encode(dv.data_ptr(), n // 2, 2.0 ** 40, L, want.data_ptr(), eb.data_ptr(), ew, stream())
    got = empty_dev((L, n))
    status = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    ctx.ckks_encode_deferred(dv.data_ptr(), n // 2, 2.0 ** 40, L, got.data_ptr(), status.data_ptr(), eb.data_ptr(),
                             ew, stream())
    assert int(status.item()) == 0
    np.testing.assert_array_equal(from_dev(got), from_dev(want))
    # values far beyond the modulus: shc_ckks_encode raises, the deferred form reports
    big = to_dev_f64(interleave(v * 2.0 ** 100))
    with pytest.raises(hip().SealHipError, match="too large"):
        ctx.ckks_encode(big.data_ptr(), n // 2, 2.0 ** 40, L, want.data_ptr(), eb.data_ptr(), ew, stream())
    ctx.ckks_encode_deferred(big.data_ptr(), n // 2, 2.0 ** 40, L, got.data_ptr(), status.data_ptr(), eb.data_ptr(),
                             ew, stream())
    assert int(status.item()) == -1
    assert not from_dev(got).any()
    # captured: replays give the eager words, with eager encodes between the replays
    ctx.ckks_encode(dv.data_ptr(), n // 2, 2.0 ** 40, L, want.data_ptr(), eb.data_ptr(), ew, stream())
    eager = from_dev(want).copy()
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(graph, stream=cap):
        ctx.ckks_encode_deferred(dv.data_ptr(), n // 2, 2.0 ** 40, L, got.data_ptr(), status.data_ptr(),
                                 eb.data_ptr(), ew, torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    for _ in range(3):
        got.zero_()
        status.fill_(7)
        graph.replay()
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        np.testing.assert_array_equal(from_dev(got), eager)
        ctx.ckks_encode_deferred(big.data_ptr(), n // 2, 2.0 ** 40, L, want.data_ptr(), status.data_ptr(),
                                 eb.data_ptr(), ew, stream())
        torch.cuda.synchronize()
        assert int(status.item()) == -1


def test_encode_decode_many_equal_single_calls():
    """A batch of plaintexts in one call (one launch per FFT layer for all of them) gives
    each plaintext the words of its own shc_ckks_encode, and the batched decode the
    values of shc_ckks_decode; per-plaintext status words flag exactly the plaintext
    whose values are too large."""
    logn, n = 13, 1 << 13
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 40, 60])
    L = 4
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(91)
    P, count, stride = 5, 3000, 2 * 4096
    vals = np.zeros((P, stride))
    for p in range(P):
        vals[p, :2 * count] = interleave(slots(rng, count))
    dv = to_dev_f64(vals)
    ew = ctx.ckks_encode_many_workspace_size(P)
    eb = ws(ew)
    plains = empty_dev((P, L, n))
    ctx.ckks_encode_many(dv.data_ptr(), count, stride, P, 2.0 ** 40, L, plains.data_ptr(), None, eb.data_ptr(), ew,
                         stream())
    got = from_dev(plains)
    ew1 = ctx.ckks_encode_workspace_size()
    eb1 = ws(ew1)
    one = empty_dev((L, n))
    for p in range(P):
        ctx.ckks_encode(dv[p].data_ptr(), count, 2.0 ** 40, L, one.data_ptr(), eb1.data_ptr(), ew1, stream())
        np.testing.assert_array_equal(got[p], from_dev(one), err_msg=f"plaintext {p}")
    dw = ctx.ckks_decode_many_workspace_size(L, P)
    db = ws(dw)
    out = torch.empty((P, n), dtype=torch.float64, device="cuda")
    ctx.ckks_decode_many(plains.data_ptr(), L, P, 2.0 ** 40, out.data_ptr(), db.data_ptr(), dw, stream())
    dw1 = ctx.ckks_decode_workspace_size(L)
    db1 = ws(dw1)
    o1 = torch.empty((n,), dtype=torch.float64, device="cuda")
    for p in range(P):
        ctx.ckks_decode(plains[p].data_ptr(), L, 2.0 ** 40, o1.data_ptr(), db1.data_ptr(), dw1, stream())
        torch.cuda.synchronize()
        assert torch.equal(out[p], o1), f"decode of plaintext {p}"
    # one plaintext far beyond the modulus: its status alone is -1 and it is left zero
    vals[3, :2 * count] *= 2.0 ** 100
    dv = to_dev_f64(vals)
    status = torch.full((P,), 7, dtype=torch.int32, device="cuda")
    ctx.ckks_encode_many(dv.data_ptr(), count, stride, P, 2.0 ** 40, L, plains.data_ptr(), status.data_ptr(),
                         eb.data_ptr(), ew, stream())
    assert status.cpu().tolist() == [0, 0, 0, -1, 0]
    g2 = from_dev(plains)
    assert not g2[3].any()
    np.testing.assert_array_equal(g2[0], got[0])
    with pytest.raises(hip().SealHipError, match="too large"):
        ctx.ckks_encode_many(dv.data_ptr(), count, stride, P, 2.0 ** 40, L, plains.data_ptr(), None, eb.data_ptr(),
                             ew, stream())


def round_half_away(x):
    """std::round of a double, as a Python integer (exact for every finite double)."""
    a = abs(x)
    f = math.floor(a)
    r = f + 1 if a - f >= 0.5 else f
    return -int(r) if x < 0 else int(r)


@pytest.mark.parametrize("logn,bits,L,scale", [
    (10, [60, 40, 40, 60], 3, 2.0 ** 40),
    (12, [60, 60, 60, 60], 3, 2.0 ** 80),
    (10, [60, 40, 60], 2, 1.0),
])
def test_encode_decode_polynomial(logn, bits, L, scale):
    """CKKSEncoderCuda::encodePolynomial / decodePolynomial (src/ckks_cuda.cu:455-572, 983-1053):
    coefficient i is std::round(values[i] * scale) (half away from zero; ties included)
    reduced into every prime and transformed to NTT form, word for word; decode returns
    every coefficient over scale."""
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(logn + L)
    count = n - 5  # the last coefficients are zero
    if scale == 1.0:
        v = rng.integers(-2000, 2000, count) / 2.0  # exact halves: every odd numerator is a tie
    else:
        v = rng.uniform(-1, 1, count)
        v[:6] = np.array([0.5, -0.5, 2.5, -2.5, 1e6 + 0.5, -(1e6 + 0.5)]) / scale  # ties after scaling
    coeffs = [round_half_away(float(x) * scale) for x in v] + [0] * (n - count)
    want = np.array([[c % q for c in coeffs] for q in moduli[:L]], dtype=np.uint64)
    want = ck.ntt_rns(want, moduli[:L], logn)
    dv = to_dev_f64(v)
    plain = empty_dev((L, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    ctx.ckks_encode_polynomial(dv.data_ptr(), count, scale, L, plain.data_ptr(), eb.data_ptr(), ew, stream())
    np.testing.assert_array_equal(from_dev(plain), want)
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    dw = ctx.ckks_decode_workspace_size(L)
    db = ws(dw)
    ctx.ckks_decode_polynomial(plain.data_ptr(), L, scale, out.data_ptr(), db.data_ptr(), dw, stream())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    exact = np.array([float(c) for c in coeffs]) / scale
    assert np.max(np.abs(got - exact) / np.maximum(1.0, np.abs(exact))) < 2.0 ** -50
    assert np.max(np.abs(got[:count] - v) - 0.5 / scale) <= 2.0 ** -50


def test_encode_polynomial_rejects():
    n = 1 << 10
    moduli = ck.coeff_modulus_create(n, [30, 30])
    ctx = hip().Context(CKKS, n, moduli)
    plain = empty_dev((2, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    big = to_dev_f64(np.array([1.0, 1e12]))
    with pytest.raises(hip().SealHipError, match="too large"):
        ctx.ckks_encode_polynomial(big.data_ptr(), 2, 2.0 ** 40, 2, plain.data_ptr(), eb.data_ptr(), ew, stream())
    bad = to_dev_f64(np.array([0.5, np.nan]))
    with pytest.raises(hip().SealHipError, match="not finite"):
        ctx.ckks_encode_polynomial(bad.data_ptr(), 2, 2.0 ** 20, 2, plain.data_ptr(), eb.data_ptr(), ew, stream())
    with pytest.raises(hip().SealHipError, match="too large"):
        ctx.ckks_encode_polynomial(big.data_ptr(), n + 1, 1.0, 2, plain.data_ptr(), eb.data_ptr(), ew, stream())
