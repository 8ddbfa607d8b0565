"""GPU parity and end-to-end correctness of sampling, key generation, encryption,
decryption and the BatchEncoder.

Word-level parity: decryption and batch encoding are compared word for word with the
oracle on arbitrary inputs; encryption is compared on the same sampled randomness.
End to end (the reference's decrypt-KAT style, test/evaluator_cuda.cu): the BFV plain
path of the CPU-runnable configuration (N = 8192, 43,43,44,44,44-bit primes,
t = 65537): encode -> encrypt -> multiply -> relinearize -> decrypt -> decode gives
the slot-wise product mod t.
"""
import numpy as np
import pytest

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2
T = 65537


def bfv_ctx(logn=13, bits=(43, 43, 44, 44, 44), t=T):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, list(bits))
    return hip().Context(BFV, n, moduli, t), moduli


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def centred(x, q):
    x = x.astype(object)
    return np.where(x > q // 2, x - q, x)


def test_samplers_distributions():
    ctx, moduli = bfv_ctx()
    n, K = ctx.n, len(moduli)
    d = empty_dev((K, n))
    ctx.sample_poly(0, 11, K, d.data_ptr(), stream())
    t = from_dev(d)
    vals = centred(t[0], moduli[0])
    assert set(np.unique(vals)) <= {-1, 0, 1}
    for i, q in enumerate(moduli):
        assert np.array_equal(centred(t[i], q), vals)
    assert abs(np.mean(vals == 0) - 1 / 3) < 0.03
    ctx.sample_poly(1, 12, K, d.data_ptr(), stream())
    e = centred(from_dev(d)[0], moduli[0]).astype(np.int64)
    assert np.abs(e).max() <= 21 and abs(e.var() - 10.5) < 1.0
    ctx.sample_poly(2, 13, K, d.data_ptr(), stream())
    u = from_dev(d)
    for i, q in enumerate(moduli):
        assert u[i].max() < q and u[i].min() >= 0
    # a seed reproduces its words
    d2 = empty_dev((K, n))
    ctx.sample_poly(2, 13, K, d2.data_ptr(), stream())
    assert np.array_equal(from_dev(d2), u)


@pytest.mark.parametrize("logn,bits,L", [(13, (43, 43, 44, 44, 44), 4), (13, (43, 43, 44, 44, 44), 2),
                                         (10, (40, 40, 50), 2)])
def test_decrypt_and_batch_words_match_oracle(logn, bits, L):
    ctx, moduli = bfv_ctx(logn, bits)
    n = ctx.n
    rng = np.random.default_rng(logn + L)
    sk = ck.ntt_rns(ck.rns_of_small(rng.integers(-1, 2, size=n), moduli), moduli, logn)
    ct = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    out = empty_dev((n,))
    w = ctx.decrypt_workspace_size(L)
    wb = ws(w)
    dct, dsk = to_dev(ct), to_dev(sk)  # keep the tensors alive across the call
    ctx.decrypt(dct.data_ptr(), 2, dsk.data_ptr(), L, out.data_ptr(), wb.data_ptr(), w, stream())
    np.testing.assert_array_equal(from_dev(out), ck.bfv_decrypt(ct, sk[:L], moduli, T, logn))
    values = rng.integers(0, T, size=n, dtype=np.uint64)
    plain = empty_dev((n,))
    dv = to_dev(values)
    ctx.batch_encode(dv.data_ptr(), n, plain.data_ptr(), stream())
    np.testing.assert_array_equal(from_dev(plain), ck.batch_encode(values, T, logn))
    back, scratch = empty_dev((n,)), empty_dev((n,))
    ctx.batch_decode(plain.data_ptr(), back.data_ptr(), scratch.data_ptr(), stream())
    np.testing.assert_array_equal(from_dev(back), values)


def test_ckks_decrypt_words_match_oracle():
    logn = 12
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 60])
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(5)
    sk = ck.ntt_rns(ck.rns_of_small(rng.integers(-1, 2, size=n), moduli), moduli, logn)
    for size in (2, 3):
        ct = ck.uniform_residues(rng, (size, 3, n), moduli[:3])
        out = empty_dev((3, n))
        dct, dsk = to_dev(ct), to_dev(sk)
        ctx.decrypt(dct.data_ptr(), size, dsk.data_ptr(), 3, out.data_ptr(), 0, 0, stream())
        np.testing.assert_array_equal(from_dev(out), ck.decrypt_phase(CKKS, ct, sk[:3], moduli, logn))


@pytest.mark.parametrize("scheme", [BFV, CKKS])
def test_encrypt_words_match_oracle_on_same_randomness(scheme):
    logn = 12
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [50, 40, 40, 50] if scheme == CKKS else [43, 43, 44, 44])
    ctx = hip().Context(scheme, n, moduli, T if scheme == BFV else 0)
    K = len(moduli)
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(1, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 2, pk.data_ptr(), stream())
    for L in (K - 1, 1):
        L1 = L + 1
        u, e0, e1 = empty_dev((L1, n)), empty_dev((L1, n)), empty_dev((L1, n))
        ctx.sample_poly(0, 21, L1, u.data_ptr(), stream())
        ctx.sample_poly(1, 22, L1, e0.data_ptr(), stream())
        ctx.sample_poly(1, 23, L1, e1.data_ptr(), stream())
        out = empty_dev((2, L, n))
        w = ctx.encrypt_workspace_size(L)
        wb = ws(w)
        ctx.encrypt_zero_asymmetric(pk.data_ptr(), u.data_ptr(), e0.data_ptr(), e1.data_ptr(), L, out.data_ptr(),
                                    wb.data_ptr(), w, stream())
        z = ck.encrypt_zero_asymmetric(scheme, from_dev(pk), from_dev(u), from_dev(e0), from_dev(e1),
                                       moduli[:L1], logn)
        want = ck.rescale(scheme, z, moduli[:L1], logn)
        np.testing.assert_array_equal(from_dev(out), want)


def test_keys_have_the_kswitch_structure():
    """Each relinearization key digit decrypts to P * s^2 (limb j) minus a CBD noise."""
    logn = 11
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [40, 40, 40, 50])
    ctx = hip().Context(CKKS, n, moduli)
    K = len(moduli)
    sk, rk = empty_dev((K, n)), empty_dev((K - 1, 2, K, n))
    ctx.keygen_secret(3, sk.data_ptr(), stream())
    w = ctx.keygen_workspace_size()
    wb = ws(w)
    ctx.keygen_relin(sk.data_ptr(), 4, rk.data_ptr(), wb.data_ptr(), w, stream())
    s = from_dev(sk)
    s2 = (s.astype(object) * s.astype(object) % np.array(moduli, dtype=object)[:, None]).astype(np.uint64)
    key = from_dev(rk)
    P = moduli[-1]
    for j in range(K - 1):
        phase = ck.ntt_rns(ck.decrypt_phase(CKKS, key[j], s, moduli, logn), moduli, logn, inverse=True)
        s2c = ck.ntt_rns(s2, moduli, logn, inverse=True)
        for i, q in enumerate(moduli):
            noise = phase[i].astype(object) - (P * s2c[i].astype(object) if i == j else 0)
            noise = centred(np.mod(noise, q), q)
            assert np.abs(noise.astype(np.int64)).max() <= 21


def test_bfv_pipeline_end_to_end():
    """encode -> encrypt -> multiply -> relinearize -> decrypt -> decode (config-1 parameters)."""
    ctx, moduli = bfv_ctx()
    n, K = ctx.n, len(moduli)
    L = K - 1
    sk, pk, rk = empty_dev((K, n)), empty_dev((2, K, n)), empty_dev((K - 1, 2, K, n))
    ctx.keygen_secret(101, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 102, pk.data_ptr(), stream())
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    ctx.keygen_relin(sk.data_ptr(), 103, rk.data_ptr(), kb.data_ptr(), kw, stream())
    rng = np.random.default_rng(1)
    x = rng.integers(0, T, size=n, dtype=np.uint64)
    y = rng.integers(0, T, size=n, dtype=np.uint64)
    px, py = empty_dev((n,)), empty_dev((n,))
    dx, dy = to_dev(x), to_dev(y)
    ctx.batch_encode(dx.data_ptr(), n, px.data_ptr(), stream())
    ctx.batch_encode(dy.data_ptr(), n, py.data_ptr(), stream())
    cts = empty_dev((2, 2, L, n))
    ew = ctx.encrypt_workspace_size(L)
    eb = ws(ew)
    ctx.encrypt(pk.data_ptr(), px.data_ptr(), L, 7, cts[0].data_ptr(), eb.data_ptr(), ew, stream())
    ctx.encrypt(pk.data_ptr(), py.data_ptr(), L, 8, cts[1].data_ptr(), eb.data_ptr(), ew, stream())
    prod = empty_dev((1, 3, L, n))
    mw = ctx.multiply_workspace_size(1, L)
    mb = ws(mw)
    ctx.multiply(cts[0].data_ptr(), cts[1].data_ptr(), prod.data_ptr(), 1, L, mb.data_ptr(), mw, stream())
    # decrypting the size-3 product already works (phase with s^2)
    dw = ctx.decrypt_workspace_size(L)
    db = ws(dw)
    out, vals, scratch = empty_dev((n,)), empty_dev((n,)), empty_dev((n,))
    ctx.decrypt(prod.data_ptr(), 3, sk.data_ptr(), L, out.data_ptr(), db.data_ptr(), dw, stream())
    ctx.batch_decode(out.data_ptr(), vals.data_ptr(), scratch.data_ptr(), stream())
    want = (x.astype(object) * y.astype(object) % T).astype(np.uint64)
    np.testing.assert_array_equal(from_dev(vals), want)
    kw2 = ctx.keyswitch_workspace_size(1, L)
    kb2 = ws(kw2)
    ctx.relinearize(prod.data_ptr(), rk.data_ptr(), 1, L, kb2.data_ptr(), kw2, stream())
    ctx.decrypt(prod.data_ptr(), 2, sk.data_ptr(), L, out.data_ptr(), db.data_ptr(), dw, stream())
    ctx.batch_decode(out.data_ptr(), vals.data_ptr(), scratch.data_ptr(), stream())
    np.testing.assert_array_equal(from_dev(vals), want)
    # the GPU ciphertext words after multiply+relinearize equal the oracle's on the same inputs
    c0 = from_dev(cts)
    o3 = ck.bfv_multiply(c0[0], c0[1], moduli[:L], T, 13)
    o2 = ck.relinearize(BFV, o3, from_dev(rk), moduli, L, 13)
    np.testing.assert_array_equal(from_dev(prod)[0, :2], o2)


def test_bfv_rotate_rows_end_to_end():
    ctx, moduli = bfv_ctx(12, (43, 43, 44, 44))
    n, K = ctx.n, len(moduli)
    L = K - 1
    sk, pk, gk = empty_dev((K, n)), empty_dev((2, K, n)), empty_dev((K - 1, 2, K, n))
    ctx.keygen_secret(201, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 202, pk.data_ptr(), stream())
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    elt = ctx.galois_elt_from_step(1)
    ctx.keygen_galois(sk.data_ptr(), elt, 203, gk.data_ptr(), kb.data_ptr(), kw, stream())
    rng = np.random.default_rng(2)
    x = rng.integers(0, T, size=n, dtype=np.uint64)
    px = empty_dev((n,))
    dx = to_dev(x)
    ctx.batch_encode(dx.data_ptr(), n, px.data_ptr(), stream())
    ct = empty_dev((1, 2, L, n))
    ew = ctx.encrypt_workspace_size(L)
    eb = ws(ew)
    ctx.encrypt(pk.data_ptr(), px.data_ptr(), L, 9, ct.data_ptr(), eb.data_ptr(), ew, stream())
    gw = ctx.galois_workspace_size(1, L)
    gb = ws(gw)
    ctx.apply_galois_keyswitch(ct.data_ptr(), gk.data_ptr(), elt, 1, L, gb.data_ptr(), gw, stream())
    dw = ctx.decrypt_workspace_size(L)
    db = ws(dw)
    out, vals, scratch = empty_dev((n,)), empty_dev((n,)), empty_dev((n,))
    ctx.decrypt(ct.data_ptr(), 2, sk.data_ptr(), L, out.data_ptr(), db.data_ptr(), dw, stream())
    ctx.batch_decode(out.data_ptr(), vals.data_ptr(), scratch.data_ptr(), stream())
    got = from_dev(vals)
    row = n // 2
    # rotate_rows(1): each row of the 2 x N/2 matrix shifts left by one
    want = np.concatenate([np.roll(x[:row], -1), np.roll(x[row:], -1)])
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("keyed", [False, True])
def test_sampler_words_match_oracle(keyed):
    """Every sampler kind gives the oracle's words: the splitmix64 stream of an unkeyed
    context and the ChaCha20 stream of a keyed one (also at a degree whose coefficient
    count is not a multiple of a keystream block's eight words)."""
    for logn, bits in ((13, (43, 43, 44, 44, 44)), (2, (30, 30))):
        n = 1 << logn
        moduli = ck.coeff_modulus_create(n, list(bits))
        ctx = hip().Context(BFV, n, moduli, T) if logn > 2 else hip().Context(CKKS, n, moduli)
        key = bytes(range(7, 39)) if keyed else None
        if keyed:
            ctx.set_sampler_key(key)
        K = len(moduli)
        d = empty_dev((K, n))
        for kind, seed in ((0, 11), (1, 12), (2, 13), (2, 1 << 63)):
            ctx.sample_poly(kind, seed, K, d.data_ptr(), stream())
            np.testing.assert_array_equal(from_dev(d), ck.sample_poly(kind, key, seed, logn, moduli),
                                          err_msg=f"kind {kind}, seed {seed}, N {n}")


def test_keyed_context_pipeline():
    """Keys and encryptions drawn from the ChaCha20 stream decrypt as before, and the
    sampler key changes every random word (the secret keys of the same seed differ)."""
    ctx, moduli = bfv_ctx()
    n, K = ctx.n, len(moduli)
    L = K - 1
    sk_plain = empty_dev((K, n))
    ctx.keygen_secret(5, sk_plain.data_ptr(), stream())
    ctx.set_sampler_key(bytes(range(32)))
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(5, sk.data_ptr(), stream())
    assert not np.array_equal(from_dev(sk), from_dev(sk_plain))
    ctx.keygen_public(sk.data_ptr(), 6, pk.data_ptr(), stream())
    rng = np.random.default_rng(4)
    x = rng.integers(0, T, size=n, dtype=np.uint64)
    px = empty_dev((n,))
    ctx.batch_encode(to_dev(x).data_ptr(), n, px.data_ptr(), stream())
    ct = empty_dev((1, 2, L, n))
    ew = ctx.encrypt_workspace_size(L)
    ctx.encrypt(pk.data_ptr(), px.data_ptr(), L, 7, ct.data_ptr(), ws(ew).data_ptr(), ew, stream())
    dw = ctx.decrypt_workspace_size(L)
    out, vals, scratch = empty_dev((n,)), empty_dev((n,)), empty_dev((n,))
    ctx.decrypt(ct.data_ptr(), 2, sk.data_ptr(), L, out.data_ptr(), ws(dw).data_ptr(), dw, stream())
    ctx.batch_decode(out.data_ptr(), vals.data_ptr(), scratch.data_ptr(), stream())
    np.testing.assert_array_equal(from_dev(vals), x)
    ctx.set_sampler_key(None)   # back to the splitmix64 stream
    ctx.keygen_secret(5, sk.data_ptr(), stream())
    np.testing.assert_array_equal(from_dev(sk), from_dev(sk_plain))


def test_key_generation_in_a_captured_graph():
    """Relinearization and Galois key generation launch kernels only (no allocation, no
    wait for the host), so they can be captured: replays of the captured graph give the
    eager keys, with eager key generation between the replays.  Keyed (ChaCha20) sampler."""
    import torch
    ctx, moduli = bfv_ctx()
    ctx.set_sampler_key(bytes(range(32)))
    n, K = ctx.n, len(moduli)
    sk = empty_dev((K, n))
    ctx.keygen_secret(5, sk.data_ptr(), stream())
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    rk, gk = empty_dev((K - 1, 2, K, n)), empty_dev((K - 1, 2, K, n))
    elt = ctx.galois_elt_from_step(3)

    def gen(st, rk_, gk_):
        ctx.keygen_relin(sk.data_ptr(), 11, rk_.data_ptr(), kb.data_ptr(), kw, st)
        ctx.keygen_galois(sk.data_ptr(), elt, 12, gk_.data_ptr(), kb.data_ptr(), kw, st)

    gen(stream(), rk, gk)
    want_r, want_g = from_dev(rk).copy(), from_dev(gk).copy()
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(graph, stream=cap):
        gen(torch.cuda.current_stream().cuda_stream, rk, gk)
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    rk2, gk2 = empty_dev((K - 1, 2, K, n)), empty_dev((K - 1, 2, K, n))
    for _ in range(3):
        rk.zero_()
        gk.zero_()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(rk), want_r)
        np.testing.assert_array_equal(from_dev(gk), want_g)
        gen(stream(), rk2, gk2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(rk2), want_r)


M64 = (1 << 64) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def sub_seed(seed, k):
    """The library's derivation of an encryption's three streams (u, e0, e1 = k 0, 1, 2)."""
    return mix64((seed ^ mix64((k + 0x5EA15EED) & M64)) & M64)


def expected_encryption(scheme, key, pk, seed, L, moduli, logn, plain=None):
    """Oracle words of a public-key encryption at level L with the streams of `seed`:
    sampled u, e0, e1 at L + 1 limbs (L at the key level), pk u + e, then the
    divide-and-round by the dropped prime, then the plaintext."""
    K = len(moduli)
    L1 = min(L + 1, K)
    u = ck.sample_poly(0, key, sub_seed(seed, 0), logn, moduli[:L1])
    e0 = ck.sample_poly(1, key, sub_seed(seed, 1), logn, moduli[:L1])
    e1 = ck.sample_poly(1, key, sub_seed(seed, 2), logn, moduli[:L1])
    z = ck.encrypt_zero_asymmetric(scheme, pk, u, e0, e1, moduli[:L1], logn)
    want = ck.rescale(scheme, z, moduli[:L1], logn) if L1 != L else z
    if plain is not None:
        if scheme == CKKS:
            want[0] = ck.mod_add(want[0], plain, moduli)
        else:
            want[0] = ck.bfv_add_plain_scaled(want[0], plain, moduli, T, logn)
    return want


@pytest.mark.parametrize("scheme,keyed,logn", [(BFV, False, 12), (CKKS, False, 12), (BFV, True, 12), (CKKS, True, 12),
                                               (CKKS, False, 14), (CKKS, True, 13), (BFV, False, 13)])
def test_encrypt_many_words_match_oracle(scheme, keyed, logn):
    """shc_encrypt_many: ciphertext g equals the oracle's encryption with the streams of
    seed + g (u, e0, e1 sampled by the oracle, pk u + e, divide-and-round, plaintext)
    and shc_encrypt(seed + g), at the top level (special prime dropped), at level 1 (a
    data prime dropped) and at the key level (nothing dropped), with and without
    plaintexts, in one pass and in slices of two.  From N = 2^13 CKKS runs the
    encryption fused into the error term's transform."""
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [50, 40, 40, 50] if scheme == CKKS else [43, 43, 44, 44])
    ctx = hip().Context(scheme, n, moduli, T if scheme == BFV else 0)
    key = bytes(range(3, 35)) if keyed else None
    if keyed:
        ctx.set_sampler_key(key)
    K = len(moduli)
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(1, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 2, pk.data_ptr(), stream())
    pk_h = from_dev(pk)
    rng = np.random.default_rng(scheme * 10 + keyed)
    count, seed = 5, 1000 + 7 * keyed
    for L in (K - 1, 1, K):
        plains = [None]
        if scheme == CKKS:
            plains.append(ck.uniform_residues(rng, (count, L, n), moduli[:L]))
        elif L < K:
            plains.append(rng.integers(0, T, size=(count, n), dtype=np.uint64))
        for plain in plains:
            dplain = to_dev(plain) if plain is not None else None
            pp = dplain.data_ptr() if dplain is not None else 0
            want = np.stack([expected_encryption(scheme, key, pk_h, seed + g, L, moduli, logn,
                                                 None if plain is None else plain[g]) for g in range(count)])
            out = empty_dev((count, 2, L, n))
            w = ctx.encrypt_many_workspace_size(L, count)
            wb = ws(w)
            ctx.encrypt_many(pk.data_ptr(), pp, count, L, seed, out.data_ptr(), wb.data_ptr(), w, stream())
            np.testing.assert_array_equal(from_dev(out), want, err_msg=f"L {L}, plain {plain is not None}")
            # slices of two ciphertexts
            w2 = ctx.encrypt_many_workspace_size(L, 2)
            out2 = empty_dev((count, 2, L, n))
            ctx.encrypt_many(pk.data_ptr(), pp, count, L, seed, out2.data_ptr(), wb.data_ptr(), w2, stream())
            np.testing.assert_array_equal(from_dev(out2), want)
            # the single-ciphertext entry point gives the same words
            ew = ctx.encrypt_workspace_size(L)
            eb = ws(ew)
            one = empty_dev((2, L, n))
            for g in (0, count - 1):
                ctx.encrypt(pk.data_ptr(), 0 if dplain is None else dplain[g].data_ptr(), L, seed + g,
                            one.data_ptr(), eb.data_ptr(), ew, stream())
                np.testing.assert_array_equal(from_dev(one), want[g])


def test_encrypt_many_rejects_and_empty():
    logn = 10
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [40, 40, 50])
    ctx = hip().Context(CKKS, n, moduli)
    K = len(moduli)
    pk, out = empty_dev((2, K, n)), empty_dev((2, 2, 2, n))
    w = ctx.encrypt_many_workspace_size(2, 2)
    wb = ws(w)
    ctx.encrypt_many(pk.data_ptr(), 0, 0, 2, 1, out.data_ptr(), 0, 0, stream())  # nothing to do
    with pytest.raises(hip().SealHipError):
        ctx.encrypt_many(pk.data_ptr(), 0, 2, 2, 1, out.data_ptr(), wb.data_ptr(), w // 2 - 8, stream())
    with pytest.raises(hip().SealHipError):
        ctx.encrypt_many(pk.data_ptr(), 0, 2, K + 1, 1, out.data_ptr(), wb.data_ptr(), w, stream())
    with pytest.raises(hip().SealHipError):
        ctx.decrypt_many(out.data_ptr(), 4, 1, pk.data_ptr(), 2, out.data_ptr(), 0, 0, stream())


@pytest.mark.parametrize("scheme", [BFV, CKKS])
def test_decrypt_many_words_match_oracle(scheme):
    """shc_decrypt_many on size-2 and size-3 ciphertexts with arbitrary words: every
    ciphertext's output is the oracle's decryption (BFV scale-and-round, CKKS phase) and
    shc_decrypt's; BFV also in slices of two ciphertexts."""
    logn = 12
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [50, 40, 40, 50] if scheme == CKKS else [43, 43, 44, 44])
    ctx = hip().Context(scheme, n, moduli, T if scheme == BFV else 0)
    K = len(moduli)
    rng = np.random.default_rng(40 + scheme)
    sk = ck.ntt_rns(ck.rns_of_small(rng.integers(-1, 2, size=n), moduli), moduli, logn)
    dsk = to_dev(sk)
    count = 5
    for L in (K - 1, 2):
        for size in (2, 3):
            cts = ck.uniform_residues(rng, (count, size, L, n), moduli[:L])
            if scheme == BFV:
                want = np.stack([ck.bfv_decrypt(ct, sk[:L], moduli, T, logn) for ct in cts])
                shape = (count, n)
            else:
                want = np.stack([ck.decrypt_phase(CKKS, ct, sk[:L], moduli, logn) for ct in cts])
                shape = (count, L, n)
            dct = to_dev(cts)
            out = empty_dev(shape)
            w = ctx.decrypt_many_workspace_size(L, count)
            wb = ws(w)
            ctx.decrypt_many(dct.data_ptr(), size, count, dsk.data_ptr(), L, out.data_ptr(), wb.data_ptr(), w,
                             stream())
            np.testing.assert_array_equal(from_dev(out), want, err_msg=f"L {L}, size {size}")
            if scheme == BFV:
                w2 = ctx.decrypt_many_workspace_size(L, 2)
                out2 = empty_dev(shape)
                ctx.decrypt_many(dct.data_ptr(), size, count, dsk.data_ptr(), L, out2.data_ptr(), wb.data_ptr(), w2,
                                 stream())
                np.testing.assert_array_equal(from_dev(out2), want)
            dw = ctx.decrypt_workspace_size(L)
            db = ws(dw)
            one = empty_dev(shape[1:])
            ctx.decrypt(dct[count - 1].data_ptr(), size, dsk.data_ptr(), L, one.data_ptr(), db.data_ptr(), dw, stream())
            np.testing.assert_array_equal(from_dev(one), want[count - 1])


def test_batch_encoder_many_matches_single_calls():
    """shc_batch_encode_many / shc_batch_decode_many: plaintext g of a batch (strided
    values, fewer than N) has the words of its own shc_batch_encode and the oracle's, and
    decodes back to its values padded with zeros."""
    ctx, moduli = bfv_ctx()
    n = ctx.n
    rng = np.random.default_rng(71)
    P, count, stride = 5, 5000, 6000
    vals = np.zeros((P, stride), dtype=np.uint64)
    vals[:, :count] = rng.integers(0, T, size=(P, count), dtype=np.uint64)
    dv = to_dev(vals)
    plains = empty_dev((P, n))
    ctx.batch_encode_many(dv.data_ptr(), count, stride, P, plains.data_ptr(), stream())
    got = from_dev(plains)
    one = empty_dev((n,))
    for g in range(P):
        padded = np.zeros(n, dtype=np.uint64)
        padded[:count] = vals[g, :count]
        np.testing.assert_array_equal(got[g], ck.batch_encode(padded, T, logn=13), err_msg=f"plaintext {g}")
        ctx.batch_encode(dv[g].data_ptr(), count, one.data_ptr(), stream())
        np.testing.assert_array_equal(got[g], from_dev(one))
    back, scratch = empty_dev((P, n)), empty_dev((P, n))
    ctx.batch_decode_many(plains.data_ptr(), P, back.data_ptr(), scratch.data_ptr(), stream())
    b = from_dev(back)
    np.testing.assert_array_equal(b[:, :count], vals[:, :count])
    assert not b[:, count:].any()
    ctx.batch_encode_many(dv.data_ptr(), count, stride, 0, plains.data_ptr(), stream())  # nothing to do
    with pytest.raises(hip().SealHipError):
        ctx.batch_encode_many(dv.data_ptr(), count, count - 1, 2, plains.data_ptr(), stream())


@pytest.mark.parametrize("keyed,logn", [(False, 11), (True, 11), (False, 13)])
def test_relin_key_words_match_oracle(keyed, logn):
    """Relinearization and Galois key words equal the oracle's key-switching key built from
    the library's documented streams: digit j encrypts with seed_j = sub_seed(seed, 100 +
    j), a from sub_seed(seed_j, 0) (uniform), e from sub_seed(seed_j, 1) (centred
    binomial), then limb j of c0 takes new_key * (P mod q_j)."""
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [40, 40, 40, 50])
    ctx = hip().Context(CKKS, n, moduli)
    key = bytes(range(40, 72)) if keyed else None
    if keyed:
        ctx.set_sampler_key(key)
    K = len(moduli)
    sk, rk, gk = empty_dev((K, n)), empty_dev((K - 1, 2, K, n)), empty_dev((K - 1, 2, K, n))
    ctx.keygen_secret(3, sk.data_ptr(), stream())
    w = ctx.keygen_workspace_size()
    wb = ws(w)
    seed = 4
    ctx.keygen_relin(sk.data_ptr(), seed, rk.data_ptr(), wb.data_ptr(), w, stream())
    elt = ctx.galois_elt_from_step(1)
    ctx.keygen_galois(sk.data_ptr(), elt, seed + 1, gk.data_ptr(), wb.data_ptr(), w, stream())
    s = from_dev(sk)

    def oracle_key(sd, new_key):
        seeds = [sub_seed(sd, 100 + j) for j in range(K - 1)]
        a = np.stack([ck.sample_poly(2, key, sub_seed(sj, 0), logn, moduli) for sj in seeds])
        e = np.stack([ck.sample_poly(1, key, sub_seed(sj, 1), logn, moduli) for sj in seeds])
        return ck.generate_kswitch_key(s, new_key, a, e, moduli, logn)

    s2 = (s.astype(object) * s.astype(object) % np.array(moduli, dtype=object)[:, None]).astype(np.uint64)
    np.testing.assert_array_equal(from_dev(rk), oracle_key(seed, s2))
    sg = np.zeros_like(s)
    for i in range(K):
        ck.oracle().or_apply_galois_ntt(logn, elt, ck.ptr(np.ascontiguousarray(s[i])), ck.ptr(sg[i]))
    np.testing.assert_array_equal(from_dev(gk), oracle_key(seed + 1, sg))
