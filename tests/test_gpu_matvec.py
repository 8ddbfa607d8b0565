"""BSGS diagonal matrix-vector product (seal-cuda_amd/matvec.py) on the GPU.

A CKKS encryption of a replicated vector v times a d x d matrix M (d = 64 = 8 x 8 baby
and giant steps, N = 2^12): the ciphertext words equal the oracle's composition of the
same operations (rotations as applyGaloisInplace, plain products, sums), and the
decryption decodes to M v within CKKS noise.  The BFV twin (t = 65537, M and v mod t,
the vector replicated in both rows of the slot matrix) decrypts to M v mod t exactly,
and its words equal the oracle's rotations, coefficient-form plain products and sums.
"""
import numpy as np
import pytest
import torch

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2
T = 65537


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def interleave(v):
    return np.stack([np.real(v), np.imag(v)], axis=1).reshape(-1)


def test_bsgs_matvec_words_and_values():
    import matvec
    logn, n = 12, 1 << 12
    slots = n // 2
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(CKKS, n, moduli)
    d, b, g = 64, 8, 8
    scale = 2.0 ** 40
    rng = np.random.default_rng(31)
    M = rng.uniform(-1, 1, (d, d))
    v = rng.uniform(-1, 1, d)
    # keys
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(7, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 8, pk.data_ptr(), stream())
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)

    def galois_key(step, seed):
        k = empty_dev((K - 1, 2, K, n))
        ctx.keygen_galois(sk.data_ptr(), ctx.galois_elt_from_step(step), seed, k.data_ptr(), kb.data_ptr(), kw,
                          stream())
        return k

    baby = [galois_key(j, 100 + j) for j in range(1, b)]
    hoisted = []
    for j, k in zip(range(1, b), baby):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(j), L, h.data_ptr(), hws.data_ptr(), hw,
                                stream())
        hoisted.append(h)
    giant = [galois_key(h * b, 200 + h) for h in range(1, g)]
    # pre-rotated diagonals, encoded at the ciphertext level
    diags = matvec.diagonals(M, slots, b, g)
    plains = empty_dev((g, b, L, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    keep = []
    for h in range(g):
        for j in range(b):
            dv = torch.from_numpy(interleave(diags[h, j])).cuda()
            keep.append(dv)
            ctx.ckks_encode(dv.data_ptr(), slots, scale, L, plains[h, j].data_ptr(), eb.data_ptr(), ew, stream())
    # the vector, replicated over the slots, encrypted
    vs = np.tile(v, slots // d).astype(np.complex128)
    dvv = torch.from_numpy(interleave(vs)).cuda()
    pv = empty_dev((L, n))
    ctx.ckks_encode(dvv.data_ptr(), slots, scale, L, pv.data_ptr(), eb.data_ptr(), ew, stream())
    ct = empty_dev((1, 2, L, n))
    cw = ctx.encrypt_workspace_size(L)
    cb = ws(cw)
    ctx.encrypt(pk.data_ptr(), pv.data_ptr(), L, 9, ct.data_ptr(), cb.data_ptr(), cw, stream())
    # the product on the GPU
    mv = matvec.BsgsMatvec(ctx, n, L, b, g, plains.data_ptr(), [k.data_ptr() for k in baby],
                           [h.data_ptr() for h in hoisted], [k.data_ptr() for k in giant])
    out = empty_dev((1, 2, L, n))
    baby_buf = empty_dev((b, 1, 2, L, n))
    inner = empty_dev((g, 1, 2, L, n))
    wsz = mv.workspace_sizes(1)
    wb = ws(wsz)
    mv.run(ct.data_ptr(), 1, out.data_ptr(), baby_buf.data_ptr(), inner.data_ptr(), wb.data_ptr(), wsz, stream())
    got = from_dev(out)[0]
    # the same operations through the oracle, word for word
    c0 = from_dev(ct)[0]
    P = from_dev(plains)
    rots = [c0] + [ck.apply_galois_keyswitch(CKKS, c0, ctx.galois_elt_from_step(j), from_dev(k), moduli, L, logn)
                   for j, k in zip(range(1, b), baby)]
    want = None
    for h in range(g):
        acc = None
        for j in range(b):
            t = ck.multiply_plain_ntt(rots[j], P[h, j], moduli, logn)
            acc = t if acc is None else ck.mod_add(acc, t, moduli)
        if h > 0:
            acc = ck.apply_galois_keyswitch(CKKS, acc, ctx.galois_elt_from_step(h * b), from_dev(giant[h - 1]),
                                            moduli, L, logn)
        want = acc if want is None else ck.mod_add(want, acc, moduli)
    np.testing.assert_array_equal(got, want)
    # and it computes M v
    dec = empty_dev((L, n))
    ctx.decrypt(out.data_ptr(), 2, sk.data_ptr(), L, dec.data_ptr(), 0, 0, stream())
    res = torch.empty(n, dtype=torch.float64, device="cuda")
    dw = ctx.ckks_decode_workspace_size(L)
    db = ws(dw)
    ctx.ckks_decode(dec.data_ptr(), L, scale * scale, res.data_ptr(), db.data_ptr(), dw, stream())
    torch.cuda.synchronize()
    r = res.cpu().numpy()
    vals = r[0::2] + 1j * r[1::2]
    np.testing.assert_allclose(vals[:d].real, M @ v, atol=1e-3)
    np.testing.assert_allclose(vals[d:2 * d].real, M @ v, atol=1e-3)


def test_bsgs_matvec_bfv_words_and_values():
    import matvec
    logn, n = 12, 1 << 12
    moduli = ck.coeff_modulus_create(n, [50, 50, 60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(BFV, n, moduli, T)
    d, b, g = 64, 8, 8
    rng = np.random.default_rng(37)
    M = rng.integers(0, T, (d, d), dtype=np.uint64)
    v = rng.integers(0, T, d, dtype=np.uint64)
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(17, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 18, pk.data_ptr(), stream())
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)

    def galois_key(step, seed):
        k = empty_dev((K - 1, 2, K, n))
        ctx.keygen_galois(sk.data_ptr(), ctx.galois_elt_from_step(step), seed, k.data_ptr(), kb.data_ptr(), kw,
                          stream())
        return k

    baby = [galois_key(j, 300 + j) for j in range(1, b)]
    hoisted = []
    for j, k in zip(range(1, b), baby):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(j), L, h.data_ptr(), hws.data_ptr(), hw,
                                stream())
        hoisted.append(h)
    giant = [galois_key(h * b, 400 + h) for h in range(1, g)]
    # pre-rotated row diagonals, batch-encoded, lifted to NTT form at level L
    diags = matvec.diagonals(M, n, b, g, rows=2)
    coeffs = empty_dev((g, b, n))
    plains = empty_dev((g, b, L, n))
    for h in range(g):
        for j in range(b):
            dv = to_dev(diags[h, j])
            ctx.batch_encode(dv.data_ptr(), n, coeffs[h, j].data_ptr(), stream())
            ctx.bfv_plain_to_ntt(coeffs[h, j].data_ptr(), L, plains[h, j].data_ptr(), stream())
    # the vector, replicated in both rows, encrypted
    pv = empty_dev((n,))
    dvv = to_dev(np.tile(v, n // d))
    ctx.batch_encode(dvv.data_ptr(), n, pv.data_ptr(), stream())
    ct = empty_dev((1, 2, L, n))
    cw = ctx.encrypt_workspace_size(L)
    cb = ws(cw)
    ctx.encrypt(pk.data_ptr(), pv.data_ptr(), L, 19, ct.data_ptr(), cb.data_ptr(), cw, stream())
    mv = matvec.BsgsMatvec(ctx, n, L, b, g, plains.data_ptr(), [k.data_ptr() for k in baby],
                           [h.data_ptr() for h in hoisted], [k.data_ptr() for k in giant], bfv=True)
    out = empty_dev((1, 2, L, n))
    baby_buf = empty_dev((b, 1, 2, L, n))
    inner = empty_dev((g, 1, 2, L, n))
    wsz = mv.workspace_sizes(1)
    wb = ws(wsz)
    mv.run(ct.data_ptr(), 1, out.data_ptr(), baby_buf.data_ptr(), inner.data_ptr(), wb.data_ptr(), wsz, stream())
    got = from_dev(out)[0]
    # the oracle's composition: rotations, coefficient-form plain products, sums
    c0 = from_dev(ct)[0]
    C = from_dev(coeffs)
    rots = [c0] + [ck.apply_galois_keyswitch(BFV, c0, ctx.galois_elt_from_step(j), from_dev(k), moduli, L, logn)
                   for j, k in zip(range(1, b), baby)]
    want = None
    for h in range(g):
        acc = None
        for j in range(b):
            t = ck.bfv_multiply_plain(rots[j], C[h, j], moduli[:L], T, logn)
            acc = t if acc is None else ck.mod_add(acc, t, moduli)
        if h > 0:
            acc = ck.apply_galois_keyswitch(BFV, acc, ctx.galois_elt_from_step(h * b), from_dev(giant[h - 1]),
                                            moduli, L, logn)
        want = acc if want is None else ck.mod_add(want, acc, moduli)
    np.testing.assert_array_equal(got, want)
    # the same request in a batch of two gives the same words
    ct2 = torch.cat([ct, to_dev(ck.uniform_residues(rng, (1, 2, L, n), moduli[:L]))])
    out2, baby2, inner2 = empty_dev((2, 2, L, n)), empty_dev((b, 2, 2, L, n)), empty_dev((g, 2, 2, L, n))
    wsz2 = mv.workspace_sizes(2)
    wb2 = ws(wsz2)
    mv.run(ct2.data_ptr(), 2, out2.data_ptr(), baby2.data_ptr(), inner2.data_ptr(), wb2.data_ptr(), wsz2, stream())
    np.testing.assert_array_equal(from_dev(out2)[0], got)
    # and it decrypts to M v mod t in every period of both rows
    dw = ctx.decrypt_workspace_size(L)
    db = ws(dw)
    dec, vals, scratch = empty_dev((n,)), empty_dev((n,)), empty_dev((n,))
    ctx.decrypt(out.data_ptr(), 2, sk.data_ptr(), L, dec.data_ptr(), db.data_ptr(), dw, stream())
    ctx.batch_decode(dec.data_ptr(), vals.data_ptr(), scratch.data_ptr(), stream())
    mv_want = (M.astype(object).dot(v.astype(object)) % T).astype(np.uint64)
    np.testing.assert_array_equal(from_dev(vals), np.tile(mv_want, n // d))


def test_bsgs_matvec_ckks_batch_equals_alone():
    """A request's words are the same in a batch of three as computed alone (uniform
    words stand in for encryptions, diagonals and keys)."""
    import matvec
    logn, n = 12, 1 << 12
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(CKKS, n, moduli)
    b = g = 4
    rng = np.random.default_rng(41)
    plains = to_dev(ck.uniform_residues(rng, (g, b, L, n), moduli[:L]))
    baby = [to_dev(ck.random_key(rng, moduli, logn)) for _ in range(1, b)]
    giant = [to_dev(ck.random_key(rng, moduli, logn)) for _ in range(1, g)]
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)
    hoisted = []
    for j, k in zip(range(1, b), baby):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(j), L, h.data_ptr(), hws.data_ptr(), hw,
                                stream())
        hoisted.append(h)
    mv = matvec.BsgsMatvec(ctx, n, L, b, g, plains.data_ptr(), [k.data_ptr() for k in baby],
                           [h.data_ptr() for h in hoisted], [k.data_ptr() for k in giant])
    cts = to_dev(ck.uniform_residues(rng, (3, 2, L, n), moduli[:L]))

    def run(x, count):
        out, bb, inner = empty_dev((count, 2, L, n)), empty_dev((b, count, 2, L, n)), empty_dev((g, count, 2, L, n))
        wsz = mv.workspace_sizes(count)
        wb = ws(wsz)
        mv.run(x.data_ptr(), count, out.data_ptr(), bb.data_ptr(), inner.data_ptr(), wb.data_ptr(), wsz, stream())
        return from_dev(out)

    batched = run(cts, 3)
    for k in range(3):
        np.testing.assert_array_equal(run(cts[k:k + 1].contiguous(), 1)[0], batched[k])
