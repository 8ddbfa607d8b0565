"""GPU parity at the measured workloads' own shapes (bit-exact against the oracle).

The smaller-N tests elsewhere pin every kernel's arithmetic; these run the launch
shapes that the benchmarked workloads use and compare words with the oracle there:

* hoisted rotations at N=2^16, CoeffModulus::Create(65536, {60, 50 x 23, 60}) (24 data
  primes + special), the 15 rotations by 2^k (four passes of the hoisted products over
  one ModUp), three ciphertexts -- shc_rotate_sum and shc_rotate_many against the
  oracle's applyGaloisInplace (src/evaluator.cpp:2153-2307) summed, for two of them;
* batched relinearization at N=2^16 with 24 + 1 primes of 60 bits, 7 ciphertexts (the
  fused key-product kernel's groups of 4, 2 and 1), CKKS and BFV, and a BFV multiply
  (BEHZ) + relinearize at that shape;
* the graph-captured CKKS chain at N=2^15, Create(32768, {60, 40 x 14, 60}), level
  groups {4, 9, 15}, three (multiply, relinearize, rescale) rounds;
* baby-step giant-step matrix-vector products at the serving sizes: BFV N=8192
  (BFVDefault(8192) primes, t=65537), 4096 x 4096, 64 x 64; CKKS N=2^15, 1024 x 1024,
  32 x 32 -- words against the oracle's composition of rotations, plain products and
  sums (app/matmul.cuh:98-121), and the decryption against M v; the diagonal products
  alone at the CKKS shape (polynomials in groups of four, XCD-ordered grid);
* the per-operation workload's CKKS shape (N=2^16, level 24 of 60 x 25): batched
  public-key encryption (fused into the error term's transform) against the oracle's
  encryption of each seed, batched decryption, and the batched CKKSEncoder against its
  single calls.

Inputs are seeded; large keys are uniform words drawn on the device by the library's
counter-based sampler and copied to the host for the oracle.  Oracle calls run on a
thread pool.
"""
import numpy as np
import pytest
import torch

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2
UNIFORM = 2
T = 65537
BFV_DEFAULT_8192 = [0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001]


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def device_uniform(ctx, shape, limbs, seed):
    """A device tensor [..., limbs, N] of uniform residues of the first `limbs` key-level
    primes, one sampler call per [limbs][N] block."""
    t = empty_dev(shape)
    blocks = t.view(-1, limbs, ctx.n)
    for b in range(blocks.shape[0]):
        ctx.sample_poly(UNIFORM, seed * 100003 + b, limbs, blocks[b].data_ptr(), stream())
    return t


def device_key(ctx, K, seed):
    return device_uniform(ctx, (K - 1, 2, K, ctx.n), K, seed)


# ---------------------------------------------------------------- hoisted rotations (N=2^16)

def test_rotate_sum_n65536_15_rotations():
    logn, n = 16, 1 << 16
    moduli = ck.coeff_modulus_create(n, [60] + [50] * 23 + [60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(CKKS, n, moduli)
    steps = [1 << k for k in range(15)]
    elts = [ctx.galois_elt_from_step(s) for s in steps]
    count = 3
    din = device_uniform(ctx, (count, 2, L, n), L, 1)
    dkeys = [device_key(ctx, K, 10 + r) for r in range(len(steps))]
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)
    hkeys = []
    for e, dk in zip(elts, dkeys):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(dk.data_ptr(), e, L, h.data_ptr(), hws.data_ptr(), hw, stream())
        hkeys.append(h)
    del hws
    out = empty_dev((count, 2, L, n))
    many = empty_dev((len(elts), count, 2, L, n))
    rw = ctx.rotate_sum_workspace_size(count, L)
    rws = ws(rw)
    hp, kp = [h.data_ptr() for h in hkeys], [k.data_ptr() for k in dkeys]
    ctx.rotate_sum(din.data_ptr(), count, L, elts, hp, kp, out.data_ptr(), rws.data_ptr(), rw, stream())
    ctx.rotate_many(din.data_ptr(), count, L, elts, hp, kp, many.data_ptr(), rws.data_ptr(), rw, stream())
    got, got_many, cts = from_dev(out), from_dev(many), from_dev(din)
    del hkeys, rws, many, out
    keys = [from_dev(k) for k in dkeys]
    del dkeys
    torch.cuda.empty_cache()
    checked = (0, count - 1)
    jobs = [(c, r) for c in checked for r in range(len(elts))]
    rots = ck.parallel_map(
        lambda cr: ck.apply_galois_keyswitch(CKKS, cts[cr[0]], elts[cr[1]], keys[cr[1]], moduli, L, logn), jobs)
    for (c, r), rot in zip(jobs, rots):
        np.testing.assert_array_equal(got_many[r, c], rot, err_msg=f"rotation by {steps[r]} of ciphertext {c}")
    for c in checked:
        want = None
        for (cc, r), rot in zip(jobs, rots):
            if cc == c:
                want = rot if want is None else ck.mod_add(want, rot, moduli)
        np.testing.assert_array_equal(got[c], want, err_msg=f"sum of 15 rotations of ciphertext {c}")
    # every ciphertext's sum equals the sum of its rotate_many outputs (closed form)
    for c in range(count):
        acc = got_many[0, c]
        for r in range(1, len(elts)):
            acc = ck.mod_add(acc, got_many[r, c], moduli)
        np.testing.assert_array_equal(got[c], acc, err_msg=f"ciphertext {c}: sum vs listed rotations")


def test_rotate_many_n65536_zero_digits():
    """Config 4's primes with zero digits in c1 (ten coefficient-form digit words of
    ciphertext 0 set to zero, in several limbs): the device-listed fix of the hoisted
    products keeps every rotation equal to the oracle's separate rotation."""
    logn, n = 16, 1 << 16
    moduli = ck.coeff_modulus_create(n, [60] + [50] * 23 + [60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(CKKS, n, moduli)
    steps = [1, 4, 64, 1024, -3]
    elts = [ctx.galois_elt_from_step(s) for s in steps]
    count = 2
    cts = from_dev(device_uniform(ctx, (count, 2, L, n), L, 7)).copy()
    c1 = ck.ntt_rns(cts[0, 1], moduli[:L], logn, inverse=True)
    for k, (limb, pos) in enumerate([(0, 5), (0, 40000), (3, 17), (11, 65535), (23, 1), (7, 32768), (7, 32769),
                                     (19, 123), (2, 4096), (15, 60001)]):
        c1[limb, pos] = 0
    cts[0, 1] = ck.ntt_rns(c1, moduli[:L], logn)
    din = to_dev(cts)
    dkeys = [device_key(ctx, K, 40 + r) for r in range(len(steps))]
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)
    hkeys = []
    for e, dk in zip(elts, dkeys):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(dk.data_ptr(), e, L, h.data_ptr(), hws.data_ptr(), hw, stream())
        hkeys.append(h)
    del hws
    many = empty_dev((len(elts), count, 2, L, n))
    rw = ctx.rotate_sum_workspace_size(count, L)
    rws = ws(rw)
    ctx.rotate_many(din.data_ptr(), count, L, elts, [h.data_ptr() for h in hkeys], [k.data_ptr() for k in dkeys],
                    many.data_ptr(), rws.data_ptr(), rw, stream())
    got = from_dev(many)
    keys = [from_dev(k) for k in dkeys]
    del hkeys, rws, many, dkeys
    torch.cuda.empty_cache()
    rots = ck.parallel_map(lambda r: ck.apply_galois_keyswitch(CKKS, cts[0], elts[r], keys[r], moduli, L, logn),
                           list(range(len(elts))))
    for r, rot in enumerate(rots):
        np.testing.assert_array_equal(got[r, 0], rot, err_msg=f"rotation by {steps[r]} with zero digits")


# ---------------------------------------------------------------- batched relinearize (N=2^16, L=24)

@pytest.mark.parametrize("scheme", [CKKS, BFV])
def test_relinearize_n65536_l24_batch7(scheme):
    """Seven ciphertexts: the fused key-product kernel runs a group of 4, a group of 2 and
    the single-ciphertext variant; one ciphertext of each is checked."""
    logn, n, K = 16, 1 << 16, 25
    L = K - 1
    moduli = ck.coeff_modulus_create(n, [60] * K)
    ctx = hip().Context(scheme, n, moduli, T if scheme == BFV else 0)
    count = 7
    d = device_uniform(ctx, (count, 3, L, n), L, 20 + scheme)
    dk = device_key(ctx, K, 30 + scheme)
    ct3, key = from_dev(d), from_dev(dk)
    wsb = ctx.keyswitch_workspace_size(count, L)
    w = ws(wsb)
    ctx.relinearize(d.data_ptr(), dk.data_ptr(), count, L, w.data_ptr(), wsb, stream())
    got = from_dev(d)[:, :2]
    del w, d, dk
    torch.cuda.empty_cache()
    checked = (0, 4, 6)
    want = ck.parallel_map(lambda c: ck.relinearize(scheme, ct3[c], key, moduli, L, logn), checked)
    for c, wnt in zip(checked, want):
        np.testing.assert_array_equal(got[c], wnt, err_msg=f"ciphertext {c}")


def test_bfv_multiply_relinearize_n65536_l24_batch5():
    """BFV multiply (BEHZ) + relinearize, N=2^16, 24 + 1 primes, t = 65537, five pairs
    (a group of 4 and a single ciphertext in the key switch); pairs 0 and 4 checked."""
    logn, n, K = 16, 1 << 16, 25
    L = K - 1
    moduli = ck.coeff_modulus_create(n, [60] * K)
    ctx = hip().Context(BFV, n, moduli, T)
    count = 5
    da = device_uniform(ctx, (count, 2, L, n), L, 40)
    db = device_uniform(ctx, (count, 2, L, n), L, 41)
    dk = device_key(ctx, K, 42)
    a, b, key = from_dev(da), from_dev(db), from_dev(dk)
    out = empty_dev((count, 3, L, n))
    mw, kw = ctx.multiply_workspace_size(count, L), ctx.keyswitch_workspace_size(count, L)
    w = ws(max(mw, kw))
    ctx.multiply(da.data_ptr(), db.data_ptr(), out.data_ptr(), count, L, w.data_ptr(), mw, stream())
    ctx.relinearize(out.data_ptr(), dk.data_ptr(), count, L, w.data_ptr(), kw, stream())
    got = from_dev(out)[:, :2]
    del w, out, da, db, dk
    torch.cuda.empty_cache()

    def oracle(c):
        return ck.relinearize(BFV, ck.bfv_multiply(a[c], b[c], moduli[:L], T, logn), key, moduli, L, logn)

    for c, wnt in zip((0, 4), ck.parallel_map(oracle, (0, 4))):
        np.testing.assert_array_equal(got[c], wnt, err_msg=f"pair {c}")


# ---------------------------------------------------------------- CKKS chain in a graph (N=2^15)

def test_graph_chain_n32768_levels_4_9_15():
    logn, n = 15, 1 << 15
    moduli = ck.coeff_modulus_create(n, [60] + [40] * 14 + [60])
    K = len(moduli)
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(23)
    dkey = device_key(ctx, K, 50)
    key_h = from_dev(dkey)
    groups = {4: 2, 9: 3, 15: 2}
    rounds = 3
    bufs = {}
    max_ws = 0
    for l, c in groups.items():
        x = ck.uniform_residues(rng, (c, 2, l, n), moduli[:l])
        ys = [ck.uniform_residues(rng, (c, 2, l - s, n), moduli[:l - s]) for s in range(rounds)]
        dev = dict(x=to_dev(x), ys=[to_dev(y) for y in ys],
                   prods=[empty_dev((c, 3, l - s, n)) for s in range(rounds)],
                   rels=[empty_dev((c, 2, l - s, n)) for s in range(rounds)],
                   accs=[empty_dev((c, 2, l - s - 1, n)) for s in range(rounds)])
        bufs[l] = (c, x, ys, dev)
        for s in range(rounds):
            max_ws = max(max_ws, ctx.keyswitch_workspace_size(c, l - s), ctx.rescale_workspace_size(c, 2, l - s))
    w = ws(max_ws)

    def chain():
        st = torch.cuda.current_stream().cuda_stream
        for l, (c, _, _, d) in bufs.items():
            cur = d["x"]
            for s in range(rounds):
                ll = l - s
                ctx.ckks_multiply(cur.data_ptr(), d["ys"][s].data_ptr(), d["prods"][s].data_ptr(), c, ll, st)
                ctx.relinearize(d["prods"][s].data_ptr(), dkey.data_ptr(), c, ll, w.data_ptr(), max_ws, st)
                d["rels"][s].copy_(d["prods"][s][:, :2])
                ctx.rescale_to_next(d["rels"][s].data_ptr(), d["accs"][s].data_ptr(), c, 2, ll, w.data_ptr(),
                                    max_ws, st)
                cur = d["accs"][s]

    chain()
    torch.cuda.synchronize()
    eager = {l: from_dev(d["accs"][-1]).copy() for l, (_, _, _, d) in bufs.items()}

    def oracle(job):
        l, i = job
        _, x, ys, _ = bufs[l]
        cur = x[i]
        for s in range(rounds):
            ll = l - s
            o3 = ck.ckks_multiply(cur, ys[s][i], moduli[:ll], logn)
            o2 = ck.relinearize(CKKS, o3, key_h, moduli, ll, logn)
            cur = ck.rescale(CKKS, o2, moduli, logn)
        return cur

    jobs = [(l, i) for l, (c, _, _, _) in bufs.items() for i in range(c)]
    for (l, i), want in zip(jobs, ck.parallel_map(oracle, jobs)):
        np.testing.assert_array_equal(eager[l][i], want, err_msg=f"level {l}, pair {i}")

    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(graph, stream=cap):
        chain()
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    for _ in range(3):
        for (_, _, _, d) in bufs.values():
            for a in d["accs"]:
                a.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for l, (_, _, _, d) in bufs.items():
            np.testing.assert_array_equal(from_dev(d["accs"][-1]), eager[l], err_msg=f"replay, level {l}")
        chain()   # eager work between replays
        torch.cuda.synchronize()


# ---------------------------------------------------------------- BSGS matvec at the serving sizes

def interleave(v):
    return np.stack([np.real(v), np.imag(v)], axis=1).reshape(-1)


def galois_keys(ctx, sk, K, steps, seed0):
    kw = ctx.keygen_workspace_size()
    kb = ws(kw)
    out = []
    for i, s in enumerate(steps):
        k = empty_dev((K - 1, 2, K, ctx.n))
        ctx.keygen_galois(sk.data_ptr(), ctx.galois_elt_from_step(s), seed0 + i, k.data_ptr(), kb.data_ptr(), kw,
                          stream())
        out.append(k)
    return out


def hoisted_keys(ctx, keys, steps, L):
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)
    out = []
    for s, k in zip(steps, keys):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(s), L, h.data_ptr(), hws.data_ptr(), hw,
                                stream())
        out.append(h)
    return out


def run_matvec(ctx, mv, ct, L, b, g):
    n = ctx.n
    count = ct.shape[0]
    out = empty_dev((count, 2, L, n))
    baby_buf = empty_dev((b, count, 2, L, n))
    inner = empty_dev((g, count, 2, L, n))
    wsz = mv.workspace_sizes(count)
    wb = ws(wsz)
    mv.run(ct.data_ptr(), count, out.data_ptr(), baby_buf.data_ptr(), inner.data_ptr(), wb.data_ptr(), wsz, stream())
    return from_dev(out)


def test_bsgs_matvec_bfv_n8192_4096x4096():
    import matvec
    logn, n = 13, 1 << 13
    moduli = BFV_DEFAULT_8192
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(BFV, n, moduli, T)
    d, b, g = 4096, 64, 64
    rng = np.random.default_rng(61)
    M = rng.integers(0, T, (d, d), dtype=np.uint64)
    v = rng.integers(0, T, d, dtype=np.uint64)
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(71, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 72, pk.data_ptr(), stream())
    baby_steps, giant_steps = list(range(1, b)), [h * b for h in range(1, g)]
    baby = galois_keys(ctx, sk, K, baby_steps, 1000)
    giant = galois_keys(ctx, sk, K, giant_steps, 2000)
    hoisted = hoisted_keys(ctx, baby, baby_steps, L)
    diags = matvec.diagonals(M, n, b, g, rows=2)
    coeffs = empty_dev((g, b, n))
    plains = empty_dev((g, b, L, n))
    for h in range(g):
        for j in range(b):
            dv = to_dev(diags[h, j])
            ctx.batch_encode(dv.data_ptr(), n, coeffs[h, j].data_ptr(), stream())
            ctx.bfv_plain_to_ntt(coeffs[h, j].data_ptr(), L, plains[h, j].data_ptr(), stream())
    pv = empty_dev((n,))
    ctx.batch_encode(to_dev(np.tile(v, n // d)).data_ptr(), n, pv.data_ptr(), stream())
    ct = empty_dev((1, 2, L, n))
    cw = ctx.encrypt_workspace_size(L)
    ctx.encrypt(pk.data_ptr(), pv.data_ptr(), L, 73, ct.data_ptr(), ws(cw).data_ptr(), cw, stream())
    mv = matvec.BsgsMatvec(ctx, n, L, b, g, plains.data_ptr(), [k.data_ptr() for k in baby],
                           [h.data_ptr() for h in hoisted], [k.data_ptr() for k in giant], bfv=True)
    got = run_matvec(ctx, mv, ct, L, b, g)[0]
    # the oracle's composition: rotations, coefficient-form plain products, sums
    c0 = from_dev(ct)[0]
    C = from_dev(coeffs)
    bk, gk = [from_dev(k) for k in baby], [from_dev(k) for k in giant]
    rots = [c0] + ck.parallel_map(
        lambda j: ck.apply_galois_keyswitch(BFV, c0, ctx.galois_elt_from_step(j), bk[j - 1], moduli, L, logn),
        baby_steps)

    def giant_term(h):
        acc = None
        for j in range(b):
            t = ck.bfv_multiply_plain(rots[j], C[h, j], moduli[:L], T, logn)
            acc = t if acc is None else ck.mod_add(acc, t, moduli)
        if h > 0:
            acc = ck.apply_galois_keyswitch(BFV, acc, ctx.galois_elt_from_step(h * b), gk[h - 1], moduli, L, logn)
        return acc

    want = None
    for acc in ck.parallel_map(giant_term, range(g)):
        want = acc if want is None else ck.mod_add(want, acc, moduli)
    np.testing.assert_array_equal(got, want)
    # it decrypts to M v mod t in both rows of the slot matrix
    dw = ctx.decrypt_workspace_size(L)
    dec, vals, scratch = empty_dev((n,)), empty_dev((n,)), empty_dev((n,))
    out = to_dev(got[None])
    ctx.decrypt(out.data_ptr(), 2, sk.data_ptr(), L, dec.data_ptr(), ws(dw).data_ptr(), dw, stream())
    ctx.batch_decode(dec.data_ptr(), vals.data_ptr(), scratch.data_ptr(), stream())
    mv_want = (M.astype(object).dot(v.astype(object)) % T).astype(np.uint64)
    np.testing.assert_array_equal(from_dev(vals), np.tile(mv_want, n // d))


def test_bsgs_matvec_ckks_n32768_1024x1024():
    import matvec
    logn, n = 15, 1 << 15
    slots = n // 2
    moduli = ck.coeff_modulus_create(n, [60] + [40] * 14 + [60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(CKKS, n, moduli)
    d, b, g = 1024, 32, 32
    scale = 2.0 ** 40
    rng = np.random.default_rng(67)
    M = rng.uniform(-1, 1, (d, d))
    v = rng.uniform(-1, 1, d)
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(81, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 82, pk.data_ptr(), stream())
    baby_steps, giant_steps = list(range(1, b)), [h * b for h in range(1, g)]
    baby = galois_keys(ctx, sk, K, baby_steps, 3000)
    giant = galois_keys(ctx, sk, K, giant_steps, 4000)
    hoisted = hoisted_keys(ctx, baby, baby_steps, L)
    diags = matvec.diagonals(M, slots, b, g)
    plains = empty_dev((g, b, L, n))
    ew = ctx.ckks_encode_workspace_size()
    eb = ws(ew)
    for h in range(g):
        for j in range(b):
            dv = torch.from_numpy(interleave(diags[h, j])).cuda()
            ctx.ckks_encode(dv.data_ptr(), slots, scale, L, plains[h, j].data_ptr(), eb.data_ptr(), ew, stream())
    dvv = torch.from_numpy(interleave(np.tile(v, slots // d).astype(np.complex128))).cuda()
    pv = empty_dev((L, n))
    ctx.ckks_encode(dvv.data_ptr(), slots, scale, L, pv.data_ptr(), eb.data_ptr(), ew, stream())
    ct = empty_dev((1, 2, L, n))
    cw = ctx.encrypt_workspace_size(L)
    ctx.encrypt(pk.data_ptr(), pv.data_ptr(), L, 83, ct.data_ptr(), ws(cw).data_ptr(), cw, stream())
    mv = matvec.BsgsMatvec(ctx, n, L, b, g, plains.data_ptr(), [k.data_ptr() for k in baby],
                           [h.data_ptr() for h in hoisted], [k.data_ptr() for k in giant])
    got = run_matvec(ctx, mv, ct, L, b, g)[0]
    c0 = from_dev(ct)[0]
    P = from_dev(plains)
    del hoisted
    bk = [from_dev(k) for k in baby]
    del baby
    gk = [from_dev(k) for k in giant]
    del giant
    torch.cuda.empty_cache()
    rots = [c0] + ck.parallel_map(
        lambda j: ck.apply_galois_keyswitch(CKKS, c0, ctx.galois_elt_from_step(j), bk[j - 1], moduli, L, logn),
        baby_steps)

    def giant_term(h):
        acc = None
        for j in range(b):
            t = ck.multiply_plain_ntt(rots[j], P[h, j], moduli, logn)
            acc = t if acc is None else ck.mod_add(acc, t, moduli)
        if h > 0:
            acc = ck.apply_galois_keyswitch(CKKS, acc, ctx.galois_elt_from_step(h * b), gk[h - 1], moduli, L, logn)
        return acc

    want = None
    for acc in ck.parallel_map(giant_term, range(g)):
        want = acc if want is None else ck.mod_add(want, acc, moduli)
    np.testing.assert_array_equal(got, want)
    # it computes M v (product scale 2^80 at level 15)
    dec = empty_dev((L, n))
    out = to_dev(got[None])
    ctx.decrypt(out.data_ptr(), 2, sk.data_ptr(), L, dec.data_ptr(), 0, 0, stream())
    res = torch.empty(n, dtype=torch.float64, device="cuda")
    dw = ctx.ckks_decode_workspace_size(L)
    ctx.ckks_decode(dec.data_ptr(), L, scale * scale, res.data_ptr(), ws(dw).data_ptr(), dw, stream())
    torch.cuda.synchronize()
    r = res.cpu().numpy()
    vals = r[0::2] + 1j * r[1::2]
    want_mv = M @ v
    for period in (0, slots // d - 1):
        np.testing.assert_allclose(vals[period * d:(period + 1) * d].real, want_mv, atol=1e-3)


# ------------------------------------------- encryption, decryption, encoders (N=2^16, L=24)

M64 = (1 << 64) - 1


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _sub_seed(seed, k):
    return _mix64((seed ^ _mix64((k + 0x5EA15EED) & M64)) & M64)


def test_encrypt_decrypt_encode_n65536_l24():
    """The per-operation workload's shapes: shc_encrypt_many of three NTT-form plaintexts
    at N=2^16, level 24 of Create(65536, {60 x 25}) (the encryption fused into the error
    term's transform, the special prime dropped) equals the oracle's encryption with the
    streams of each seed; shc_decrypt_many of those ciphertexts equals the oracle's
    phase; shc_ckks_encode_many / shc_ckks_decode_many of three plaintexts equal their
    single calls, and decode(encode(x)) is within 2^-20 of x."""
    logn, n = 16, 1 << 16
    moduli = ck.coeff_modulus_create(n, [60] * 25)
    K, L = len(moduli), len(moduli) - 1
    ctx = hip().Context(CKKS, n, moduli)
    sk, pk = empty_dev((K, n)), empty_dev((2, K, n))
    ctx.keygen_secret(3, sk.data_ptr(), stream())
    ctx.keygen_public(sk.data_ptr(), 4, pk.data_ptr(), stream())
    count, seed = 3, 77
    plain = device_uniform(ctx, (count, L, n), L, 5)
    out = empty_dev((count, 2, L, n))
    w = ctx.encrypt_many_workspace_size(L, count)
    wb = ws(w)
    ctx.encrypt_many(pk.data_ptr(), plain.data_ptr(), count, L, seed, out.data_ptr(), wb.data_ptr(), w, stream())
    got, pk_h, pl_h = from_dev(out), from_dev(pk), from_dev(plain)

    def expected(g):
        u = ck.sample_poly(0, None, _sub_seed(seed + g, 0), logn, moduli)
        e0 = ck.sample_poly(1, None, _sub_seed(seed + g, 1), logn, moduli)
        e1 = ck.sample_poly(1, None, _sub_seed(seed + g, 2), logn, moduli)
        want = ck.rescale(CKKS, ck.encrypt_zero_asymmetric(CKKS, pk_h, u, e0, e1, moduli, logn), moduli, logn)
        want[0] = ck.mod_add(want[0], pl_h[g], moduli)
        return want

    for g, want in enumerate(ck.parallel_map(expected, range(count))):
        np.testing.assert_array_equal(got[g], want, err_msg=f"ciphertext {g}")
    dec = empty_dev((count, L, n))
    ctx.decrypt_many(out.data_ptr(), 2, count, sk.data_ptr(), L, dec.data_ptr(), 0, 0, stream())
    sk_h = from_dev(sk)
    phases = ck.parallel_map(lambda g: ck.decrypt_phase(CKKS, got[g], sk_h[:L], moduli, logn), range(count))
    np.testing.assert_array_equal(from_dev(dec), np.stack(phases))

    g = torch.Generator(device="cuda").manual_seed(9)
    vals = torch.rand((count, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    cw = ctx.ckks_encode_many_workspace_size(count)
    cb = ws(cw)
    plains = empty_dev((count, L, n))
    status = torch.zeros((count,), dtype=torch.int32, device="cuda")
    ctx.ckks_encode_many(vals.data_ptr(), n // 2, n, count, 2.0 ** 40, L, plains.data_ptr(), status.data_ptr(),
                         cb.data_ptr(), cw, stream())
    ew1 = ctx.ckks_encode_workspace_size()
    eb1 = ws(ew1)
    one = empty_dev((L, n))
    many = from_dev(plains)
    assert not status.any()
    for p in range(count):
        ctx.ckks_encode(vals[p].data_ptr(), n // 2, 2.0 ** 40, L, one.data_ptr(), eb1.data_ptr(), ew1, stream())
        np.testing.assert_array_equal(many[p], from_dev(one), err_msg=f"plaintext {p}")
    dw = ctx.ckks_decode_many_workspace_size(L, count)
    db = ws(dw)
    back = torch.empty((count, n), dtype=torch.float64, device="cuda")
    ctx.ckks_decode_many(plains.data_ptr(), L, count, 2.0 ** 40, back.data_ptr(), db.data_ptr(), dw, stream())
    torch.cuda.synchronize()
    assert float((back - vals).abs().max()) < 2.0 ** -20


def test_plain_dot_many_matvec_shape():
    """The diagonal products at the CKKS matvec's shape (N=2^15, 15 primes): 12
    polynomials in groups of four over 4 terms, 9 outputs (an 8-wide and a 1-wide pass),
    positions in the XCD-ordered grid, against the oracle's plain products and sums."""
    logn, n = 15, 1 << 15
    moduli = ck.coeff_modulus_create(n, [60] + [40] * 14 + [60])
    L, terms, outputs, polys = 15, 4, 9, 12
    ctx = hip().Context(CKKS, n, moduli)
    cts = device_uniform(ctx, (terms, polys, L, n), L, 21)
    plains = device_uniform(ctx, (outputs, terms, L, n), L, 22)
    out = empty_dev((outputs, polys, L, n))
    ctx.multiply_plain_accumulate_many(cts.data_ptr(), plains.data_ptr(), terms, outputs, polys, L, out.data_ptr(),
                                       stream())
    got, c_h, p_h = from_dev(out), from_dev(cts), from_dev(plains)

    def expected(o):
        want = None
        for j in range(terms):
            t = ck.multiply_plain_ntt(c_h[j], p_h[o, j], moduli[:L], logn)
            want = t if want is None else ck.mod_add(want, t, moduli)
        return want

    for o, want in enumerate(ck.parallel_map(expected, range(outputs))):
        np.testing.assert_array_equal(got[o], want, err_msg=f"output {o}")
