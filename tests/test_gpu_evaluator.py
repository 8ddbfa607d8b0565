"""GPU parity of multiply / relinearize / rescale against the oracle (bit-exact words).

Shapes follow the reference's evaluator tests (small N, test/evaluator_cuda.cu) and
the measured workloads: mul+relin at N=2^16 with 24 data primes + 1 special prime,
and the CKKS chain at N=2^15 with 60 + 14x40 + 60 bits.  Inputs and keys are
injected uniform words (the same words go to both paths).
"""
import numpy as np
import pytest
import torch

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def make_ctx(scheme, logn, bits, plain=0):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    return hip().Context(scheme, n, moduli, plain), moduli


def run_relin(ctx, ct3_batch, key, L):
    count = ct3_batch.shape[0]
    d = to_dev(ct3_batch)
    dk = to_dev(key)
    ws_bytes = ctx.keyswitch_workspace_size(count, L)
    ws = empty_dev((ws_bytes + 7) // 8)
    ctx.relinearize(d.data_ptr(), dk.data_ptr(), count, L, ws.data_ptr(), ws_bytes, stream())
    return from_dev(d)[:, :2]


@pytest.mark.parametrize("logn,bits,L,count", [
    (10, [60, 40, 40, 60], 3, 3),
    (12, [60, 40, 40, 40, 60], 2, 2),
    (13, [50] * 6, 5, 1),
])
def test_ckks_multiply(logn, bits, L, count):
    ctx, moduli = make_ctx(CKKS, logn, bits)
    n = 1 << logn
    rng = np.random.default_rng(logn)
    a = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    da, db, do = to_dev(a), to_dev(b), empty_dev((count, 3, L, n))
    ctx.ckks_multiply(da.data_ptr(), db.data_ptr(), do.data_ptr(), count, L, stream())
    got = from_dev(do)
    for c in range(count):
        np.testing.assert_array_equal(got[c], ck.ckks_multiply(a[c], b[c], moduli[:L], logn))


@pytest.mark.parametrize("scheme,logn,bits,L,count", [
    (CKKS, 10, [60, 40, 40, 60], 3, 3),
    (CKKS, 11, [60, 40, 40, 40, 60], 1, 2),
    (BFV, 10, [40, 40, 50], 2, 2),
    (BFV, 13, [43, 43, 44, 44, 44], 4, 2),
    (CKKS, 15, [60] + [40] * 14 + [60], 15, 1),
    # 60-bit primes at 20 digits, and 5 ciphertexts: a partial last group of 4
    (CKKS, 13, [60] * 21, 20, 5),
    # the ModUp rows pass fused with the key products (N = 2^15, 2^16): groups of 4
    # ciphertexts with a partial last one, CKKS (own-prime limbs from the input) and BFV
    (CKKS, 15, [60] * 9, 8, 6),
    (BFV, 15, [60] * 7, 6, 3),
    (CKKS, 16, [60] * 4, 3, 5),
    # 33 digits: the fused kernel reduces the transformed words below 4q before the
    # products (the 128-bit sum bound); groups of 4, 2 and 1
    (CKKS, 15, [60] * 34, 33, 7),
    # the largest ring degree, and a level of one prime
    (CKKS, 17, [60] * 4, 3, 2),
    (CKKS, 12, [50, 50], 1, 2),
])
def test_relinearize(scheme, logn, bits, L, count):
    ctx, moduli = make_ctx(scheme, logn, bits, 65537 if scheme == BFV else 0)
    n = 1 << logn
    rng = np.random.default_rng(100 + logn)
    ct3 = ck.uniform_residues(rng, (count, 3, L, n), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    got = run_relin(ctx, ct3, key, L)
    for c in range(count):
        np.testing.assert_array_equal(got[c], ck.relinearize(scheme, ct3[c], key, moduli, L, logn))


def test_relinearize_columns_all_targets():
    """A batch large enough that each ModUp columns block transforms its digit tile for
    all L + 1 targets (104 x 20 x 8 tiles >= 16384 blocks); three ciphertexts against the
    oracle, spread over the batch."""
    logn, L, count = 13, 20, 104
    ctx, moduli = make_ctx(CKKS, logn, [60] * (L + 1))
    rng = np.random.default_rng(13)
    ct3 = ck.uniform_residues(rng, (count, 3, L, 1 << logn), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    got = run_relin(ctx, ct3, key, L)
    for c in (0, count // 2 - 1, count - 1):
        np.testing.assert_array_equal(got[c], ck.relinearize(CKKS, ct3[c], key, moduli, L, logn))


def test_relinearize_target_shape():
    """N = 2^16, 24 data primes + 1 special prime (60 bits each), one ciphertext."""
    logn, L = 16, 24
    ctx, moduli = make_ctx(CKKS, logn, [60] * 25)
    rng = np.random.default_rng(7)
    ct3 = ck.uniform_residues(rng, (1, 3, L, 1 << logn), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    got = run_relin(ctx, ct3, key, L)
    np.testing.assert_array_equal(got[0], ck.relinearize(CKKS, ct3[0], key, moduli, L, logn))


def test_batch_equals_alone():
    """Each ciphertext gives the same words alone and inside a batch."""
    logn, bits, L = 12, [60, 40, 40, 40, 60], 4
    ctx, moduli = make_ctx(CKKS, logn, bits)
    rng = np.random.default_rng(11)
    ct3 = ck.uniform_residues(rng, (7, 3, L, 1 << logn), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    batch = run_relin(ctx, ct3, key, L)
    for c in (0, 3, 6):
        alone = run_relin(ctx, ct3[c:c + 1], key, L)
        np.testing.assert_array_equal(batch[c], alone[0])


@pytest.mark.parametrize("scheme,logn,bits,L,count,polys", [
    (CKKS, 10, [60, 40, 40, 60], 3, 2, 2),
    (CKKS, 15, [60] + [40] * 14 + [60], 15, 2, 2),
    (CKKS, 13, [60, 40, 40, 60], 3, 1, 3),       # an odd number of polynomials
    (CKKS, 13, [60, 40, 40, 60], 4, 3, 2),       # key level: the special prime dropped
    (CKKS, 16, [60] * 25, 24, 3, 2),             # the headline shape
    (CKKS, 17, [60] * 4, 2, 2, 3),
    (BFV, 12, [43, 43, 44, 44, 44], 4, 2, 2),
    (BFV, 14, [50, 50, 50, 50], 3, 3, 3),
])
def test_rescale(scheme, logn, bits, L, count, polys):
    """rescale_to_next words equal the oracle's per polynomial: CKKS through the
    divide-and-round fused into the correction's transform (N >= 2^13), BFV in one pass."""
    ctx, moduli = make_ctx(scheme, logn, bits, 65537 if scheme == BFV else 0)
    n = 1 << logn
    rng = np.random.default_rng(200 + logn)
    ct = ck.uniform_residues(rng, (count, polys, L, n), moduli[:L])
    din, dout = to_dev(ct), empty_dev((count, polys, L - 1, n))
    ws_bytes = ctx.rescale_workspace_size(count, polys, L)
    ws = empty_dev((ws_bytes + 7) // 8)
    ctx.rescale_to_next(din.data_ptr(), dout.data_ptr(), count, polys, L, ws.data_ptr(), ws_bytes, stream())
    got = from_dev(dout)
    for c in range(count):
        np.testing.assert_array_equal(got[c], ck.rescale(scheme, ct[c], moduli, logn))


@pytest.mark.parametrize("logn,bits,L,t,count", [
    (10, [40, 40, 50], 2, 65537, 2),
    (13, [43, 43, 44, 44, 44], 4, 65537, 2),   # BFV plain path: BFVDefault-sized primes, t = 65537
    (11, [60, 60, 60], 2, 1024, 1),
    (10, [30, 30, 30, 40], 3, 786433, 3),
    (4, [30, 30, 40], 2, 97, 2),
    # a 40-bit t with two 60-bit primes: 32 + bits(t) + bits(q) >= 61 (L + 1), so B takes
    # one more prime (|B| = L + 1, |Bsk| = L + 2: uneven output slices in the conversions)
    (12, [60, 60, 60], 2, None, 3),
    (15, [60] * 5, 4, None, 5),
])
def test_bfv_multiply(logn, bits, L, t, count):
    if t is None:  # a 40-bit prime plain modulus (= 1 mod 2N, coprime to the 60-bit q)
        t = ck.coeff_modulus_create(1 << logn, [40])[0]
    ctx, moduli = make_ctx(BFV, logn, bits, t)
    n = 1 << logn
    rng = np.random.default_rng(300 + logn)
    a = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    da, db, do = to_dev(a), to_dev(b), empty_dev((count, 3, L, n))
    ws_bytes = ctx.multiply_workspace_size(count, L)
    ws = empty_dev((ws_bytes + 7) // 8)
    ctx.multiply(da.data_ptr(), db.data_ptr(), do.data_ptr(), count, L, ws.data_ptr(), ws_bytes, stream())
    got = from_dev(do)
    for c in range(count):
        np.testing.assert_array_equal(got[c], ck.bfv_multiply(a[c], b[c], moduli[:L], t, logn))


def test_bfv_multiply_target_shape():
    """N = 2^16, 24 data primes of 60 bits, t = 65537 (one pair)."""
    logn, L, t = 16, 24, 65537
    ctx, moduli = make_ctx(BFV, logn, [60] * 25, t)
    rng = np.random.default_rng(8)
    a = ck.uniform_residues(rng, (1, 2, L, 1 << logn), moduli[:L])
    b = ck.uniform_residues(rng, (1, 2, L, 1 << logn), moduli[:L])
    da, db, do = to_dev(a), to_dev(b), empty_dev((1, 3, L, 1 << logn))
    ws_bytes = ctx.multiply_workspace_size(1, L)
    ws = empty_dev((ws_bytes + 7) // 8)
    ctx.multiply(da.data_ptr(), db.data_ptr(), do.data_ptr(), 1, L, ws.data_ptr(), ws_bytes, stream())
    np.testing.assert_array_equal(from_dev(do)[0], ck.bfv_multiply(a[0], b[0], moduli[:L], t, logn))


def test_keyswitch_rejects_small_workspace():
    ctx, moduli = make_ctx(CKKS, 10, [60, 40, 60])
    d = to_dev(np.zeros((1, 3, 2, 1024), dtype=np.uint64))
    k = to_dev(np.zeros((2, 2, 3, 1024), dtype=np.uint64))
    ws = empty_dev(16)
    with pytest.raises(hip().SealHipError):
        ctx.relinearize(d.data_ptr(), k.data_ptr(), 1, 2, ws.data_ptr(), 128, stream())


def test_empty_batches():
    """count = 0 is a valid batch for every batched entry point: nothing runs, no error."""
    ctx, moduli = make_ctx(CKKS, 12, [60, 40, 60])
    bctx, _ = make_ctx(BFV, 12, [60, 60, 60], 65537)
    ws = empty_dev((1 << 16,))
    d = empty_dev((8,))
    ctx.ckks_multiply(d.data_ptr(), d.data_ptr(), d.data_ptr(), 0, 2, stream())
    ctx.relinearize(d.data_ptr(), d.data_ptr(), 0, 2, ws.data_ptr(), 8 << 16, stream())
    ctx.rescale_to_next(d.data_ptr(), d.data_ptr(), 0, 2, 2, ws.data_ptr(), 8 << 16, stream())
    bctx.multiply(d.data_ptr(), d.data_ptr(), d.data_ptr(), 0, 2, ws.data_ptr(), 8 << 16, stream())
    torch.cuda.synchronize()
