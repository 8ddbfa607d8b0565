"""The key switch's launch-shape autotuner (shc_autotune_relinearize and the per-context
shape table).

Every candidate shape -- target primes per ModUp columns-pass block, and 4 / 2 / 1
ciphertexts per fused rows-pass block -- must give the built-in shape's words (and the
oracle's): shapes change the launch, never the arithmetic.  The tuner keeps the
fastest candidate, which is never slower than the built-in shape on its own run, and
the table survives an export / import round trip.
"""
import numpy as np
import pytest

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def relin(ctx, ct3, key, count, L):
    d = to_dev(ct3)
    w = ctx.keyswitch_workspace_size(count, L)
    b = ws(w)
    ctx.relinearize(d.data_ptr(), key.data_ptr(), count, L, b.data_ptr(), w, stream())
    return from_dev(d)[:, :2]


@pytest.mark.parametrize("scheme,logn,bits,count", [
    (CKKS, 15, [60] + [40] * 6 + [60], 7),     # fused path: groups 4 + 2 + 1, 2 x 3 + 1, 7 singles
    (BFV, 15, [60] * 5, 5),
    (CKKS, 13, [60, 40, 40, 40, 60], 3),       # unfused path: only the columns-pass choice
])
def test_every_candidate_shape_gives_the_same_words(scheme, logn, bits, count):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(logn * 7 + count)
    ct3 = ck.uniform_residues(rng, (count, 3, L, n), moduli[:L])
    key_h = ck.random_key(rng, moduli, logn)
    key = to_dev(key_h)
    base = relin(ctx, ct3, key, count, L)
    np.testing.assert_array_equal(base[count - 1], ck.relinearize(scheme, ct3[count - 1], key_h, moduli, L, logn))
    chunks = (1, 2, 3, 4, 6, 8, 12, 16, L + 1)
    targets = sorted({(L + c) // c for c in chunks if c <= L + 1})
    groups = (4, 2, 1) if logn in (15, 16) else (0,)
    for g in groups:
        for tc in targets:
            assert ctx.keyswitch_shape(L, count, (tc, g)) == (tc, g)
            np.testing.assert_array_equal(relin(ctx, ct3, key, count, L), base, err_msg=f"shape ({tc}, {g})")
    assert ctx.keyswitch_shape(L, count, clear=True) == (0, 0)


def test_autotune_keeps_the_fastest_and_round_trips():
    logn, n = 15, 1 << 15
    moduli = ck.coeff_modulus_create(n, [60] + [40] * 8 + [60])
    K = len(moduli)
    L = K - 1
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(3)
    count = 6
    key = to_dev(ck.random_key(rng, moduli, logn))
    d = to_dev(ck.uniform_residues(rng, (count, 3, L, n), moduli[:L]))
    w = ctx.keyswitch_workspace_size(count, L)
    b = ws(w)
    res = ctx.autotune_relinearize(d.data_ptr(), key.data_ptr(), count, L, b.data_ptr(), w, reps=2, stream=stream())
    assert res["best_ms"] <= res["builtin_ms"] and res["best_ms"] > 0
    assert ctx.keyswitch_shape(L, count) == res["best"]
    ctx.keyswitch_shape(L, 1, (3, 2))
    text = ctx.keyswitch_shapes_export()
    ctx2 = hip().Context(CKKS, n, moduli)
    ctx2.keyswitch_shapes_import(text)
    assert ctx2.keyswitch_shapes_export() == text
    assert ctx2.keyswitch_shape(L, 1) == (3, 2)
    with pytest.raises(hip().SealHipError):
        ctx2.keyswitch_shapes_import(f"{L} 4 {L + 2} 4\n")     # more targets than primes
    with pytest.raises(hip().SealHipError):
        ctx2.keyswitch_shape(L, 4, (1, 3))                      # no group of 3
    # the tuned shape still gives the built-in words
    ct3 = ck.uniform_residues(rng, (count, 3, L, n), moduli[:L])
    tuned = relin(ctx, ct3, key, count, L)
    ctx.keyswitch_shape(L, count, clear=True)
    np.testing.assert_array_equal(tuned, relin(ctx, ct3, key, count, L))
