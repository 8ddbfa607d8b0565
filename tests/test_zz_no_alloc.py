"""shc_relinearize_noalloc: relinearization with the key-switch scratch in the library's
fixed device array (no workspace argument, no allocation).

Through the C ABI, its output words must equal shc_relinearize's (the entry point with a
caller workspace, itself checked against the oracle elsewhere) on two calls in a row,
the second with a smaller input, for CKKS (NTT form) and BFV (coefficient form); one
result is also checked against the oracle.  A shape whose key switch needs more than the
fixed 128 MiB is rejected with SHC_OUT_OF_RANGE before any GPU work.
"""
import numpy as np
import pytest

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def _with_workspace(ctx, ct3, key, count, L):
    d = to_dev(ct3)
    w = ctx.keyswitch_workspace_size(count, L)
    ws = empty_dev(((w + 7) // 8,))
    ctx.relinearize(d.data_ptr(), key.data_ptr(), count, L, ws.data_ptr(), w, stream())
    return from_dev(d)[:, :2]


def _noalloc(ctx, ct3, key, count, L):
    d = to_dev(ct3)
    ctx.relinearize_noalloc(d.data_ptr(), key.data_ptr(), count, L, stream())
    return from_dev(d)[:, :2]


@pytest.mark.parametrize("scheme,n,bits,plain", [(CKKS, 8192, [60, 40, 40, 40, 60], 0),
                                                  (BFV, 8192, [43, 43, 44, 44, 44], 65537)])
def test_noalloc_relinearize_matches_workspace_entry(scheme, n, bits, plain):
    sh = hip()
    logn = n.bit_length() - 1
    moduli = ck.coeff_modulus_create(n, bits)
    K = len(moduli)
    L = K - 1
    ctx = sh.Context(scheme, n, moduli, plain)
    rng = np.random.default_rng(23 + scheme)
    key_h = ck.random_key(rng, moduli, logn)
    key = to_dev(key_h)
    # first call: 6 ciphertexts at the top level; second: 1 ciphertext one level down
    for count, lv in ((6, L), (1, L - 1)):
        ct3 = ck.uniform_residues(rng, (count, 3, lv, n), moduli[:lv])
        want = _with_workspace(ctx, ct3, key, count, lv)
        got = _noalloc(ctx, ct3, key, count, lv)
        assert np.array_equal(got, want), f"count={count} L={lv}: noalloc differs from the workspace entry point"
    assert np.array_equal(got[0], ck.relinearize(scheme, ct3[0], key_h, moduli, L - 1, logn))


def test_noalloc_rejects_shapes_beyond_the_scratch():
    sh = hip()
    n = 1 << 16
    moduli = ck.coeff_modulus_create(n, [60] * 25)
    ctx = sh.Context(CKKS, n, moduli)
    assert ctx.keyswitch_workspace_size(1, 24) > 128 << 20
    with pytest.raises(sh.SealHipError) as e:
        ctx.relinearize_noalloc(0, 0, 1, 24, stream())
    assert e.value.status == -3 and "scratch" in str(e.value)
