"""GPU parity of Galois automorphisms and rotation key switching against the oracle.

Mirrors the reference's rotation tests (test/evaluator_cuda.cu:4180-4497, 5377-5511)
at the word level: injected ciphertexts and Galois keys, bit-exact outputs.
"""
import numpy as np
import pytest

import checkers as ck
from checkers import ptr
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


@pytest.mark.parametrize("logn", [3, 10, 13, 16])
@pytest.mark.parametrize("ntt_form", [True, False])
def test_apply_galois(logn, ntt_form):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, [50, 40, 50])
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(logn)
    L = 2
    a = ck.uniform_residues(rng, (3, L, n), moduli[:L])
    d = to_dev(a)
    o = empty_dev((3, L, n))
    for step in (1, -1, 0, 3, (n >> 1) - 1 if n > 4 else 1):
        elt = ctx.galois_elt_from_step(step)
        assert elt == ck.galois_elt(logn, step)
        ctx.apply_galois(d.data_ptr(), o.data_ptr(), 3, L, elt, ntt_form, stream())
        got = from_dev(o)
        want = np.zeros_like(a)
        for p in range(3):
            for l in range(L):
                if ntt_form:
                    ck.oracle().or_apply_galois_ntt(logn, elt, ptr(np.ascontiguousarray(a[p, l])),
                                                   ptr(want[p, l]))
                else:
                    ck.oracle().or_apply_galois(logn, elt, moduli[l], ptr(np.ascontiguousarray(a[p, l])),
                                               ptr(want[p, l]))
        np.testing.assert_array_equal(got, want)


def test_galois_kat_on_gpu():
    # test/utils/galois.cpp:74-109 (N = 8, q = 17, elt 3)
    ctx = hip().Context(CKKS, 8, [17, 97])
    d = to_dev(np.arange(8, dtype=np.uint64))
    o = empty_dev((8,))
    ctx.apply_galois(d.data_ptr(), o.data_ptr(), 1, 1, 3, False, stream())
    assert [int(x) for x in from_dev(o)] == [0, 14, 6, 1, 13, 7, 2, 12]
    ctx.apply_galois(d.data_ptr(), o.data_ptr(), 1, 1, 3, True, stream())
    assert [int(x) for x in from_dev(o)] == [4, 5, 7, 6, 1, 0, 2, 3]


@pytest.mark.parametrize("scheme,logn,bits,L,count,steps", [
    (CKKS, 10, [60, 40, 40, 60], 3, 2, [1, -1, 0]),
    (CKKS, 16, [60] + [50] * 23 + [60], 24, 1, [4]),      # hoisted-rotation workload's shape
    (CKKS, 15, [60] * 6, 5, 7, [3, 0]),   # fused products over groups of 4, 2, 1 reading sigma(c0) in place
    (CKKS, 13, [60] * 4, 3, 2, [2]),      # unfused key products: c1 zeroed first
    (BFV, 12, [43, 43, 44, 44, 44], 4, 3, [1, -7]),
    (BFV, 13, [43, 43, 44, 44, 44], 4, 1, [0]),          # row swap (conjugation element)
])
def test_rotation_keyswitch(scheme, logn, bits, L, count, steps):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(logn * 7 + L)
    ws_bytes = ctx.galois_workspace_size(count, L)
    ws = empty_dev((ws_bytes + 7) // 8)
    for step in steps:
        elt = ctx.galois_elt_from_step(step)
        ct = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
        key = ck.random_key(rng, moduli, logn)
        d, dk = to_dev(ct), to_dev(key)
        ctx.apply_galois_keyswitch(d.data_ptr(), dk.data_ptr(), elt, count, L, ws.data_ptr(), ws_bytes, stream())
        got = from_dev(d)
        for c in range(count):
            np.testing.assert_array_equal(got[c], ck.apply_galois_keyswitch(scheme, ct[c], elt, key, moduli, L, logn))
