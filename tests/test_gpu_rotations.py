"""Hoisted rotations on the GPU (sums and separate outputs).

shc_rotate_sum and shc_rotate_many (one ModUp of c1 for every rotation, keys prepared with
the sign-pattern correction) must give the words of summing (or listing) separate rotations -- the oracle's
applyGaloisInplace restatement, itself pinned against the reference -- for CKKS (NTT
form) and BFV (coefficient form), for positive, negative and conjugation elements, and
when c1 has zero digits (fixed on the device), is entirely zero, or has many zeros; and a
captured HIP graph of the call replays to the same words.
"""
import numpy as np
import torch
import pytest

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def run_case(scheme, logn, bits, L, count, steps, zero_digits=False, seed=0):
    """zero_digits: False; True (two zeroed coefficients of ciphertext 0's c1); or
    "heavy" (ciphertext 0's c1 entirely zero, every 16th coefficient of ciphertext 1's c1
    zero)."""
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    K = len(moduli)
    ctx = hip().Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(seed)
    cts = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    if zero_digits == "heavy":
        cts[0, 1] = 0
        c1 = cts[1, 1]
        if scheme == CKKS:
            c1 = ck.ntt_rns(c1, moduli[:L], logn, inverse=True)
        c1[:, ::16] = 0
        cts[1, 1] = ck.ntt_rns(c1, moduli[:L], logn) if scheme == CKKS else c1
    elif zero_digits:
        # zero some coefficient-form digits of c1 (CKKS holds c1 in NTT form)
        c1 = cts[0, 1]
        if scheme == CKKS:
            c1 = ck.ntt_rns(c1, moduli[:L], logn, inverse=True)
        c1[:, 5] = 0
        c1[1, 17] = 0
        cts[0, 1] = ck.ntt_rns(c1, moduli[:L], logn) if scheme == CKKS else c1
    elts = [ck.galois_elt(logn, s) for s in steps]
    keys = [ck.random_key(rng, moduli, logn) for _ in steps]
    dkeys = [to_dev(k) for k in keys]
    hb = ctx.hoisted_key_size(L)
    hw = ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)
    hkeys = []
    for e, dk in zip(elts, dkeys):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(dk.data_ptr(), e, L, h.data_ptr(), hws.data_ptr(), hw, stream())
        hkeys.append(h)
    din = to_dev(cts)
    out = empty_dev((count, 2, L, n))
    many = empty_dev((len(elts), count, 2, L, n))
    rw = ctx.rotate_sum_workspace_size(count, L)
    rws = ws(rw)
    hp, kp = [h.data_ptr() for h in hkeys], [k.data_ptr() for k in dkeys]
    ctx.rotate_sum(din.data_ptr(), count, L, elts, hp, kp, out.data_ptr(), rws.data_ptr(), rw, stream())
    ctx.rotate_many(din.data_ptr(), count, L, elts, hp, kp, many.data_ptr(), rws.data_ptr(), rw, stream())
    got, got_many = from_dev(out), from_dev(many)
    for c in range(count):
        want = None
        for r, (e, k) in enumerate(zip(elts, keys)):
            rot = ck.apply_galois_keyswitch(scheme, cts[c], e, k, moduli, L, logn)
            np.testing.assert_array_equal(got_many[r, c], rot, err_msg=f"rotation {r} of ciphertext {c}")
            want = rot if want is None else ck.mod_add(want, rot, moduli)
        np.testing.assert_array_equal(got[c], want, err_msg=f"ciphertext {c}")
    # the input is untouched
    np.testing.assert_array_equal(from_dev(din), cts)


@pytest.mark.parametrize("scheme,logn,bits,L,count,steps", [
    (CKKS, 10, [60, 40, 40, 60], 3, 2, [1, 2, -1]),
    (CKKS, 13, [60, 40, 40, 40, 60], 4, 3, [1, 4, -3, 0]),
    (BFV, 12, [40, 40, 50], 2, 2, [1, -2, 0]),
    (BFV, 13, [43, 43, 44, 44, 44], 4, 1, [3, 7]),
    (CKKS, 15, [60] + [40] * 6 + [60], 5, 2, [1, 2, 4, 8]),   # a level below the first
])
def test_rotate_sum_equals_separate_rotations(scheme, logn, bits, L, count, steps):
    run_case(scheme, logn, bits, L, count, steps, seed=logn + L)


@pytest.mark.parametrize("scheme", [CKKS, BFV])
def test_rotate_sum_zero_digits_exact(scheme):
    """A zero digit at a negated position has the reference digit 0: the device lists the
    zeros and puts the products right, so the words still equal separate rotations."""
    run_case(scheme, 12, [40, 40, 50], 2, 2, [1, 2], zero_digits=True, seed=5)


@pytest.mark.parametrize("scheme", [CKKS, BFV])
def test_rotate_sum_many_zero_digits_and_zero_c1(scheme):
    run_case(scheme, 12, [40, 40, 50], 2, 2, [1, 2, -1, 3, 5], zero_digits="heavy", seed=6)


def test_rotate_sum_in_a_captured_graph():
    """The hoisted rotation sum waits for nothing on the host, so it can be captured: a
    captured call (with zero digits in the batch, so the fix runs) replays to the eager
    words three times, with eager work between the replays."""
    logn, n = 12, 1 << 12
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 60])
    L, count = 3, 2
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(41)
    cts = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
    c1 = ck.ntt_rns(cts[1, 1], moduli[:L], logn, inverse=True)
    c1[:, 7] = 0
    cts[1, 1] = ck.ntt_rns(c1, moduli[:L], logn)
    steps = [1, 2, 5]
    elts = [ck.galois_elt(logn, s) for s in steps]
    keys = [ck.random_key(rng, moduli, logn) for _ in steps]
    dkeys = [to_dev(k) for k in keys]
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = ws(hw)
    hkeys = []
    for e, dk in zip(elts, dkeys):
        h = empty_dev(((hb + 7) // 8,))
        ctx.hoisted_key_prepare(dk.data_ptr(), e, L, h.data_ptr(), hws.data_ptr(), hw, stream())
        hkeys.append(h)
    din = to_dev(cts)
    out = empty_dev((count, 2, L, n))
    rw = ctx.rotate_sum_workspace_size(count, L)
    rws = ws(rw)
    hp, kp = [h.data_ptr() for h in hkeys], [k.data_ptr() for k in dkeys]
    ctx.rotate_sum(din.data_ptr(), count, L, elts, hp, kp, out.data_ptr(), rws.data_ptr(), rw, stream())
    eager = from_dev(out).copy()
    for c in range(count):
        want = None
        for e, k in zip(elts, keys):
            rot = ck.apply_galois_keyswitch(CKKS, cts[c], e, k, moduli, L, logn)
            want = rot if want is None else ck.mod_add(want, rot, moduli)
        np.testing.assert_array_equal(eager[c], want, err_msg=f"eager, ciphertext {c}")
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(graph, stream=cap):
        ctx.rotate_sum(din.data_ptr(), count, L, elts, hp, kp, out.data_ptr(), rws.data_ptr(), rw,
                       torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    other = empty_dev((count, 2, L, n))
    for _ in range(3):
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(out), eager, err_msg="replay")
        # eager work between replays, on the same workspace
        ctx.rotate_sum(din.data_ptr(), count, L, elts, hp, kp, other.data_ptr(), rws.data_ptr(), rw, stream())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(other), eager, err_msg="eager between replays")


def test_hoisted_key_bound_to_level_and_element():
    """A prepared key names its level and Galois element; using it at another level or
    for another element is an invalid argument, not wrong words."""
    logn, n = 12, 1 << 12
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 60])
    K = len(moduli)
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(3)
    key = to_dev(ck.random_key(rng, moduli, logn))
    e1, e2 = ck.galois_elt(logn, 1), ck.galois_elt(logn, 2)
    L = 3
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    h = empty_dev(((hb + 7) // 8,))
    ctx.hoisted_key_prepare(key.data_ptr(), e1, L, h.data_ptr(), ws(hw).data_ptr(), hw, stream())
    for lvl, elt in ((L, e2), (L - 1, e1)):
        din = to_dev(ck.uniform_residues(rng, (1, 2, lvl, n), moduli[:lvl]))
        out = empty_dev((1, 2, lvl, n))
        rw = ctx.rotate_sum_workspace_size(1, lvl)
        with pytest.raises(hip().SealHipError, match="prepared for another level"):
            ctx.rotate_sum(din.data_ptr(), 1, lvl, [elt], [h.data_ptr()], [key.data_ptr()], out.data_ptr(),
                           ws(rw).data_ptr(), rw, stream())
    # the matching level and element still work
    din = to_dev(ck.uniform_residues(rng, (1, 2, L, n), moduli[:L]))
    out = empty_dev((1, 2, L, n))
    rw = ctx.rotate_sum_workspace_size(1, L)
    ctx.rotate_sum(din.data_ptr(), 1, L, [e1], [h.data_ptr()], [key.data_ptr()], out.data_ptr(), ws(rw).data_ptr(),
                   rw, stream())
    want = ck.apply_galois_keyswitch(CKKS, from_dev(din)[0], e1, from_dev(key), moduli, L, logn)
    np.testing.assert_array_equal(from_dev(out)[0], want)


def test_rotate_sum_seven_rotations_two_passes():
    """Seven rotations: two passes of the hoisted products (4 + 3) accumulate into the
    output and reuse the product buffers."""
    run_case(CKKS, 12, [60, 40, 40, 60], 3, 3, [1, 2, 3, 5, -1, -4, 7], seed=77)
