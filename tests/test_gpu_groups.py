"""Level groups in one launch per stage (shc_*_groups): ciphertexts at uneven levels.

Each group's words must equal the one-level calls on that group (shc_ckks_multiply,
shc_relinearize, shc_rescale_to_next), which the oracle pins elsewhere; one ciphertext
per case is also checked against the oracle.  Covered: the fused CKKS path at N = 2^15
(config 3's primes, levels 2..15, groups of 1..7 so that groups of 4 / 2 / 1 all run)
and N = 2^16, more than 16 groups (taken 16 at a time), the group-by-group paths (BFV,
N = 2^13), rescaling straight from size-3 ciphertexts, and a captured graph of a
grouped multiply -> relinearize -> rescale replayed with eager work between replays.
"""
import numpy as np
import pytest
import torch

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


def make_case(scheme, logn, bits, spec, seed):
    n = 1 << logn
    moduli = ck.coeff_modulus_create(n, bits)
    ctx = hip().Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
    rng = np.random.default_rng(seed)
    key_h = ck.random_key(rng, moduli, logn)
    gs = []
    for L, count in spec:
        a = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
        b = ck.uniform_residues(rng, (count, 2, L, n), moduli[:L])
        ct3 = ck.uniform_residues(rng, (count, 3, L, n), moduli[:L])
        gs.append((L, count, a, b, ct3))
    return ctx, moduli, key_h, to_dev(key_h), gs, n


def relin_one(ctx, ct3, key, count, L):
    d = to_dev(ct3)
    w = ctx.keyswitch_workspace_size(count, L)
    b = ws(w)
    ctx.relinearize(d.data_ptr(), key.data_ptr(), count, L, b.data_ptr(), w, stream())
    return from_dev(d)


def rescale_one(ctx, ct2, count, L):
    d = to_dev(ct2)
    out = empty_dev((count, 2, L - 1, ct2.shape[-1]))
    w = ctx.rescale_workspace_size(count, 2, L)
    b = ws(w)
    ctx.rescale_to_next(d.data_ptr(), out.data_ptr(), count, 2, L, b.data_ptr(), w, stream())
    return from_dev(out)


@pytest.mark.parametrize("scheme,logn,bits,spec", [
    (CKKS, 15, [60] + [40] * 14 + [60], [(15, 2), (9, 5), (4, 3), (7, 1), (12, 7), (2, 2)]),
    (CKKS, 16, [60] * 6, [(5, 3), (3, 2), (4, 1)]),
    (CKKS, 15, [60] + [40] * 6 + [60], [(2 + k % 6, 1 + k % 3) for k in range(18)]),  # 18 groups: 16 + 2
    (CKKS, 13, [60, 40, 40, 40, 60], [(4, 2), (2, 3)]),                               # group by group
    (BFV, 15, [60] * 5, [(4, 2), (2, 3)]),                                             # group by group
])
def test_relinearize_groups_equal_one_level_calls(scheme, logn, bits, spec):
    ctx, moduli, key_h, key, gs, n = make_case(scheme, logn, bits, spec, seed=logn + len(spec))
    devs = [to_dev(ct3) for (_, _, _, _, ct3) in gs]
    items = [(d.data_ptr(), c, L) for d, (L, c, _, _, _) in zip(devs, gs)]
    w = ctx.relinearize_groups_workspace_size(items)
    b = ws(w)
    ctx.relinearize_groups(items, key.data_ptr(), b.data_ptr(), w, stream())
    for d, (L, c, _, _, ct3) in zip(devs, gs):
        got = from_dev(d)[:, :2]
        want = relin_one(ctx, ct3, key, c, L)[:, :2]
        np.testing.assert_array_equal(got, want, err_msg=f"level {L}, {c} ciphertexts")
    L, c, _, _, ct3 = gs[0]
    np.testing.assert_array_equal(from_dev(devs[0])[c - 1, :2],
                                  ck.relinearize(scheme, ct3[c - 1], key_h, moduli, L, int(np.log2(n))))


@pytest.mark.parametrize("logn,bits,spec", [
    (15, [60] + [40] * 14 + [60], [(15, 2), (9, 5), (4, 3), (7, 1), (12, 7), (2, 2)]),
    (16, [60] * 6, [(5, 3), (3, 2), (4, 1)]),
])
def test_multiply_relinearize_rescale_groups(logn, bits, spec):
    """Tensor products of groups, relinearization, and a rescale read straight from the
    size-3 products, against the one-level calls (and the oracle for one ciphertext)."""
    ctx, moduli, key_h, key, gs, n = make_case(CKKS, logn, bits, spec, seed=3 * logn)
    da = [to_dev(a) for (_, _, a, _, _) in gs]
    db = [to_dev(b) for (_, _, _, b, _) in gs]
    prods = [empty_dev((c, 3, L, n)) for (L, c, _, _, _) in gs]
    outs = [empty_dev((c, 2, L - 1, n)) for (L, c, _, _, _) in gs]
    it = lambda ts: [(t.data_ptr(), c, L) for t, (L, c, _, _, _) in zip(ts, gs)]
    ctx.ckks_multiply_groups(it(da), it(db), it(prods), stream())
    w = ctx.relinearize_groups_workspace_size(it(prods))
    b = ws(w)
    ctx.relinearize_groups(it(prods), key.data_ptr(), b.data_ptr(), w, stream())
    rin = it(prods)
    rout = [(t.data_ptr(), c, L - 1) for t, (L, c, _, _, _) in zip(outs, gs)]
    rw = ctx.rescale_groups_workspace_size(rin)
    rb = ws(rw)
    ctx.rescale_to_next_groups(rin, 3, rout, rb.data_ptr(), rw, stream())
    for k, (L, c, a, bb, _) in enumerate(gs):
        p = empty_dev((c, 3, L, n))
        ctx.ckks_multiply(da[k].data_ptr(), db[k].data_ptr(), p.data_ptr(), c, L, stream())
        relin = relin_one(ctx, from_dev(p), key, c, L)[:, :2]
        np.testing.assert_array_equal(from_dev(prods[k])[:, :2], relin, err_msg=f"relin, level {L}")
        np.testing.assert_array_equal(from_dev(outs[k]), rescale_one(ctx, np.ascontiguousarray(relin), c, L),
                                      err_msg=f"rescale, level {L}")
    L, c, a, bb, _ = gs[1]
    o3 = ck.ckks_multiply(a[0], bb[0], moduli[:L], logn)
    o2 = ck.relinearize(CKKS, o3, key_h, moduli, L, logn)
    np.testing.assert_array_equal(from_dev(outs[1])[0], ck.rescale(CKKS, o2, moduli, logn))


def test_grouped_chain_in_a_captured_graph():
    logn, n = 15, 1 << 15
    spec = [(15, 2), (9, 3), (5, 1)]
    ctx, moduli, key_h, key, gs, n = make_case(CKKS, logn, [60] + [40] * 14 + [60], spec, seed=8)
    da = [to_dev(a) for (_, _, a, _, _) in gs]
    db = [to_dev(b) for (_, _, _, b, _) in gs]
    prods = [empty_dev((c, 3, L, n)) for (L, c, _, _, _) in gs]
    outs = [empty_dev((c, 2, L - 1, n)) for (L, c, _, _, _) in gs]
    it = lambda ts, d=0: [(t.data_ptr(), c, L - d) for t, (L, c, _, _, _) in zip(ts, gs)]
    w = max(ctx.relinearize_groups_workspace_size(it(prods)), ctx.rescale_groups_workspace_size(it(prods)))
    b = ws(w)

    def step(st):
        ctx.ckks_multiply_groups(it(da), it(db), it(prods), st)
        ctx.relinearize_groups(it(prods), key.data_ptr(), b.data_ptr(), w, st)
        ctx.rescale_to_next_groups(it(prods), 3, it(outs, 1), b.data_ptr(), w, st)

    step(stream())
    eager = [from_dev(o).copy() for o in outs]
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(graph, stream=cap):
        step(torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    for _ in range(3):
        for o in outs:
            o.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            np.testing.assert_array_equal(from_dev(o), e)
        step(stream())
        torch.cuda.synchronize()
