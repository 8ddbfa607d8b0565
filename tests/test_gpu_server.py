"""The continuous-batching request server (shc_server_*, seal-cuda_amd/csrc/server.cpp).

Requests of three types -- a BFV and a CKKS diagonal matrix-vector product (baby-step
giant-step rotations, as app/matmul.cuh:98-121 computes one diagonal at a time) and a
CKKS multiply + relinearize + rescale -- are submitted as one burst to a server with two
lanes, so batches of different types run on the device at once.  Each request's output
words must equal (a) the same request served alone by a one-lane, batch-of-one server,
(b) the library's direct calls (matvec.BsgsMatvec.run over the C ABI; multiply,
relinearize, rescale), and for the multiply type (c) the oracle's words.
"""
import numpy as np
import pytest

import checkers as ck
from gpu_util import empty_dev, from_dev, hip, stream, to_dev

pytestmark = pytest.mark.gpu
BFV, CKKS = 1, 2


def ws(nbytes):
    return empty_dev((max(nbytes, 8) + 7) // 8)


class MatvecType:
    """A matvec request type over synthetic residues (keys and diagonals uniform mod q:
    the words, not decryption, are what the server must preserve)."""

    def __init__(self, scheme, n, moduli, plain, b, g, seed):
        import matvec
        sh = hip()
        rng = np.random.default_rng(seed)
        K = len(moduli)
        self.L, self.n, self.b, self.g = K - 1, n, b, g
        L = self.L
        self.moduli = moduli
        self.ctx = sh.Context(scheme, n, moduli, plain)
        ctx = self.ctx
        logn = n.bit_length() - 1
        self.plains = to_dev(ck.uniform_residues(rng, (g, b, L, n), moduli[:L]))
        self.baby = [to_dev(ck.random_key(rng, moduli, logn)) for _ in range(1, b)]
        self.giant = [to_dev(ck.random_key(rng, moduli, logn)) for _ in range(1, g)]
        hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
        hws = ws(hw)
        self.hoisted = []
        for j, k in zip(range(1, b), self.baby):
            h = empty_dev(((hb + 7) // 8,))
            ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(j), L, h.data_ptr(), hws.data_ptr(), hw,
                                    stream())
            self.hoisted.append(h)
        self.mv = matvec.BsgsMatvec(ctx, n, L, b, g, self.plains.data_ptr(), [k.data_ptr() for k in self.baby],
                                    [h.data_ptr() for h in self.hoisted], [k.data_ptr() for k in self.giant],
                                    bfv=scheme == BFV)

    def register(self, server):
        return server.add_matvec(self.ctx, self.L, self.b, self.g, self.plains.data_ptr(),
                                 [k.data_ptr() for k in self.baby], [h.data_ptr() for h in self.hoisted],
                                 [k.data_ptr() for k in self.giant])

    def inputs(self, rng, count):
        return [ck.uniform_residues(rng, (2, self.L, self.n), self.moduli[:self.L]) for _ in range(count)]

    def direct(self, x):
        """The product through matvec.BsgsMatvec.run (library calls on the test's stream)."""
        L, n, b, g = self.L, self.n, self.b, self.g
        d_in = to_dev(x[None])
        out = empty_dev((1, 2, L, n))
        baby = empty_dev((b, 1, 2, L, n))
        inner = empty_dev((g, 1, 2, L, n))
        wsz = self.mv.workspace_sizes(1)
        w = ws(wsz)
        self.mv.run(d_in.data_ptr(), 1, out.data_ptr(), baby.data_ptr(), inner.data_ptr(), w.data_ptr(), wsz, stream())
        return from_dev(out)[0]


class MulRelinType:
    """multiply + relinearize (+ rescale / modulus switch) requests; BFV inputs are
    coefficient form (BEHZ multiply), CKKS NTT form."""

    def __init__(self, n, moduli, seed, scheme=CKKS, rescale=True):
        sh = hip()
        rng = np.random.default_rng(seed)
        K = len(moduli)
        self.L, self.n, self.moduli = K - 1, n, moduli
        self.logn = n.bit_length() - 1
        self.scheme, self.rescale = scheme, rescale
        self.ctx = sh.Context(scheme, n, moduli, 65537 if scheme == BFV else 0)
        self.key_h = ck.random_key(rng, moduli, self.logn)
        self.key = to_dev(self.key_h)

    def register(self, server):
        return server.add_mulrelin(self.ctx, self.L, self.key.data_ptr(), rescale=self.rescale)

    def inputs(self, rng, count):
        return [ck.uniform_residues(rng, (2, 2, self.L, self.n), self.moduli[:self.L]) for _ in range(count)]

    def direct(self, x):
        ctx, L, n = self.ctx, self.L, self.n
        a, b = to_dev(x[0][None]), to_dev(x[1][None])
        prod = empty_dev((1, 3, L, n))
        mw = ctx.multiply_workspace_size(1, L)
        mb = ws(mw)
        ctx.multiply(a.data_ptr(), b.data_ptr(), prod.data_ptr(), 1, L, mb.data_ptr(), mw, stream())
        kw = ctx.keyswitch_workspace_size(1, L)
        kb = ws(kw)
        ctx.relinearize(prod.data_ptr(), self.key.data_ptr(), 1, L, kb.data_ptr(), kw, stream())
        relin = prod[:, :2].contiguous()
        if not self.rescale:
            return from_dev(relin)[0]
        out = empty_dev((1, 2, L - 1, n))
        rw = ctx.rescale_workspace_size(1, 2, L)
        rb = ws(rw)
        ctx.rescale_to_next(relin.data_ptr(), out.data_ptr(), 1, 2, L, rb.data_ptr(), rw, stream())
        return from_dev(out)[0]

    def oracle(self, x):
        if self.scheme == BFV:
            o3 = ck.bfv_multiply(x[0], x[1], self.moduli[:self.L], 65537, self.logn)
        else:
            o3 = ck.ckks_multiply(x[0], x[1], self.moduli[:self.L], self.logn)
        o2 = ck.relinearize(self.scheme, o3, self.key_h, self.moduli, self.L, self.logn)
        return ck.rescale(self.scheme, o2, self.moduli, self.logn) if self.rescale else o2


def serve(server, reqs, out_words):
    """Submits every (type, input) at once, then waits for all; outputs and infos."""
    outs = [np.zeros(out_words[t], dtype=np.uint64) for t, _ in reqs]
    ids = [server.submit(t, np.ascontiguousarray(x, dtype=np.uint64).reshape(-1), o) for (t, x), o in zip(reqs, outs)]
    infos = [server.wait(i, 120.0) for i in ids]
    assert all(inf is not None for inf in infos), "a request did not finish within 120 s"
    return outs, infos


def test_server_mixed_stream_matches_alone_and_direct():
    sh = hip()
    bfv = MatvecType(BFV, 4096, ck.coeff_modulus_create(4096, [40, 40, 50]), 65537, 4, 4, seed=1)
    ckks = MatvecType(CKKS, 4096, ck.coeff_modulus_create(4096, [60, 40, 40, 60]), 0, 4, 4, seed=2)
    mr = MulRelinType(4096, ck.coeff_modulus_create(4096, [60, 40, 40, 40, 60]), seed=3)
    kinds = [bfv, ckks, mr]
    rng = np.random.default_rng(5)
    # a burst: 8 of each type interleaved, so each lane admits a batch of one type while
    # the other lane runs another
    per = [k.inputs(rng, 8) for k in kinds]
    order = [(t, i) for i in range(8) for t in range(3)]

    mixed = sh.Server(device_bytes=2 << 30, staging_bytes=256 << 20, lanes=2, max_batch=4)
    types = [k.register(mixed) for k in kinds]
    out_words = {types[t]: mixed.request_size(types[t])[1] for t in range(3)}
    outs, infos = serve(mixed, [(types[t], per[t][i]) for t, i in order], out_words)
    stats = mixed.stats()
    assert stats["completed"] == 24 and stats["failed"] == 0
    assert stats["max_batches_in_flight"] >= 2, "no two batches overlapped on the device"
    assert max(inf.batch_size for inf in infos) > 1, "the burst was never batched"
    assert len({inf.lane for inf in infos}) == 2
    mixed.close()

    alone = sh.Server(device_bytes=2 << 30, staging_bytes=256 << 20, lanes=1, max_batch=1)
    types1 = [k.register(alone) for k in kinds]
    for (t, i), o in zip(order, outs):
        got, inf = serve(alone, [(types1[t], per[t][i])], {types1[t]: out_words[types[t]]})
        assert inf[0].batch_size == 1
        assert np.array_equal(got[0], o), f"type {t} request {i}: words differ alone and in the mixed stream"
    alone.close()

    # the direct library calls (and, for mul+relin+rescale, the oracle) give the same words
    for t in range(3):
        for i in (0, 7):
            want = kinds[t].direct(per[t][i]).reshape(-1)
            got = outs[order.index((t, i))]
            assert np.array_equal(got, want), f"type {t} request {i}: server differs from the direct calls"
    assert np.array_equal(outs[order.index((2, 3))], mr.oracle(per[2][3]).reshape(-1))


def test_server_admission_by_device_memory():
    """With room for only one request of a type at a time, a burst is served one by one
    (never more device memory than the arena holds), and every result is still exact."""
    sh = hip()
    mr = MulRelinType(4096, ck.coeff_modulus_create(4096, [60, 40, 40, 60]), seed=9)
    probe = sh.Server(device_bytes=1 << 30, staging_bytes=64 << 20, lanes=1, max_batch=1)
    one = probe.request_size(mr.register(probe))[2]
    probe.close()
    server = sh.Server(device_bytes=one + one // 2, staging_bytes=64 << 20, lanes=2, max_batch=8)
    t = mr.register(server)
    rng = np.random.default_rng(10)
    xs = mr.inputs(rng, 6)
    outs, infos = serve(server, [(t, x) for x in xs], {t: server.request_size(t)[1]})
    st = server.stats()
    assert all(inf.batch_size == 1 for inf in infos)
    assert st["max_batches_in_flight"] == 1 and st["peak_device_bytes"] <= one + one // 2
    for x, o in zip(xs, outs):
        assert np.array_equal(o, mr.direct(x).reshape(-1))
    server.close()


def test_server_rejects_requests_that_cannot_fit():
    sh = hip()
    mr = MulRelinType(4096, ck.coeff_modulus_create(4096, [60, 40, 40, 60]), seed=11)
    server = sh.Server(device_bytes=1 << 20, staging_bytes=64 << 20, lanes=1, max_batch=1)
    with pytest.raises(sh.SealHipError, match="does not fit"):
        mr.register(server)
    server.close()


def test_server_pinned_buffers_in_place():
    """Inputs and outputs in the server's pinned memory are uploaded from and downloaded
    into in place; the words equal those of staged (ordinary host) buffers."""
    sh = hip()
    mr = MulRelinType(4096, ck.coeff_modulus_create(4096, [60, 40, 40, 60]), seed=12)
    server = sh.Server(device_bytes=1 << 30, staging_bytes=16 << 20, lanes=2, max_batch=4)
    t = mr.register(server)
    in_w, out_w, _ = server.request_size(t)
    rng = np.random.default_rng(13)
    xs = mr.inputs(rng, 6)
    pins, pouts, ids = [], [], []
    for x in xs:
        p = server.host_array(in_w)
        p[:] = x.reshape(-1)
        o = server.host_array(out_w)
        pins.append(p)
        pouts.append(o)
        ids.append(server.submit(t, p, o))
    for i in ids:
        assert server.wait(i, 60.0) is not None
    staged, _ = serve(server, [(t, x) for x in xs], {t: out_w})
    for o, st in zip(pouts, staged):
        assert np.array_equal(o, st)
    assert np.array_equal(pouts[0], mr.oracle(xs[0]).reshape(-1))
    server.close()


@pytest.mark.parametrize("scheme,rescale", [(BFV, True), (BFV, False), (CKKS, False)])
def test_server_mulrelin_types(scheme, rescale):
    """BFV (BEHZ multiply, modulus switch) and CKKS without rescale through the server:
    batches equal the direct calls and one request the oracle."""
    sh = hip()
    bits = [40, 40, 50] if scheme == BFV else [60, 40, 40, 60]
    mr = MulRelinType(4096, ck.coeff_modulus_create(4096, bits), seed=20 + scheme, scheme=scheme, rescale=rescale)
    server = sh.Server(device_bytes=1 << 30, staging_bytes=64 << 20, lanes=2, max_batch=4)
    t = mr.register(server)
    rng = np.random.default_rng(21)
    xs = mr.inputs(rng, 5)
    outs, infos = serve(server, [(t, x) for x in xs], {t: server.request_size(t)[1]})
    assert max(inf.batch_size for inf in infos) > 1
    for x, o in zip(xs, outs):
        assert np.array_equal(o, mr.direct(x).reshape(-1))
    assert np.array_equal(outs[2], mr.oracle(xs[2]).reshape(-1))
    server.close()
