"""Key switching, tensor product and rescale captured in a HIP graph.

A CKKS chain over ciphertexts of uneven levels (one batched call per level group:
multiply -> relinearize -> pack -> rescale, twice) is captured once and replayed three
times with eager work and synchronisations between the replays; every replay must give
the eager chain's words, and the eager words equal the oracle's.
"""
import numpy as np
import pytest
import torch

import checkers as ck
from gpu_util import from_dev, hip, to_dev

pytestmark = pytest.mark.gpu
CKKS = 2


def test_graph_replays_equal_eager_chain():
    logn, n = 13, 1 << 13
    moduli = ck.coeff_modulus_create(n, [60, 40, 40, 40, 40, 60])
    K = len(moduli)
    ctx = hip().Context(CKKS, n, moduli)
    rng = np.random.default_rng(21)
    key_h = ck.random_key(rng, moduli, logn)
    key = to_dev(key_h)
    groups = {3: 2, 4: 1, 5: 3}
    bufs = {}
    max_ws = 0
    for l, c in groups.items():
        x = ck.uniform_residues(rng, (c, 2, l, n), moduli[:l])
        ys = [ck.uniform_residues(rng, (c, 2, l - s, n), moduli[:l - s]) for s in range(2)]
        dev = dict(x=to_dev(x), ys=[to_dev(y) for y in ys],
                   prods=[torch.empty((c, 3, l - s, n), dtype=torch.int64, device="cuda") for s in range(2)],
                   rels=[torch.empty((c, 2, l - s, n), dtype=torch.int64, device="cuda") for s in range(2)],
                   accs=[torch.empty((c, 2, l - s - 1, n), dtype=torch.int64, device="cuda") for s in range(2)])
        bufs[l] = (c, x, ys, dev)
        for s in range(2):
            max_ws = max(max_ws, ctx.keyswitch_workspace_size(c, l - s), ctx.rescale_workspace_size(c, 2, l - s))
    ws = torch.empty(((max_ws + 7) // 8,), dtype=torch.int64, device="cuda")

    def chain():
        st = torch.cuda.current_stream().cuda_stream
        for l, (c, _, _, d) in bufs.items():
            cur = d["x"]
            for s in range(2):
                ll = l - s
                ctx.ckks_multiply(cur.data_ptr(), d["ys"][s].data_ptr(), d["prods"][s].data_ptr(), c, ll, st)
                ctx.relinearize(d["prods"][s].data_ptr(), key.data_ptr(), c, ll, ws.data_ptr(), max_ws, st)
                d["rels"][s].copy_(d["prods"][s][:, :2])
                ctx.rescale_to_next(d["rels"][s].data_ptr(), d["accs"][s].data_ptr(), c, 2, ll, ws.data_ptr(),
                                    max_ws, st)
                cur = d["accs"][s]

    chain()
    torch.cuda.synchronize()
    eager = {l: from_dev(d["accs"][1]).copy() for l, (_, _, _, d) in bufs.items()}
    # the eager words are the oracle's
    for l, (c, x, ys, _) in bufs.items():
        for i in range(c):
            cur = x[i]
            for s in range(2):
                ll = l - s
                o3 = ck.ckks_multiply(cur, ys[s][i], moduli[:ll], logn)
                o2 = ck.relinearize(CKKS, o3, key_h, moduli, ll, logn)
                cur = ck.rescale(CKKS, o2, moduli, logn)
            np.testing.assert_array_equal(eager[l][i], cur)

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
            np.testing.assert_array_equal(from_dev(d["accs"][1]), eager[l])
        chain()   # eager work between replays
        torch.cuda.synchronize()
