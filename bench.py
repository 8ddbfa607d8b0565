#!/usr/bin/env python3
"""Benchmark of the MI355X RNS polynomial core.

Default workload (`--workload mulrelin`): CKKS multiply + relinearize at N = 2^16 with
25 primes of 60 bits (24 data primes + 1 special prime, CoeffModulus::Create(65536,
{60 x 25})), a batch of B = 128 ciphertext pairs in NTT form.  One step = tensor
product of all 128 pairs + relinearization of all 128 size-3 results, inputs and the
relinearization key resident in HBM.  The same work is also timed through the plain
per-ciphertext path (one call per ciphertext).

`--workload ntt`: batched negacyclic NTT, N = 2^16, 24 primes of 60 bits,
128 ciphertexts x 2 polys x 24 limbs, forward then inverse.

`--workload chain`: the CKKS chain (N=2^15, 64 pairs at uneven levels, 3 x mul -> relin ->
rescale) captured in a HIP graph.

`--workload rotsum`: hoisted rotations (N=2^16, 16 ciphertexts, 15 rotations each, summed)
against the same sum through separate rotations.

`--workload matvec`: CKKS 1024 x 1024 matrix-vector products by baby-step giant-step
rotations (N=2^15, batches of 16 requests).

`--workload copy`: a 1 GiB device copy, the calibration run for HBM byte counters.

Prints one JSON line.  Data are seeded uniform residues (synthetic) and injected
uniform key words; every word is fully reduced.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "seal-cuda_amd"))
sys.path.insert(0, os.path.join(HERE, "tests"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel at the default workload (batch 128),
# measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB device copy
# (profiles/mulrelin_pmc_traffic.md); a profiler cannot run inside this process.
KEYPROD_TRAFFIC_B128 = 40.93e9 + 3.36e9
# VALU wave instructions per launch of the same kernel (rocprofv3 --pmc SQ_INSTS_VALU, same
# file) and the VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles
# per SIMD at the 2.4 GHz peak engine clock.  The kernel is bound by VALU issue (exact
# 64-bit modular arithmetic from 32-bit multiplies), so this is its figure of merit.
KEYPROD_VALU_B128 = 8084848640
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4
CKKS = 2


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="mulrelin",
                   choices=["mulrelin", "ntt", "copy", "chain", "rotsum", "matvec", "serve"])
    p.add_argument("--no-bfv", action="store_true", help="skip the BFV leg of the mulrelin workload")
    p.add_argument("--batch", type=int, default=128)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-plain", action="store_true", help="skip the plain per-ciphertext leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-serve", action="store_true", help="skip the served-latency leg")
    p.add_argument("--serve-seconds", type=float, default=1.5, help="arrival span of the served-latency leg")
    return p.parse_args()


def fill_uniform(torch, t, moduli, limb_dim, gen):
    """Fill a [..., L, N] int64 device tensor with uniform residues of moduli[limb]."""
    for i, q in enumerate(moduli):
        sl = t.select(limb_dim, i)
        sl.copy_(torch.randint(0, q, sl.shape, generator=gen, device="cuda", dtype=torch.int64))


class Timer:
    def __init__(self, torch):
        self.torch = torch
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        self.a.record()
        return self

    def __exit__(self, *exc):
        self.b.record()
        self.torch.cuda.synchronize()
        self.ms = self.a.elapsed_time(self.b)


def cpu_baseline_mulrelin(moduli, logn, budget_s):
    """The reference CPU path (oracle/_ref, src/evaluator.cpp multiply + relinearize) when
    it was shipped with this tree, else the oracle restatement, on one host thread."""
    import numpy as np
    import checkers as ck
    n, K = 1 << logn, len(moduli)
    L = K - 1
    rng = np.random.default_rng(5)
    a = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    kind = "reference" if ck.ref() is not None else "port"
    sess = ck.RefSession(CKKS, n, moduli) if kind == "reference" else None
    ops = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s or ops == 0:
        if sess is not None:
            ct3 = sess.multiply(a, b, L)
            sess.relinearize(ct3, key, L)
        else:
            ct3 = ck.ckks_multiply(a, b, moduli[:L], logn)
            ck.relinearize(CKKS, ct3, key, moduli, L, logn)
        ops += 1
    dt = time.perf_counter() - t0
    return {"value": ops / dt, "unit": "mul+relin ops/s", "cores": 1, "kind": kind,
            "sample": f"{ops} CKKS mul+relin at N=2^{logn}, L={L} (+1 special), one thread, {dt:.1f} s"}


def served_latency(torch, ctx, a, b, out, key, ws, ws_bytes, st, B, L, n, rate, requests, max_batch, seed=7):
    """Open-loop serving: `requests` mul+relin requests arrive as a Poisson process at
    `rate` per second; each request's ciphertext pair sits in a ring of B resident
    slots.  The server takes every request that has arrived (at most max_batch), runs
    them as one batched multiply + relinearize and waits for it; a request's latency is
    its batch's completion time minus its arrival time.  max_batch = 1 is the plain
    per-ciphertext path under the same arrivals."""
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=requests))
    s2, s3 = 2 * L * n * 8, 3 * L * n * 8

    def run(slot, k):
        ctx.ckks_multiply(a.data_ptr() + slot * s2, b.data_ptr() + slot * s2, out.data_ptr() + slot * s3, k, L, st)
        ctx.relinearize(out.data_ptr() + slot * s3, key.data_ptr(), k, L, ws.data_ptr(), ws_bytes, st)

    lat = np.empty(requests)
    batches = 0
    done = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while done < requests:
        now = time.perf_counter() - t0
        if arrivals[done] > now:
            time.sleep(min(arrivals[done] - now, 1e-3))
            continue
        ready = int(np.searchsorted(arrivals, now, side="right")) - done
        k = min(ready, max_batch)
        slot = done % B
        k1 = min(k, B - slot)
        run(slot, k1)
        if k > k1:
            run(0, k - k1)
        torch.cuda.synchronize()
        fin = time.perf_counter() - t0
        lat[done:done + k] = fin - arrivals[done:done + k]
        done += k
        batches += 1
    span = time.perf_counter() - t0
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_batch": requests / batches, "served_per_s": requests / span}


def bench_mulrelin(args):
    import torch
    import sealhip
    sealhip.init(0)
    logn, n = 16, 1 << 16
    K, B = 25, args.batch
    L = K - 1
    moduli = sealhip.coeff_modulus_create(n, [60] * K)
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    gen = torch.Generator(device="cuda").manual_seed(1)
    a = torch.empty((B, 2, L, n), dtype=torch.int64, device="cuda")
    b = torch.empty((B, 2, L, n), dtype=torch.int64, device="cuda")
    key = torch.empty((K - 1, 2, K, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, a, moduli[:L], 2, gen)
    fill_uniform(torch, b, moduli[:L], 2, gen)
    fill_uniform(torch, key, moduli, 2, gen)
    out = torch.empty((B, 3, L, n), dtype=torch.int64, device="cuda")
    ws_bytes = ctx.keyswitch_workspace_size(B, L)
    ws = torch.empty(((ws_bytes + 7) // 8,), dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream

    def step_batched():
        ctx.ckks_multiply(a.data_ptr(), b.data_ptr(), out.data_ptr(), B, L, st)
        ctx.relinearize(out.data_ptr(), key.data_ptr(), B, L, ws.data_ptr(), ws_bytes, st)

    ws1_bytes = ctx.keyswitch_workspace_size(1, L)
    stride3 = 3 * L * n * 8
    stride2 = 2 * L * n * 8

    def step_plain():
        for i in range(B):
            ctx.ckks_multiply(a.data_ptr() + i * stride2, b.data_ptr() + i * stride2, out.data_ptr() + i * stride3,
                              1, L, st)
            ctx.relinearize(out.data_ptr() + i * stride3, key.data_ptr(), 1, L, ws.data_ptr(), ws1_bytes, st)

    for _ in range(args.warmup):
        step_batched()
    torch.cuda.synchronize()
    # dominant kernel: the key switch's ModUp rows pass, bracketed by HIP events that the
    # library records on its launch stream inside every timed step
    probes = [(sealhip.HipEvent(), sealhip.HipEvent()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    with Timer(torch) as tm:
        for s in range(args.steps):
            ctx.probe_keyswitch(*probes[s])
            step_batched()
    wall = time.perf_counter() - t0
    ms = tm.ms / args.steps
    ops_s = B / (ms * 1e-3)
    rows_ms = sum(b.elapsed_ms(e) for b, e in probes) / args.steps

    plain = None
    if not args.no_plain:
        # the batch must give each ciphertext the same words as the per-ciphertext path
        probe = [0, B // 2 + 1, B - 1]
        batched_words = out[probe, :2].clone()
        step_plain()
        torch.cuda.synchronize()
        with Timer(torch) as tp:
            step_plain()
        same = bool(torch.equal(out[probe, :2], batched_words))
        if not same:
            raise SystemExit("bench self-check failed: batched and per-ciphertext results differ")
        plain = {"ops_s": B / (tp.ms * 1e-3), "ms_per_step": tp.ms, "calls": 2 * B,
                 "self_check": f"batched == per-ciphertext words on ciphertexts {probe}"}

    served = None
    if not args.no_serve:
        # arrivals at half the batched throughput, the batched server against the plain one
        rate = 0.5 * ops_s
        reqs = max(200, int(rate * args.serve_seconds))
        served = {"arrivals": "poisson", "rate_per_s": rate, "requests": reqs,
                  "batched": served_latency(torch, ctx, a, b, out, key, ws, ws_bytes, st, B, L, n, rate, reqs, B),
                  "plain": served_latency(torch, ctx, a, b, out, key, ws, ws_bytes, st, B, L, n, rate, reqs, 1)}

    # The dominant kernel: the ModUp rows pass fused with the products by the key.  It
    # reads every ModUp limb once (B x L x L: each data prime's own digit is the NTT-form
    # input limb, so L - 1 digits go to each of the L data primes and L to the special
    # prime), those B x L input limbs, the key's 2 L (L+1) limbs and the L+1 twiddle
    # tables once, and writes the 2 B (L+1) product limbs; it runs the last 8 butterfly
    # layers of every ModUp limb and 2 B L (L+1) N multiply-adds.
    tlimbs = B * L * L
    rows_bytes = (tlimbs + B * L + 2 * L * (L + 1) + 2 * B * (L + 1)) * n * 8 + K * n * 16
    rows_bfly = tlimbs * (n // 2) * (logn - logn // 2)
    rows_macs = 2 * B * L * (L + 1) * n
    limb_bytes = n * 8
    unit_bytes = 6 * L * limb_bytes + (2 * (K - 1) * K * limb_bytes) / B  # inputs 4L + output 2L + key share
    step_bytes = unit_bytes * B
    achieved_step = step_bytes / (ms * 1e-3) / 1e9
    achieved_rows = rows_bytes / (rows_ms * 1e-3) / 1e9
    out_json = {
        "metric": "HE mul+relin ops/s at N=2^16 batched vs plain; p50/p99 served latency",
        "value": ops_s,
        "unit": "mul+relin ops/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded uniform residues; injected uniform relinearization key)",
        "config": {"workload": f"CKKS multiply+relinearize, N=2^16, 24 data + 1 special 60-bit primes, "
                               f"batch of {B} ciphertext pairs, NTT form, key resident",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved_rows, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_rows / HBM_PEAK_GBS,
                     "traffic": KEYPROD_TRAFFIC_B128 if B == 128 else None,
                     "butterflies_per_launch": rows_bfly, "macs_per_launch": rows_macs,
                     "kernel": f"ntt_rows_keyprod_kernel<16,8,4,16,4,1,false>: key-switch ModUp NTT rows pass over "
                               f"{tlimbs} limbs (last {logn - logn // 2} layers) fused with the key products",
                     "algorithmic_bytes": rows_bytes, "avg_ms": rows_ms,
                     "butterflies_per_s": rows_bfly / (rows_ms * 1e-3),
                     "macs_per_s": rows_macs / (rows_ms * 1e-3),
                     "valu": ({"wave_instructions": KEYPROD_VALU_B128,
                               "achieved_per_s": KEYPROD_VALU_B128 / (rows_ms * 1e-3),
                               "peak_per_s": VALU_PEAK_WAVE_INSTR_S,
                               "frac": KEYPROD_VALU_B128 / (rows_ms * 1e-3) / VALU_PEAK_WAVE_INSTR_S}
                              if B == 128 else None)},
        "step_roofline": {"achieved": achieved_step, "unit": "GB/s", "frac": achieved_step / HBM_PEAK_GBS,
                          "algorithmic_bytes_per_op": unit_bytes},
        "plain": plain,
        "batched_over_plain": (ops_s / plain["ops_s"]) if plain else None,
        "served_latency": served,
        "wall_s": wall,
    }
    if not args.no_bfv:
        del a, b, out, ws
        torch.cuda.empty_cache()
        out_json["bfv"] = bench_bfv_leg(torch, sealhip, B, args.steps, args.warmup)
    if not args.no_cpu:
        out_json["cpu_baseline"] = cpu_baseline_mulrelin(moduli, logn, args.cpu_seconds)
    return out_json


def cpu_baseline_ntt(moduli, logn, budget_s):
    import numpy as np
    import checkers as ck
    n = 1 << logn
    rng = np.random.default_rng(3)
    a = rng.integers(0, moduli[0], size=n, dtype=np.uint64)
    ref = ck.ref()
    kind = "reference" if ref is not None else "port"
    t0 = time.perf_counter()
    limbs = 0
    while time.perf_counter() - t0 < budget_s:
        q = moduli[limbs % len(moduli)]
        b = a % np.uint64(q)
        if ref is not None:
            ref.ref_ntt(logn, q, ck.ptr(b), 1, 0)
            ref.ref_ntt(logn, q, ck.ptr(b), 1, 1)
        else:
            b = ck.ntt(ck.ntt(b, q, logn), q, logn, inverse=True)
        limbs += 1
    dt = time.perf_counter() - t0
    return {"value": limbs / dt, "unit": "limb fwd+inv NTT/s", "cores": 1, "kind": kind,
            "sample": f"{limbs} limbs of N=2^{logn} (forward+inverse), one thread"}


def bench_ntt(args):
    import torch
    import sealhip
    sealhip.init(0)
    logn, n, L, polys = 16, 1 << 16, 24, 256
    moduli = sealhip.coeff_modulus_create(n, [60] * L)
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    gen = torch.Generator(device="cuda").manual_seed(1)
    d = torch.empty((polys, L, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, d, moduli, 1, gen)
    st = torch.cuda.current_stream().cuda_stream

    def step():
        ctx.ntt_forward(d.data_ptr(), polys, L, st)
        ctx.ntt_inverse(d.data_ptr(), polys, L, st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with Timer(torch) as tm:
        for _ in range(args.steps):
            step()
    wall = time.perf_counter() - t0
    ms = tm.ms / args.steps
    with Timer(torch) as tf:
        ctx.ntt_forward(d.data_ptr(), polys, L, st)
    ctx.ntt_inverse(d.data_ptr(), polys, L, st)
    torch.cuda.synchronize()
    limbs = polys * L
    fwd_bytes = 2 * limbs * n * 8 + L * n * 16
    achieved = fwd_bytes / (tf.ms * 1e-3) / 1e9
    out = {
        "metric": "batched NTT limbs/s (forward+inverse), N=2^16, 24x60-bit primes",
        "value": limbs / (ms * 1e-3), "unit": "limb fwd+inv/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (uniform residues, seeded)",
        "config": {"workload": "batched NTT: N=2^16, 24x60-bit primes, 128 ct x 2 polys x 24 limbs, fwd+inv",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "forward NTT (columns + rows pass)", "algorithmic_bytes": fwd_bytes,
                     "avg_ms": tf.ms},
        "wall_s": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ntt(moduli, logn, args.cpu_seconds)
    return out


def bench_bfv_leg(torch, sealhip, B, steps, warmup):
    """BFV multiply (BEHZ) + relinearize at the headline shape: N=2^16, 24 data + 1 special
    60-bit primes, t = 65537, coefficient form, batch B, inputs and key resident."""
    logn, n, K = 16, 1 << 16, 25
    L = K - 1
    moduli = sealhip.coeff_modulus_create(n, [60] * K)
    t = 65537
    ctx = sealhip.Context(sealhip.SCHEME_BFV, n, moduli, t)
    gen = torch.Generator(device="cuda").manual_seed(3)
    a = torch.empty((B, 2, L, n), dtype=torch.int64, device="cuda")
    b = torch.empty((B, 2, L, n), dtype=torch.int64, device="cuda")
    key = torch.empty((K - 1, 2, K, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, a, moduli[:L], 2, gen)
    fill_uniform(torch, b, moduli[:L], 2, gen)
    fill_uniform(torch, key, moduli, 2, gen)
    out = torch.empty((B, 3, L, n), dtype=torch.int64, device="cuda")
    mw = ctx.multiply_workspace_size(B, L)
    kw = ctx.keyswitch_workspace_size(B, L)
    ws = torch.empty(((max(mw, kw) + 7) // 8,), dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream

    def step():
        ctx.multiply(a.data_ptr(), b.data_ptr(), out.data_ptr(), B, L, ws.data_ptr(), mw, st)
        ctx.relinearize(out.data_ptr(), key.data_ptr(), B, L, ws.data_ptr(), kw, st)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    with Timer(torch) as tm:
        for _ in range(steps):
            step()
    ms = tm.ms / steps
    del a, b, key, out, ws
    torch.cuda.empty_cache()
    return {"ops_s": B / (ms * 1e-3), "ms_per_step": ms, "unit": "mul+relin ops/s",
            "config": f"BFV multiply (BEHZ) + relinearize, N=2^16, 24 data + 1 special 60-bit primes, t={t}, "
                      f"batch of {B} ciphertext pairs, coefficient form, key resident"}


def bench_chain(args):
    """CKKS chain at N=2^15, CoeffModulus::Create(32768, {60, 40 x 14, 60}) (15 data primes
    + special), scale 2^40 words: 64 ciphertext pairs at levels drawn from U[4, 15]; per
    pair, 3 x (multiply by the partner at the current level -> relinearize -> rescale),
    ciphertexts of one level batched together, the whole chain captured in one HIP graph
    and replayed.  A graph replay is one step; value = mul+relin+rescale per second."""
    import numpy as np
    import torch
    import sealhip
    sealhip.init(0)
    logn, n = 15, 1 << 15
    moduli = sealhip.coeff_modulus_create(n, [60] + [40] * 14 + [60])
    K = len(moduli)
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    rng = np.random.default_rng(9)
    pairs = 64
    levels = rng.integers(4, K, size=pairs)           # data levels 4..15
    groups = {int(l): int(c) for l, c in zip(*np.unique(levels, return_counts=True))}
    gen = torch.Generator(device="cuda").manual_seed(4)
    key = torch.empty((K - 1, 2, K, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, key, moduli, 2, gen)
    bufs = {}
    max_ws = 0
    for l, c in groups.items():
        x = torch.empty((c, 2, l, n), dtype=torch.int64, device="cuda")
        fill_uniform(torch, x, moduli[:l], 2, gen)
        ys = []
        for s in range(3):   # the partner at each level of the chain
            y = torch.empty((c, 2, l - s, n), dtype=torch.int64, device="cuda")
            fill_uniform(torch, y, moduli[:l - s], 2, gen)
            ys.append(y)
        prods = [torch.empty((c, 3, l - s, n), dtype=torch.int64, device="cuda") for s in range(3)]
        rels = [torch.empty((c, 2, l - s, n), dtype=torch.int64, device="cuda") for s in range(3)]
        accs = [torch.empty((c, 2, l - s - 1, n), dtype=torch.int64, device="cuda") for s in range(3)]
        bufs[l] = (c, x, ys, prods, accs, rels)
        for s in range(3):
            max_ws = max(max_ws, ctx.keyswitch_workspace_size(c, l - s), ctx.rescale_workspace_size(c, 2, l - s))
    ws = torch.empty(((max_ws + 7) // 8,), dtype=torch.int64, device="cuda")

    def chain(st):
        for l, (c, x, ys, prods, accs, rels) in bufs.items():
            cur = x
            for s in range(3):
                ll = l - s
                ctx.ckks_multiply(cur.data_ptr(), ys[s].data_ptr(), prods[s].data_ptr(), c, ll, st)
                ctx.relinearize(prods[s].data_ptr(), key.data_ptr(), c, ll, ws.data_ptr(), max_ws, st)
                rels[s].copy_(prods[s][:, :2])   # the size-2 results, packed
                ctx.rescale_to_next(rels[s].data_ptr(), accs[s].data_ptr(), c, 2, ll, ws.data_ptr(), max_ws, st)
                cur = accs[s]

    st = torch.cuda.current_stream().cuda_stream
    chain(st)
    torch.cuda.synchronize()
    eager = {l: b[4][2].clone() for l, b in bufs.items()}
    t0 = time.perf_counter()
    with Timer(torch) as te:
        for _ in range(args.steps):
            chain(st)
    eager_ms = te.ms / args.steps
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(graph, stream=cap):
            chain(torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(cap)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    # every replay must give the eager words (HIP graphs replayed with eager work between)
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        for l, b in bufs.items():
            if not torch.equal(b[4][2], eager[l]):
                raise SystemExit("bench self-check failed: graph replay differs from the eager chain")
        chain(st)
        torch.cuda.synchronize()
    with Timer(torch) as tg:
        for _ in range(args.steps):
            graph.replay()
    wall = time.perf_counter() - t0
    ms = tg.ms / args.steps
    ops = pairs * 3
    return {"metric": "CKKS mul+relin+rescale per second, N=2^15 chain of 3, uneven levels, HIP graph",
            "value": ops / (ms * 1e-3), "unit": "mul+relin+rescale/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded uniform residues; injected key)",
            "config": {"workload": f"CKKS chain: N=2^15, 15+1 primes (60, 40x14, 60), {pairs} pairs at levels "
                                   f"U[4,15] ({len(groups)} level groups), 3 x (mul, relin, rescale)",
                       "parallelism": "single GPU"},
            "eager_ms_per_step": eager_ms, "graph_over_eager": eager_ms / ms,
            "self_check": "every graph replay equals the eager chain's words", "wall_s": wall}


def bench_rotsum(args):
    """Hoisted rotations: N=2^16, CoeffModulus::Create(65536, {60, 50 x 23, 60}) (24 data
    primes + special), 16 ciphertexts in NTT form, each replaced by the sum of its 15
    rotations by 2^k (k = 0..14) -- one ModUp per ciphertext (shc_rotate_sum).  The same
    sum through 15 separate rotations per ciphertext is timed beside it, and both must
    give the same words.  value = rotations per second."""
    import torch
    import sealhip
    sealhip.init(0)
    logn, n = 16, 1 << 16
    moduli = sealhip.coeff_modulus_create(n, [60] + [50] * 23 + [60])
    K = len(moduli)
    L = K - 1
    cts, steps = 16, [1 << k for k in range(15)]
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    gen = torch.Generator(device="cuda").manual_seed(5)
    x = torch.empty((cts, 2, L, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, x, moduli[:L], 2, gen)
    elts = []
    keys, hkeys = [], []
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = torch.empty(((hw + 7) // 8,), dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for s_ in steps:
        elts.append(ctx.galois_elt_from_step(s_))
        k = torch.empty((K - 1, 2, K, n), dtype=torch.int64, device="cuda")
        fill_uniform(torch, k, moduli, 2, gen)
        h = torch.empty(((hb + 7) // 8,), dtype=torch.int64, device="cuda")
        ctx.hoisted_key_prepare(k.data_ptr(), elts[-1], L, h.data_ptr(), hws.data_ptr(), hw, st)
        keys.append(k)
        hkeys.append(h)
    out = torch.empty_like(x)
    rw = ctx.rotate_sum_workspace_size(cts, L)
    gw = ctx.galois_workspace_size(cts, L)
    ws = torch.empty(((max(rw, gw) + 7) // 8,), dtype=torch.int64, device="cuda")
    hptrs, kptrs = [h.data_ptr() for h in hkeys], [k.data_ptr() for k in keys]

    def hoisted():
        ctx.rotate_sum(x.data_ptr(), cts, L, elts, hptrs, kptrs, out.data_ptr(), ws.data_ptr(), rw, st)

    tmp = torch.empty_like(x)
    acc = torch.empty_like(x)

    def separate():
        for r, e in enumerate(elts):
            tmp.copy_(x)
            ctx.apply_galois_keyswitch(tmp.data_ptr(), kptrs[r], e, cts, L, ws.data_ptr(), gw, st)
            if r == 0:
                acc.copy_(tmp)
            else:
                ctx.add(acc.data_ptr(), tmp.data_ptr(), acc.data_ptr(), 2 * cts, L, st)

    for _ in range(args.warmup):
        hoisted()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with Timer(torch) as th:
        for _ in range(args.steps):
            hoisted()
    wall = time.perf_counter() - t0
    separate()
    torch.cuda.synchronize()
    if not torch.equal(out, acc):
        raise SystemExit("bench self-check failed: hoisted rotation sum differs from separate rotations")
    with Timer(torch) as ts:
        separate()
    ms = th.ms / args.steps
    rots = cts * len(steps)
    return {"metric": "CKKS rotations/s, N=2^16, 16 ct x 15 rotations summed, hoisted",
            "value": rots / (ms * 1e-3), "unit": "rotations/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded uniform residues; injected Galois keys)",
            "config": {"workload": "hoisted rotations: N=2^16, 24+1 primes (60, 50x23, 60), 16 ciphertexts, "
                                   "rotations by 2^k, k=0..14, summed",
                       "parallelism": "single GPU"},
            "separate_ms_per_step": ts.ms, "hoisted_over_separate": ts.ms / ms,
            "self_check": "hoisted words == sum of 15 separate rotations", "wall_s": wall}


def bench_matvec(args):
    """CKKS matrix-vector products (the CKKS request type of the serving workload): N=2^15,
    CoeffModulus::Create(32768, {60, 40 x 14, 60}), ciphertexts at the first level (15
    primes), a 1024 x 1024 matrix as 1024 NTT-form diagonals, baby-step giant-step 32 x 32
    (seal-cuda_amd/matvec.py).  A step is one batch of 16 requests; value = matvec/s; a
    single request's latency is timed beside it.  Self-check: the batched product of one
    ciphertext equals its product computed alone."""
    import torch
    import sealhip
    import matvec
    sealhip.init(0)
    logn, n = 15, 1 << 15
    moduli = sealhip.coeff_modulus_create(n, [60] + [40] * 14 + [60])
    K = len(moduli)
    L = K - 1
    b = g = 32
    batch = 16
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    gen = torch.Generator(device="cuda").manual_seed(6)
    st = torch.cuda.current_stream().cuda_stream
    plains = torch.empty((g, b, L, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, plains, moduli[:L], 2, gen)

    def key():
        k = torch.empty((K - 1, 2, K, n), dtype=torch.int64, device="cuda")
        fill_uniform(torch, k, moduli, 2, gen)
        return k

    baby = [key() for _ in range(1, b)]
    giant = [key() for _ in range(1, g)]
    hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
    hws = torch.empty(((hw + 7) // 8,), dtype=torch.int64, device="cuda")
    hoisted = []
    for j, k in zip(range(1, b), baby):
        h = torch.empty(((hb + 7) // 8,), dtype=torch.int64, device="cuda")
        ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(j), L, h.data_ptr(), hws.data_ptr(), hw, st)
        hoisted.append(h)
    mv = matvec.BsgsMatvec(ctx, n, L, b, g, plains.data_ptr(), [k.data_ptr() for k in baby],
                           [h.data_ptr() for h in hoisted], [k.data_ptr() for k in giant])
    x = torch.empty((batch, 2, L, n), dtype=torch.int64, device="cuda")
    fill_uniform(torch, x, moduli[:L], 2, gen)
    out = torch.empty_like(x)
    baby_buf = torch.empty((b, batch, 2, L, n), dtype=torch.int64, device="cuda")
    inner = torch.empty((g, batch, 2, L, n), dtype=torch.int64, device="cuda")
    wsz = mv.workspace_sizes(batch)
    ws = torch.empty(((wsz + 7) // 8,), dtype=torch.int64, device="cuda")

    def run(count):
        mv.run(x.data_ptr(), count, out.data_ptr(), baby_buf.data_ptr(), inner.data_ptr(), ws.data_ptr(), wsz, st)

    for _ in range(args.warmup):
        run(batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with Timer(torch) as tb:
        for _ in range(args.steps):
            run(batch)
    wall = time.perf_counter() - t0
    first = out[0].clone()
    run(1)
    torch.cuda.synchronize()
    if not torch.equal(out[0], first):
        raise SystemExit("bench self-check failed: batched matvec differs from the single-request product")
    with Timer(torch) as t1:
        for _ in range(args.steps):
            run(1)
    ms = tb.ms / args.steps
    return {"metric": "CKKS BSGS matvec/s, N=2^15, 1024x1024, batch 16", "value": batch / (ms * 1e-3),
            "unit": "matvec/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": "synthetic (seeded uniform residues for ciphertexts, diagonals and keys)",
            "config": {"workload": "CKKS matvec: N=2^15, 15+1 primes (60, 40x14, 60), 1024x1024 diagonals, "
                                   "BSGS 32x32, 16 requests per batch", "parallelism": "single GPU"},
            "single_request_ms": t1.ms / args.steps,
            "self_check": "batched product of request 0 == its product alone", "wall_s": wall}


class MatvecServer:
    """One request type of the serving workload: a context, one matrix's diagonals (NTT
    form, synthetic residues), resident Galois keys (synthetic residues; the baby steps
    also prepared for hoisting), a pool of resident input ciphertexts, and batch buffers.
    Admission by HBM footprint: the batch buffers of m requests take m ciphertexts each
    for the input, the output, the b baby steps and the g inner sums, plus the key-switch
    workspace of m ciphertexts; max_batch is the largest m <= cap whose buffers fit in
    `budget` bytes (keys and diagonals are resident and counted apart)."""

    def __init__(self, torch, sealhip, matvec, scheme, n, moduli, plain, b, g, pool, cap, budget, gen, st):
        K = len(moduli)
        L = K - 1
        self.torch, self.n, self.L, self.st = torch, n, L, st
        self.ctx = sealhip.Context(scheme, n, moduli, plain)
        ctx = self.ctx
        self.plains = torch.empty((g, b, L, n), dtype=torch.int64, device="cuda")
        fill_uniform(torch, self.plains, moduli[:L], 2, gen)

        def key():
            k = torch.empty((K - 1, 2, K, n), dtype=torch.int64, device="cuda")
            fill_uniform(torch, k, moduli, 2, gen)
            return k

        self.baby = [key() for _ in range(1, b)]
        self.giant = [key() for _ in range(1, g)]
        hb, hw = ctx.hoisted_key_size(L), ctx.hoisted_key_workspace_size(L)
        hws = torch.empty(((hw + 7) // 8,), dtype=torch.int64, device="cuda")
        self.hoisted = []
        for j, k in zip(range(1, b), self.baby):
            h = torch.empty(((hb + 7) // 8,), dtype=torch.int64, device="cuda")
            ctx.hoisted_key_prepare(k.data_ptr(), ctx.galois_elt_from_step(j), L, h.data_ptr(), hws.data_ptr(), hw,
                                    st)
            self.hoisted.append(h)
        self.mv = matvec.BsgsMatvec(ctx, n, L, b, g, self.plains.data_ptr(), [k.data_ptr() for k in self.baby],
                                    [h.data_ptr() for h in self.hoisted], [k.data_ptr() for k in self.giant],
                                    bfv=scheme == sealhip.SCHEME_BFV)
        self.pool = torch.empty((pool, 2, L, n), dtype=torch.int64, device="cuda")
        fill_uniform(torch, self.pool, moduli[:L], 2, gen)
        ct_bytes = 2 * L * n * 8
        footprint = lambda m: m * ct_bytes * (2 + b + g) + self.mv.workspace_sizes(m)
        max_batch = cap
        while max_batch > 1 and footprint(max_batch) > budget:
            max_batch -= 1
        self.max_batch = max_batch
        self.batch_bytes = footprint(max_batch)
        self.per_request_bytes = footprint(2) - footprint(1)
        self.resident_bytes = (self.plains.numel() + sum(k.numel() for k in self.baby + self.giant)
                               + sum(h.numel() for h in self.hoisted)) * 8
        self.x = torch.empty((max_batch, 2, L, n), dtype=torch.int64, device="cuda")
        self.out = torch.empty_like(self.x)
        self.babybuf = torch.empty((b, max_batch, 2, L, n), dtype=torch.int64, device="cuda")
        self.inner = torch.empty((g, max_batch, 2, L, n), dtype=torch.int64, device="cuda")
        self.wsz = self.mv.workspace_sizes(max_batch)
        self.ws = torch.empty(((self.wsz + 7) // 8,), dtype=torch.int64, device="cuda")

    def run(self, inputs):
        """One batch: the pool entries `inputs` gathered into the batch buffer, then one
        batched matvec into self.out[:len(inputs)]."""
        for k, i in enumerate(inputs):
            self.x[k].copy_(self.pool[i])
        self.mv.run(self.x.data_ptr(), len(inputs), self.out.data_ptr(), self.babybuf.data_ptr(),
                    self.inner.data_ptr(), self.ws.data_ptr(), self.wsz, self.st)


def serve_mixed(torch, servers, rate, requests, max_batch, seed):
    """Open-loop serving of a mixed stream: Poisson arrivals at `rate` per second, each
    request of type 0 or 1 with probability 1/2 and an input drawn from that type's pool.
    The server picks the type whose oldest waiting request arrived first, runs every
    waiting request of that type (at most max_batch, and at most what that type's
    admission allows) as one batch, and waits for it; a request's latency is its batch's
    completion minus its arrival.  max_batch = 1 is the plain one-request-at-a-time path
    under the same arrivals."""
    import collections
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=requests))
    types = (rng.random(requests) < 0.5).astype(int)
    inputs = rng.integers(0, servers[0].pool.shape[0], size=requests)
    pending = [collections.deque(), collections.deque()]
    lat = np.empty(requests)
    batches = [0, 0]
    nxt = done = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while done < requests:
        now = time.perf_counter() - t0
        while nxt < requests and arrivals[nxt] <= now:
            pending[types[nxt]].append(nxt)
            nxt += 1
        if not pending[0] and not pending[1]:
            time.sleep(min(arrivals[nxt] - now, 1e-3))
            continue
        ty = 0 if not pending[1] or (pending[0] and pending[0][0] < pending[1][0]) else 1
        batch = [pending[ty].popleft() for _ in range(min(max_batch, servers[ty].max_batch, len(pending[ty])))]
        servers[ty].run([int(inputs[r]) for r in batch])
        torch.cuda.synchronize()
        fin = time.perf_counter() - t0
        lat[batch] = fin - arrivals[batch]
        done += len(batch)
        batches[ty] += 1
    span = time.perf_counter() - t0
    res = {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "served_per_s": requests / span}
    for ty, name in enumerate(("bfv", "ckks")):
        sel = lat[types == ty]
        res[name] = {"requests": int(sel.size), "p50_ms": float(np.percentile(sel, 50) * 1e3),
                     "p99_ms": float(np.percentile(sel, 99) * 1e3),
                     "mean_batch": float(sel.size / max(batches[ty], 1))}
    return res


def bench_serve(args):
    """The serving workload: a Poisson stream of matrix-vector requests, half BFV (N=8192,
    BFVDefault(8192) primes, t = 65537, a 4096 x 4096 matrix mod t, BSGS 64 x 64 in each
    row of the slot matrix) and half CKKS (N=2^15, Create(32768, {60, 40 x 14, 60}), a
    1024 x 1024 matrix, BSGS 32 x 32), keys and diagonals resident (synthetic residues).
    Arrivals at half the batched server's capacity for the mix (measured first with full
    batches); value = requests/s served by the batched server, with p50/p99 latency, and
    the plain one-at-a-time server under the same arrivals beside it.  Self-check: for
    each type, a request's words in a full batch equal its words computed alone."""
    import torch
    import sealhip
    import matvec
    sealhip.init(0)
    st = torch.cuda.current_stream().cuda_stream
    gen = torch.Generator(device="cuda").manual_seed(8)
    B = 16
    budget = 8 << 30  # HBM for each request type's batch buffers
    bfv_moduli = [0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001]
    ckks_moduli = sealhip.coeff_modulus_create(1 << 15, [60] + [40] * 14 + [60])
    servers = [MatvecServer(torch, sealhip, matvec, sealhip.SCHEME_BFV, 8192, bfv_moduli, 65537, 64, 64, 32, B,
                            budget, gen, st),
               MatvecServer(torch, sealhip, matvec, sealhip.SCHEME_CKKS, 1 << 15, ckks_moduli, 0, 32, 32, 32, B,
                            budget, gen, st)]
    # capacity with full batches, per type
    cap = []
    for sv in servers:
        full = list(range(sv.max_batch))
        for _ in range(max(args.warmup, 1)):
            sv.run(full)
        with Timer(torch) as t:
            for _ in range(args.steps):
                sv.run(full)
        cap.append(sv.max_batch * args.steps / (t.ms * 1e-3))
    mix_cap = 1.0 / (0.5 / cap[0] + 0.5 / cap[1])
    # determinism: a request's words in a full batch equal its words alone
    for sv, name in zip(servers, ("bfv", "ckks")):
        last = sv.max_batch - 1
        sv.run(list(range(sv.max_batch)))
        torch.cuda.synchronize()
        batched = [sv.out[0].clone(), sv.out[last].clone()]
        for k, i in ((0, 0), (1, last)):
            sv.run([i])
            torch.cuda.synchronize()
            if not torch.equal(sv.out[0], batched[k]):
                raise SystemExit(f"bench self-check failed: {name} request {i} differs alone and in a batch")
    rate = 0.5 * mix_cap
    reqs = max(200, int(rate * args.serve_seconds))
    batched = serve_mixed(torch, servers, rate, reqs, B, seed=11)
    plain = serve_mixed(torch, servers, rate, reqs, 1, seed=11)
    return {"metric": "mixed BFV/CKKS matvec requests/s served; p50/p99 latency", "value": batched["served_per_s"],
            "unit": "requests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 / batched["served_per_s"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64",
            "data": "synthetic (seeded uniform residues for ciphertexts, diagonals and keys; Poisson arrivals)",
            "config": {"workload": "serving: 50% BFV matvec (N=8192, BFVDefault(8192), t=65537, 4096x4096, "
                                   "BSGS 64x64), 50% CKKS matvec (N=2^15, 60+40x14+60, 1024x1024, BSGS 32x32), "
                                   f"batches of at most {B} within {budget >> 30} GiB of batch buffers per type",
                       "parallelism": "single GPU"},
            "capacity_per_s": {"bfv": cap[0], "ckks": cap[1], "mix": mix_cap},
            "admission": {"budget_bytes_per_type": budget,
                          "max_batch": {"bfv": servers[0].max_batch, "ckks": servers[1].max_batch},
                          "per_request_bytes": {"bfv": servers[0].per_request_bytes,
                                                "ckks": servers[1].per_request_bytes},
                          "batch_bytes": {"bfv": servers[0].batch_bytes, "ckks": servers[1].batch_bytes},
                          "resident_key_and_diagonal_bytes": {"bfv": servers[0].resident_bytes,
                                                              "ckks": servers[1].resident_bytes}},
            "arrivals": {"kind": "poisson", "rate_per_s": rate, "requests": reqs},
            "batched": batched, "plain": plain, "batched_p99_over_plain_p99": batched["p99_ms"] / plain["p99_ms"],
            "self_check": "per type, the first and last request of a full batch give the same words alone"}


def bench_copy(args):
    """HBM copy of 1 GiB (torch's device copy kernel): the calibration run for the
    FETCH_SIZE / WRITE_SIZE counters (known 1 GiB read + 1 GiB written per launch)."""
    import torch
    words = 1 << 27
    src = torch.ones(words, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    with Timer(torch) as tm:
        for _ in range(args.steps):
            dst.copy_(src)
    ms = tm.ms / args.steps
    gbs = 2 * words * 8 / (ms * 1e-3) / 1e9
    return {"metric": "HBM copy GB/s (read + write), 1 GiB", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic (ones)",
            "config": {"workload": "device copy of 2^27 int64 words", "parallelism": "single GPU"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus N > 1 is not supported "
                                   "(independent ciphertexts shard across GPUs with no exchange)"}))
        sys.exit(2)
    out = {"mulrelin": bench_mulrelin, "ntt": bench_ntt, "copy": bench_copy, "chain": bench_chain,
           "rotsum": bench_rotsum, "matvec": bench_matvec, "serve": bench_serve}[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
