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

`--workload primitives`: rotation, rescale, encryption, decryption, both encoders, key
generation and serialization, each timed alone (CKKS at the headline shape, BFV at
N=8192), batches of 128.

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
KEYPROD_TRAFFIC_B128 = 44.16e9 + 3.36e9
# VALU wave instructions per launch of the same kernel (rocprofv3 --pmc SQ_INSTS_VALU, same
# file) and the VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles
# per SIMD at the 2.4 GHz peak engine clock.  The kernel is bound by VALU issue (exact
# 64-bit modular arithmetic from 32-bit multiplies), so this is its figure of merit.
KEYPROD_VALU_B128 = 8140554240
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4
CKKS = 2
BFV_SCHEME = 1
T_BFV = 65537


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="mulrelin",
                   choices=["mulrelin", "ntt", "copy", "chain", "rotsum", "matvec", "serve", "bfv8192", "bfv",
                            "primitives"])
    p.add_argument("--no-bfv", action="store_true", help="skip the BFV leg of the mulrelin workload")
    p.add_argument("--batch", type=int, default=128)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-plain", action="store_true", help="skip the plain per-ciphertext leg")
    p.add_argument("--cpu-seconds", type=float, default=6.0,
                   help="CPU-baseline budget per workload (half one thread, half all threads)")
    p.add_argument("--no-serve", action="store_true", help="skip the served-latency leg")
    p.add_argument("--serve-seconds", type=float, default=1.5, help="arrival span of the served-latency leg")
    p.add_argument("--headline-only", action="store_true",
                   help="mulrelin: skip the other workloads' sub-objects in the JSON line")
    p.add_argument("--sub-steps", type=int, default=5, help="timed steps of each sub-workload in the default line")
    p.add_argument("--no-tune", action="store_true", help="keep the built-in key-switch launch shapes")
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


def cpu_threads():
    """Host threads for the all-core CPU baseline: every core the process may use, at
    most 16 (the CPU share of one GPU's host)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(make_worker, unit, units_per_call, budget_s, sample):
    """Times the CPU path on this host: make_worker() returns a callable that does one
    call of independent work (its own session and buffers; `units_per_call` units of the
    metric).  Half the budget runs one worker on one thread (latency per call, rate);
    the other half runs one worker per thread on cpu_threads() threads at once
    (all-core throughput; the reference CPU path and the oracle hold no global mutable
    state, and ctypes releases the GIL during their calls).  The reference CPU library
    (oracle/_ref) is used where it was shipped with this tree, else the oracle."""
    import threading
    import checkers as ck
    kind = "reference" if ck.ref() is not None else "port"
    half = budget_s / 2
    w = make_worker(kind)
    calls = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < half or calls == 0:
        w()
        calls += 1
    dt1 = time.perf_counter() - t0
    nt = cpu_threads()
    workers = ck.parallel_map(lambda _: make_worker(kind), range(nt), workers=nt)  # contexts built side by side
    counts = [0] * nt
    start = time.perf_counter()
    stop = start + half

    def loop(i):
        while time.perf_counter() < stop or counts[i] == 0:
            workers[i]()
            counts[i] += 1

    th = [threading.Thread(target=loop, args=(i,)) for i in range(nt)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dtn = time.perf_counter() - start
    return {"value": calls * units_per_call / dt1, "unit": unit, "cores": 1, "kind": kind,
            "latency_ms": dt1 / calls * 1e3,
            "sample": f"{sample}; one thread: {calls} call(s) in {dt1:.1f} s",
            "all_core": {"value": sum(counts) * units_per_call / dtn, "unit": unit, "threads": nt,
                         "nproc": os.cpu_count(),
                         "sample": f"{nt} threads, one independent call each at a time: {sum(counts)} call(s) "
                                   f"in {dtn:.1f} s"}}


def cpu_baseline_mulrelin(moduli, logn, budget_s):
    """CKKS multiply + relinearize at the headline shape through the reference CPU path
    (src/evaluator.cpp multiply + relinearize; oracle/_ref), else the oracle."""
    import numpy as np
    import checkers as ck
    n, K = 1 << logn, len(moduli)
    L = K - 1
    rng = np.random.default_rng(5)
    a = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    b = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    key = ck.random_key(rng, moduli, logn)

    def make(kind):
        if kind == "reference":
            sess = ck.RefSession(CKKS, n, moduli)
            return lambda: sess.relinearize(sess.multiply(a, b, L), key, L)
        return lambda: ck.relinearize(CKKS, ck.ckks_multiply(a, b, moduli[:L], logn), key, moduli, L, logn)

    return cpu_baseline(make, "mul+relin ops/s", 1, budget_s,
                        f"CKKS mul+relin at N=2^{logn}, L={L} (+1 special), injected uniform words")


def served_mulrelin(torch, sealhip, ctx, a, b, key, L, n, seconds):
    """The served leg of the headline: multiply + relinearize requests through the
    library's continuous-batching server, each request a pair of host ciphertexts
    (N=2^16, L=24: 50 MB in, 25 MB out, over PCIe inside the latency).  Capacity is
    measured with a burst; Poisson arrivals at half of it then go to the batched server
    (2 lanes, batches of at most 32, a 32 GiB device arena) and to the plain server (one
    lane, one request at a time).  Self-check: a request's words served alone equal its
    words in the burst and the resident batch's words for the same pair."""
    import numpy as np
    pool = []
    for i in range(4):
        pair = torch.stack([a[i], b[i]]).cpu()
        pool.append(pair.reshape(-1).numpy().view(np.uint64).copy())
    ring, B = 24, 32
    staging = 4 * pool[0].nbytes  # requests use pinned buffers in place; staging serves the self-check
    batched_sv = sealhip.Server(32 << 30, staging, lanes=2, max_batch=B)
    plain_sv = sealhip.Server(2 << 30, staging, lanes=1, max_batch=1)
    tb, tp = batched_sv.add_mulrelin(ctx, L, key.data_ptr()), plain_sv.add_mulrelin(ctx, L, key.data_ptr())
    pool_b, pool_p = pinned_copy(batched_sv, pool), pinned_copy(plain_sv, pool)
    serve_open_loop(batched_sv, [(tb, pool_b)], 0, 8, seed=1, ring=ring, burst=True)
    cap = serve_open_loop(batched_sv, [(tb, pool_b)], 0, 48, seed=2, ring=ring, burst=True)
    words = batched_sv.request_size(tb)[1]
    outs = [np.empty(words, dtype=np.uint64) for _ in range(8)]
    ids = [batched_sv.submit(tb, pool[i % 4], outs[i]) for i in range(8)]
    for i in ids:
        batched_sv.wait(i)
    alone = np.empty(words, dtype=np.uint64)
    plain_sv.wait(plain_sv.submit(tp, pool[1], alone))
    if not (np.array_equal(alone, outs[1]) and np.array_equal(alone, outs[5])):
        raise SystemExit("bench self-check failed: a served mul+relin request differs alone and in a batch")
    rate = 0.5 * cap["served_per_s"]
    reqs = max(100, int(rate * seconds))
    res = {"server": "library (shc_server_*): host ciphertexts in and out (pinned buffers used in place), 2 "
                     f"lanes, batches of at most {B}, 32 GiB device arena; plain = 1 lane, one request at a time",
           "capacity_per_s": cap["served_per_s"], "arrivals": "poisson", "rate_per_s": rate, "requests": reqs,
           "batched": serve_open_loop(batched_sv, [(tb, pool_b)], rate, reqs, seed=7, ring=ring),
           "plain": serve_open_loop(plain_sv, [(tp, pool_p)], rate, reqs, seed=7, ring=ring),
           "max_batches_in_flight": batched_sv.stats()["max_batches_in_flight"]}
    batched_sv.close()
    plain_sv.close()
    return res, alone


def facade_leg(ops=128, warmup=16):
    """The plain drop-in path as a reference program would call it: the troyn:: class API
    (include/troyhip.hpp) one ciphertext at a time on the default stream
    (evaluator.multiply(a, b, out); evaluator.relinearizeInplace(out, rk)), at the same
    shape, timed by seal-cuda_amd/bench_facade (a separate process, built with the
    library)."""
    import subprocess
    exe = os.path.join(HERE, "seal-cuda_amd", "bench_facade")
    if not os.path.exists(exe):
        return {"error": f"{exe} not built"}
    r = subprocess.run([exe, str(ops), str(warmup)], capture_output=True, text=True, timeout=300)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"bench_facade exited {r.returncode}: {r.stderr[-300:]}"}
    res = json.loads(lines[-1])
    res["path"] = ("troyn::Evaluator multiply + relinearizeInplace per ciphertext (C++ API over the C ABI, "
                   "default stream, pooled device blocks), CKKS N=2^16, 24+1 primes")
    return res


def facade_io_leg(reps=8):
    """Ciphertext and relinearization-key save / load and key generation through the
    troyn:: API at the headline shape (seal-cuda_amd/bench_facade io, a separate process
    like the drop-in timing)."""
    import subprocess
    exe = os.path.join(HERE, "seal-cuda_amd", "bench_facade")
    if not os.path.exists(exe):
        return {"error": "bench_facade not built"}
    r = subprocess.run([exe, "io", str(reps)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": f"bench_facade io exited {r.returncode}: {r.stderr[-300:]}"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def bench_mulrelin(args):
    import numpy as np
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

    # launch shapes of the key switch's ModUp kernels, chosen by the library's autotuner
    # for this batch and for the per-ciphertext path (the products buffer is its scratch)
    tune = None
    if not args.no_tune:
        ctx.ckks_multiply(a.data_ptr(), b.data_ptr(), out.data_ptr(), B, L, st)
        tune = {"batched": ctx.autotune_relinearize(out.data_ptr(), key.data_ptr(), B, L, ws.data_ptr(), ws_bytes,
                                                    reps=2, stream=st),
                "per_ciphertext": ctx.autotune_relinearize(out.data_ptr(), key.data_ptr(), 1, L, ws.data_ptr(),
                                                           ws_bytes, reps=8, stream=st)}
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

    facade = None
    if not args.no_plain:
        facade = facade_leg()

    served = None
    if not args.no_serve:
        served, served_words = served_mulrelin(torch, sealhip, ctx, a, b, key, L, n, args.serve_seconds)
        if not np.array_equal(served_words.reshape(2, L, n), out[1, :2].cpu().numpy().view(np.uint64)):
            raise SystemExit("bench self-check failed: a served request differs from the resident batch's words")

    # The dominant kernel: the ModUp rows pass fused with the products by the key.  It
    # reads every ModUp limb once (B x L x L: each data prime's own digit is the NTT-form
    # input limb, so L - 1 digits go to each of the L data primes and L to the special
    # prime), those B x L input limbs, the 2 B L limbs of (c0, c1) (their p multiples join
    # the sums, so that the ModDown writes them unread), the key's 2 L (L+1) limbs and the
    # L+1 twiddle tables once, and writes the 2 B (L+1) product limbs; it runs the last 8
    # butterfly layers of every ModUp limb and 2 B L (L+1) N multiply-adds.
    grp = ctx.keyswitch_shape(L, B)[1] or 4
    kname = {4: "4,1", 2: "2,1", 1: "1,2"}[grp]
    valu_known = B == 128 and grp == 4  # the SQ_INSTS_VALU count was measured for groups of 4
    tlimbs = B * L * L
    rows_bytes = (tlimbs + B * L + 2 * B * L + 2 * L * (L + 1) + 2 * B * (L + 1)) * n * 8 + K * n * 16
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
                     "traffic": KEYPROD_TRAFFIC_B128 if valu_known else None,
                     "butterflies_per_launch": rows_bfly, "macs_per_launch": rows_macs,
                     "kernel": f"ntt_rows_keyprod_kernel<16,8,4,16,{kname},false>: key-switch ModUp NTT rows pass over "
                               f"{tlimbs} limbs (last {logn - logn // 2} layers) fused with the key products",
                     "algorithmic_bytes": rows_bytes, "avg_ms": rows_ms,
                     "butterflies_per_s": rows_bfly / (rows_ms * 1e-3),
                     "macs_per_s": rows_macs / (rows_ms * 1e-3),
                     "valu": ({"wave_instructions": KEYPROD_VALU_B128,
                               "achieved_per_s": KEYPROD_VALU_B128 / (rows_ms * 1e-3),
                               "peak_per_s": VALU_PEAK_WAVE_INSTR_S,
                               "frac": KEYPROD_VALU_B128 / (rows_ms * 1e-3) / VALU_PEAK_WAVE_INSTR_S}
                              if valu_known else None)},
        "step_roofline": {"achieved": achieved_step, "unit": "GB/s", "frac": achieved_step / HBM_PEAK_GBS,
                          "algorithmic_bytes_per_op": unit_bytes},
        "plain": plain,
        "batched_over_plain": (ops_s / plain["ops_s"]) if plain else None,
        "plain_facade": facade,
        "autotune": tune,
        "batched_over_facade": (ops_s / facade["ops_s"]) if facade and "ops_s" in facade else None,
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
    """Forward + inverse negacyclic NTT of limbs at N = 2^logn through the reference CPU
    path (nttNegacyclicHarvey / inverseNttNegacyclicHarvey, src/utils/ntt.cpp:158-213),
    64 limbs of one prime per call (the tables are built once per call), cycling
    through the primes."""
    import numpy as np
    import checkers as ck
    n = 1 << logn
    per_call = 64
    rng = np.random.default_rng(3)
    base = [rng.integers(0, q, size=per_call * n, dtype=np.uint64) for q in moduli[:4]]

    def make(kind):
        state = {"k": 0}
        buf = np.empty(per_call * n, dtype=np.uint64)

        def run():
            k = state["k"] % len(base)
            state["k"] += 1
            q = moduli[k]
            buf[:] = base[k]
            if kind == "reference":
                ck.ref().ref_ntt(logn, q, ck.ptr(buf), per_call, 0)
                ck.ref().ref_ntt(logn, q, ck.ptr(buf), per_call, 1)
            else:
                ck.oracle().or_ntt_batch(logn, q, ck.ptr(buf), per_call, 0)
                ck.oracle().or_ntt_batch(logn, q, ck.ptr(buf), per_call, 1)
        return run

    return cpu_baseline(make, "limb fwd+inv/s", per_call, budget_s,
                        f"forward + inverse NTT, N=2^{logn}, 60-bit primes, {per_call} limbs per call")


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
    # the batch must give each pair the same words as the per-ciphertext path
    probe = [0, B // 2 + 1, B - 1]
    batched_words = out[probe, :2].clone()
    mw1, kw1 = ctx.multiply_workspace_size(1, L), ctx.keyswitch_workspace_size(1, L)
    stride2, stride3 = 2 * L * n * 8, 3 * L * n * 8
    for i in probe:
        ctx.multiply(a.data_ptr() + i * stride2, b.data_ptr() + i * stride2, out.data_ptr() + i * stride3, 1, L,
                     ws.data_ptr(), mw1, st)
        ctx.relinearize(out.data_ptr() + i * stride3, key.data_ptr(), 1, L, ws.data_ptr(), kw1, st)
    torch.cuda.synchronize()
    if not torch.equal(out[probe, :2], batched_words):
        raise SystemExit("bench self-check failed: BFV batched and per-ciphertext results differ")
    del a, b, key, out, ws
    torch.cuda.empty_cache()
    return {"ops_s": B / (ms * 1e-3), "ms_per_step": ms, "unit": "mul+relin ops/s",
            "self_check": f"batched == per-ciphertext words on pairs {probe}",
            "config": f"BFV multiply (BEHZ) + relinearize, N=2^16, 24 data + 1 special 60-bit primes, t={t}, "
                      f"batch of {B} ciphertext pairs, coefficient form, key resident"}


def bench_bfv(args):
    """The BFV leg of the headline alone (for profiling): BFV multiply (BEHZ) + relinearize,
    N=2^16, 24+1 primes of 60 bits, t = 65537, batch --batch."""
    import torch
    import sealhip
    sealhip.init(0)
    r = bench_bfv_leg(torch, sealhip, args.batch, args.steps, args.warmup)
    return {"metric": "BFV mul+relin ops/s at N=2^16, L=24", "value": r["ops_s"], "unit": r["unit"], "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded uniform residues)",
            "config": {"workload": r["config"], "parallelism": "single GPU"}, "self_check": r["self_check"]}


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

    def chain(st, probe_at=None, probe=None):
        for l, (c, x, ys, prods, accs, rels) in bufs.items():
            cur = x
            for s in range(3):
                ll = l - s
                ctx.ckks_multiply(cur.data_ptr(), ys[s].data_ptr(), prods[s].data_ptr(), c, ll, st)
                if probe_at == (l, s):
                    ctx.probe_keyswitch(*probe)
                ctx.relinearize(prods[s].data_ptr(), key.data_ptr(), c, ll, ws.data_ptr(), max_ws, st)
                rels[s].copy_(prods[s][:, :2])   # the size-2 results, packed
                ctx.rescale_to_next(rels[s].data_ptr(), accs[s].data_ptr(), c, 2, ll, ws.data_ptr(), max_ws, st)
                cur = accs[s]

    # the same chain with every level group in one launch per stage (shc_*_groups): per
    # round, one grouped multiply, relinearize and rescale (read straight from the
    # size-3 products)
    order = sorted(bufs)
    gw = 0
    for s in range(3):
        pr = [(bufs[l][3][s].data_ptr(), bufs[l][0], l - s) for l in order]
        gw = max(gw, ctx.relinearize_groups_workspace_size(pr), ctx.rescale_groups_workspace_size(pr))
    gws = torch.empty(((gw + 7) // 8,), dtype=torch.int64, device="cuda")

    def chain_grouped(st):
        for s in range(3):
            cur = [(bufs[l][1] if s == 0 else bufs[l][4][s - 1]).data_ptr() for l in order]
            a = [(p_, bufs[l][0], l - s) for p_, l in zip(cur, order)]
            b = [(bufs[l][2][s].data_ptr(), bufs[l][0], l - s) for l in order]
            pr = [(bufs[l][3][s].data_ptr(), bufs[l][0], l - s) for l in order]
            out = [(bufs[l][4][s].data_ptr(), bufs[l][0], l - s - 1) for l in order]
            ctx.ckks_multiply_groups(a, b, pr, st)
            ctx.relinearize_groups(pr, key.data_ptr(), gws.data_ptr(), gw, st)
            ctx.rescale_to_next_groups(pr, 3, out, gws.data_ptr(), gw, st)

    st = torch.cuda.current_stream().cuda_stream
    chain(st)
    torch.cuda.synchronize()
    eager = {l: b[4][2].clone() for l, b in bufs.items()}
    for b in bufs.values():
        b[4][2].zero_()
    chain_grouped(st)
    torch.cuda.synchronize()
    for l, b in bufs.items():
        if not torch.equal(b[4][2], eager[l]):
            raise SystemExit("bench self-check failed: the grouped chain differs from the per-level chain")
    # dominant kernel: the fused ModUp rows pass + key products of the largest level's
    # first key switch, bracketed by HIP events on the library's launch stream (eager)
    top = max(groups)
    probe = (sealhip.HipEvent(), sealhip.HipEvent())
    chain(st, probe_at=(top, 0), probe=probe)
    torch.cuda.synchronize()
    kp_ms = probe[0].elapsed_ms(probe[1])
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
    per_level_ms = tg.ms / args.steps
    # the grouped chain: eager, then captured; every replay must give the same words
    with Timer(torch) as tge:
        for _ in range(args.steps):
            chain_grouped(st)
    grouped_eager_ms = tge.ms / args.steps
    ggraph = torch.cuda.CUDAGraph()
    cap2 = torch.cuda.Stream()
    cap2.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap2):
        with torch.cuda.graph(ggraph, stream=cap2):
            chain_grouped(torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(cap2)
    for _ in range(2):
        for b in bufs.values():
            b[4][2].zero_()
        ggraph.replay()
        torch.cuda.synchronize()
        for l, b in bufs.items():
            if not torch.equal(b[4][2], eager[l]):
                raise SystemExit("bench self-check failed: a grouped graph replay differs from the eager chain")
        chain(st)
        torch.cuda.synchronize()
    with Timer(torch) as tgg:
        for _ in range(args.steps):
            ggraph.replay()
    grouped_ms = tgg.ms / args.steps
    wall = time.perf_counter() - t0
    ms = min(per_level_ms, grouped_ms)
    ops = pairs * 3
    limb = n * 8
    # step: per (group, round) at level m, inputs 4m limbs and output 2(m-1) limbs per
    # pair, the key's 2m(m+1) limbs once per key-switch call
    step_bytes = sum((c * (4 * (l - s) + 2 * (l - s - 1)) + 2 * (l - s) * (l - s + 1)) * limb
                     for l, c in groups.items() for s in range(3))
    # the probed launch at level L = top for c pairs: L x L ModUp limbs, L input limbs and
    # the 2L limbs of (c0, c1) per pair, the key's 2L(L+1) limbs, 2(L+1) product limbs per
    # pair, L+1 twiddle tables
    c_top, L_top = groups[top], top
    kp_bytes = ((c_top * (L_top * L_top + 3 * L_top + 2 * (L_top + 1)) + 2 * L_top * (L_top + 1)) * limb
                + (L_top + 1) * n * 16)
    out = {"metric": "CKKS mul+relin+rescale per second, N=2^15 chain of 3, uneven levels, HIP graph",
           "value": ops / (ms * 1e-3), "unit": "mul+relin+rescale/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded uniform residues; injected key)",
           "config": {"workload": f"CKKS chain: N=2^15, 15+1 primes (60, 40x14, 60), {pairs} pairs at levels "
                                  f"U[4,15] ({len(groups)} level groups), 3 x (mul, relin, rescale)",
                      "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": kp_bytes / (kp_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": kp_bytes / (kp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "kernel": f"ntt_rows_keyprod_kernel<15,...>: fused ModUp rows pass + key products of the "
                                  f"level-{L_top} group ({c_top} pairs), eager pass",
                        "algorithmic_bytes": kp_bytes, "avg_ms": kp_ms},
           "step_roofline": {"achieved": step_bytes / (ms * 1e-3) / 1e9, "unit": "GB/s",
                             "frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "algorithmic_bytes_per_step": step_bytes},
           "eager_ms_per_step": eager_ms, "graph_over_eager": eager_ms / per_level_ms,
           "per_level_graph_ms": per_level_ms, "grouped_graph_ms": grouped_ms, "grouped_eager_ms": grouped_eager_ms,
           "path": ("level groups in one launch per stage (shc_*_groups), HIP graph" if grouped_ms <= per_level_ms
                    else "one batched call per level group and stage, HIP graph"),
           "self_check": "every graph replay (per-level and grouped) equals the eager per-level chain's words",
           "wall_s": wall}
    del bufs, ws, gws, key, graph, ggraph
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_chain(moduli, logn, [int(x) for x in levels], args.cpu_seconds)
    return out


def cpu_baseline_chain(moduli, logn, levels, budget_s):
    """The chain's pairs in order (their levels as drawn for the GPU batch), one pair per
    call: 3 x (multiply, relinearize, rescale) through the reference CPU path
    (src/evaluator.cpp), injected uniform words."""
    import numpy as np
    import checkers as ck
    n, K = 1 << logn, len(moduli)
    rng = np.random.default_rng(19)
    key = ck.random_key(rng, moduli, logn)
    top = max(levels)
    xs = ck.uniform_residues(rng, (2, top, n), moduli[:top])
    ys = ck.uniform_residues(rng, (2, top, n), moduli[:top])

    def make(kind):
        state = {"k": 0}
        sess = ck.RefSession(CKKS, n, moduli) if kind == "reference" else None

        def run():
            l = levels[state["k"] % len(levels)]
            state["k"] += 1
            cur = np.ascontiguousarray(xs[:, :l])
            for s in range(3):
                ll = l - s
                y = np.ascontiguousarray(ys[:, :ll])
                if sess is not None:
                    cur = sess.rescale(sess.relinearize(sess.multiply(cur, y, ll), key, ll), ll)
                else:
                    o2 = ck.relinearize(CKKS, ck.ckks_multiply(cur, y, moduli[:ll], logn), key, moduli, ll, logn)
                    cur = ck.rescale(CKKS, o2, moduli, logn)
        return run

    return cpu_baseline(make, "mul+relin+rescale/s", 3, budget_s,
                        "one pair per call (levels as drawn for the GPU batch, U[4,15]): "
                        "3 x (multiply, relinearize, rescale), N=2^15")


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
    # dominant kernel: the first pass of hoisted products (4 rotations over the shared
    # ModUp limbs), bracketed by HIP events on the library's launch stream
    probes = [(sealhip.HipEvent(), sealhip.HipEvent()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    with Timer(torch) as th:
        for s_ in range(args.steps):
            ctx.probe_keyswitch(*probes[s_])
            hoisted()
    wall = time.perf_counter() - t0
    prod_ms = sum(b.elapsed_ms(e) for b, e in probes) / args.steps
    separate()
    torch.cuda.synchronize()
    if not torch.equal(out, acc):
        raise SystemExit("bench self-check failed: hoisted rotation sum differs from separate rotations")
    with Timer(torch) as ts:
        separate()
    ms = th.ms / args.steps
    rots = cts * len(steps)
    limb = n * 8
    # step (per ciphertext: 2L limbs in, 2L out; each prepared key once: 2L(L+1) key' limbs
    # + 2(L+1) correction limbs)
    key_limbs = 2 * L * (L + 1) + 2 * (L + 1)
    step_bytes = (cts * 4 * L + len(steps) * key_limbs) * limb
    # products pass, R = 4 rotations: the L x L ModUp limbs and L own-prime input limbs per
    # ciphertext, R prepared keys, R x 2(L+1) product limbs per ciphertext
    R = 4
    prod_bytes = (cts * (L * L + L + R * 2 * (L + 1)) + R * key_limbs) * limb
    out_json = {"metric": "CKKS rotations/s, N=2^16, 16 ct x 15 rotations summed, hoisted",
                "value": rots / (ms * 1e-3), "unit": "rotations/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "u64",
                "data": "synthetic (seeded uniform residues; injected Galois keys)",
                "config": {"workload": "hoisted rotations: N=2^16, 24+1 primes (60, 50x23, 60), 16 ciphertexts, "
                                       "rotations by 2^k, k=0..14, summed",
                           "parallelism": "single GPU"},
                "roofline": {"bound": "hbm", "achieved": prod_bytes / (prod_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": prod_bytes / (prod_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "traffic": None,
                             "kernel": "rotation_products_kernel<4>: products of 4 rotations' prepared keys over "
                                       "the shared ModUp limbs (first pass of each step)",
                             "algorithmic_bytes": prod_bytes, "avg_ms": prod_ms},
                "step_roofline": {"achieved": step_bytes / (ms * 1e-3) / 1e9, "unit": "GB/s",
                                  "frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "algorithmic_bytes_per_step": step_bytes},
                "separate_ms_per_step": ts.ms, "hoisted_over_separate": ts.ms / ms,
                "self_check": "hoisted words == sum of 15 separate rotations", "wall_s": wall}
    del keys, hkeys, ws, x, out, tmp, acc
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out_json["cpu_baseline"] = cpu_baseline_rotation(moduli, logn, steps, args.cpu_seconds)
    return out_json


def cpu_baseline_rotation(moduli, logn, steps, budget_s):
    """Rotations at N=2^16 with 24 + 1 primes through the reference CPU path
    (Evaluator::applyGaloisInplace, src/evaluator.cpp:2153-2307: one full key switch per
    rotation; the reference has no hoisting), one rotation per call, cycling through the
    15 steps, injected uniform words (one key shared by the steps)."""
    import numpy as np
    import checkers as ck
    n, K = 1 << logn, len(moduli)
    L = K - 1
    rng = np.random.default_rng(29)
    ct = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    key = ck.random_key(rng, moduli, logn)
    elts = [ck.galois_elt(logn, s) for s in steps]

    def make(kind):
        state = {"k": 0}
        sess = ck.RefSession(CKKS, n, moduli) if kind == "reference" else None

        def run():
            e = elts[state["k"] % len(elts)]
            state["k"] += 1
            if sess is not None:
                sess.apply_galois(ct, e, key, L)
            else:
                ck.apply_galois_keyswitch(CKKS, ct, e, key, moduli, L, logn)
        return run

    return cpu_baseline(make, "rotations/s", 1, budget_s,
                        f"one rotation (applyGaloisInplace) per call at N=2^{logn}, L={L} (+1 special)")


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


class MatvecResident:
    """One request type of the serving workload, its device-resident part: a context, one
    matrix's diagonals (NTT form, synthetic residues), Galois keys (synthetic residues;
    the baby steps also prepared for hoisting at the ciphertexts' level) and a host pool
    of input ciphertexts.  Requests run through the library's server (shc_server_*)."""

    def __init__(self, torch, sealhip, scheme, n, moduli, plain, b, g, pool, gen, st):
        K = len(moduli)
        L = K - 1
        self.n, self.L, self.b, self.g = n, L, b, g
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
        x = torch.empty((pool, 2, L, n), dtype=torch.int64, device="cuda")
        fill_uniform(torch, x, moduli[:L], 2, gen)
        torch.cuda.synchronize()
        self.inputs = [r.reshape(-1).numpy().view("uint64").copy() for r in x.cpu()]
        self.resident_bytes = (self.plains.numel() + sum(k.numel() for k in self.baby + self.giant)
                               + sum(h.numel() for h in self.hoisted)) * 8

    def register(self, server):
        return server.add_matvec(self.ctx, self.L, self.b, self.g, self.plains.data_ptr(),
                                 [k.data_ptr() for k in self.baby], [h.data_ptr() for h in self.hoisted],
                                 [k.data_ptr() for k in self.giant])


def pinned_copy(server, arrays):
    """Copies of host arrays in the server's pinned memory (used in place by requests)."""
    out = []
    for a in arrays:
        p = server.host_array(a.size)
        p[:] = a
        out.append(p)
    return out


def serve_open_loop(server, kinds, rate, requests, seed, ring=64, burst=False):
    """Open-loop serving through the library's server.  kinds: [(type id, [host input
    arrays])]; request r picks a kind uniformly at random and an input from its list, and
    arrives at the r-th time of a Poisson process of `rate` per second (burst: all at
    once).  It is submitted at its arrival time (submit copies the input into the server's
    pinned staging), and its latency runs from arrival to the server's completion, when
    the result has been copied back into host memory -- upload, queueing, batching,
    device work and download included.  At most `ring` requests are outstanding (their
    output buffers, pinned memory of the server, are reused)."""
    import collections
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.zeros(requests) if burst else np.cumsum(rng.exponential(1.0 / rate, size=requests))
    kind = rng.integers(0, len(kinds), size=requests)
    pick = rng.integers(0, 1 << 30, size=requests)
    outs = [[server.host_array(server.request_size(t)[1]) for _ in range(ring)] for t, _ in kinds]
    lat = np.empty(requests)
    bsz = np.empty(requests)
    done_at = np.empty(requests)
    pending = collections.deque()

    def reap(block):
        while pending:
            r, rid = pending[0]
            info = server.wait(rid, -1.0 if block else 0.0)
            if info is None:
                return
            pending.popleft()
            lat[r] = info.done_s - (t0 + arrivals[r])
            done_at[r] = info.done_s
            bsz[r] = info.batch_size
            if block:
                return

    t0 = server.clock() + 0.005
    for r in range(requests):
        while len(pending) >= ring:
            reap(True)
        target = t0 + arrivals[r]
        while True:
            now = server.clock()
            if now >= target:
                break
            reap(False)
            time.sleep(min(target - now, 5e-4))
        t, pool = kinds[kind[r]]
        pending.append((r, server.submit(t, pool[pick[r] % len(pool)], outs[kind[r]][r % ring])))
    while pending:
        reap(True)
    span = float(done_at.max() - t0)
    res = {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "served_per_s": requests / span, "mean_batch": float(bsz.mean()), "requests": requests}
    if len(kinds) > 1:
        res["per_type"] = [{"requests": int((kind == k).sum()),
                            "p50_ms": float(np.percentile(lat[kind == k], 50) * 1e3),
                            "p99_ms": float(np.percentile(lat[kind == k], 99) * 1e3),
                            "mean_batch": float(bsz[kind == k].mean())} for k in range(len(kinds))]
    return res


def bench_serve(args):
    """The serving workload: a Poisson stream of matrix-vector requests, half BFV (N=8192,
    BFVDefault(8192) primes, t = 65537, a 4096 x 4096 matrix mod t, BSGS 64 x 64 in each
    row of the slot matrix) and half CKKS (N=2^15, Create(32768, {60, 40 x 14, 60}), a
    1024 x 1024 matrix, BSGS 32 x 32), keys and diagonals resident (synthetic residues),
    requests arriving as host ciphertexts.  The library's continuous-batching server
    (2 lanes, batches of at most 16, a 24 GiB device arena) against the plain server (one
    lane, one request at a time) under the same arrivals, at half the batched server's
    capacity for the mix (measured first with a burst); value = requests/s served by the
    batched server, with p50/p99 latency from arrival to the result in host memory.
    Self-check: for each type, a request's words in a burst equal its words served alone."""
    import numpy as np
    import torch
    import sealhip
    sealhip.init(0)
    st = torch.cuda.current_stream().cuda_stream
    gen = torch.Generator(device="cuda").manual_seed(8)
    B, lanes, arena = 16, 2, 24 << 30
    bfv_moduli = [0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001]
    ckks_moduli = sealhip.coeff_modulus_create(1 << 15, [60] + [40] * 14 + [60])
    res = [MatvecResident(torch, sealhip, sealhip.SCHEME_BFV, 8192, bfv_moduli, 65537, 64, 64, 8, gen, st),
           MatvecResident(torch, sealhip, sealhip.SCHEME_CKKS, 1 << 15, ckks_moduli, 0, 32, 32, 8, gen, st)]
    ring = 64
    staging = 8 * max(2 * r.inputs[0].size * 8 for r in res)  # the self-check's unpinned outputs
    batched_sv = sealhip.Server(arena, staging, lanes=lanes, max_batch=B)
    plain_sv = sealhip.Server(arena // 4, staging, lanes=1, max_batch=1)
    kinds_b = [(r.register(batched_sv), pinned_copy(batched_sv, r.inputs)) for r in res]
    kinds_p = [(r.register(plain_sv), pinned_copy(plain_sv, r.inputs)) for r in res]
    # warm up, then capacity per type with a burst of 4 full batches
    for t, pool in kinds_b:
        serve_open_loop(batched_sv, [(t, pool)], 0, 2 * B, seed=1, burst=True)
    cap = [serve_open_loop(batched_sv, [k], 0, 4 * B, seed=2, burst=True)["served_per_s"] for k in kinds_b]
    mix_cap = 1.0 / (0.5 / cap[0] + 0.5 / cap[1])
    # determinism: a request's words in a burst equal its words alone
    for (t, pool), (tp, _), name in zip(kinds_b, kinds_p, ("bfv", "ckks")):
        words = batched_sv.request_size(t)[1]
        outs = [np.empty(words, dtype=np.uint64) for _ in range(B)]
        ids = [batched_sv.submit(t, pool[i % len(pool)], outs[i]) for i in range(B)]
        infos = [batched_sv.wait(i) for i in ids]
        alone = np.empty(words, dtype=np.uint64)
        for i in (0, B - 1):
            plain_sv.wait(plain_sv.submit(tp, pool[i % len(pool)], alone))
            if not np.array_equal(alone, outs[i]):
                raise SystemExit(f"bench self-check failed: {name} request {i} differs alone and in a batch")
        if max(inf.batch_size for inf in infos) < 2:
            raise SystemExit(f"bench self-check failed: the {name} burst was never batched")
    rate = 0.5 * mix_cap
    reqs = max(200, int(rate * args.serve_seconds))
    # the dominant kernel of the mix: the inner sums of a full CKKS batch (every giant
    # step's plain-product sum over the b baby rotations, shc_multiply_plain_accumulate_many)
    ck_res = res[1]
    Lc, nc, bb, gg = ck_res.L, ck_res.n, ck_res.b, ck_res.g
    babies = torch.empty((bb, B, 2, Lc, nc), dtype=torch.int64, device="cuda")
    fill_uniform(torch, babies, ckks_moduli[:Lc], 3, gen)
    inner = torch.empty((gg, B, 2, Lc, nc), dtype=torch.int64, device="cuda")
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    sums_ms = []
    for _ in range(max(args.steps, 1) + 1):
        ev[0].record()
        ck_res.ctx.multiply_plain_accumulate_many(babies.data_ptr(), ck_res.plains.data_ptr(), bb, gg, B * 2, Lc,
                                                  inner.data_ptr(), st)
        ev[1].record()
        torch.cuda.synchronize()
        sums_ms.append(ev[0].elapsed_time(ev[1]))
    sums_ms = sum(sums_ms[1:]) / len(sums_ms[1:])
    del babies, inner
    # reads the b baby rotations of B requests (2L limbs each) and the g x b diagonals
    # (L limbs each), writes g inner sums per request (2L limbs)
    sums_bytes = (bb * B * 2 * Lc + gg * bb * Lc + gg * B * 2 * Lc) * nc * 8
    batched = serve_open_loop(batched_sv, kinds_b, rate, reqs, seed=11)
    plain = serve_open_loop(plain_sv, kinds_p, rate, reqs, seed=11)
    stats = batched_sv.stats()
    out = {"metric": "mixed BFV/CKKS matvec requests/s served; p50/p99 latency", "value": batched["served_per_s"],
           "unit": "requests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 / batched["served_per_s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u64",
           "data": "synthetic (seeded uniform residues for ciphertexts, diagonals and keys; Poisson arrivals)",
           "config": {"workload": "serving: 50% BFV matvec (N=8192, BFVDefault(8192), t=65537, 4096x4096, "
                                  "BSGS 64x64), 50% CKKS matvec (N=2^15, 60+40x14+60, 1024x1024, BSGS 32x32); "
                                  f"library server, {lanes} lanes, batches of at most {B}, "
                                  f"{arena >> 30} GiB device arena; host ciphertexts in and out",
                      "parallelism": "single GPU"},
           "capacity_per_s": {"bfv": cap[0], "ckks": cap[1], "mix": mix_cap},
           "admission": {"device_arena_bytes": arena,
                         "device_bytes_of_one_request": {"bfv": batched_sv.request_size(kinds_b[0][0])[2],
                                                         "ckks": batched_sv.request_size(kinds_b[1][0])[2]},
                         "resident_key_and_diagonal_bytes": {"bfv": res[0].resident_bytes,
                                                             "ckks": res[1].resident_bytes},
                         "max_batches_in_flight": stats["max_batches_in_flight"],
                         "peak_device_bytes": stats["peak_device_bytes"]},
           "arrivals": {"kind": "poisson", "rate_per_s": rate, "requests": reqs},
           "batched": batched, "plain": plain, "batched_p99_over_plain_p99": batched["p99_ms"] / plain["p99_ms"],
           "roofline": {"bound": "hbm", "achieved": sums_bytes / (sums_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": sums_bytes / (sums_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "kernel": f"plain_dot_many_kernel: the {gg} inner sums of a full CKKS batch ({B} requests, "
                                  f"{bb} baby rotations, L={Lc})",
                        "algorithmic_bytes": sums_bytes, "avg_ms": sums_ms},
           "self_check": "per type, the first and last request of a burst give the same words served alone"}
    batched_sv.close()
    plain_sv.close()
    del res
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_serve(bfv_moduli, ckks_moduli, args.cpu_seconds)
    return out


def cpu_baseline_serve(bfv_moduli, ckks_moduli, budget_s):
    """The mixed matvec request stream through the reference CPU path, by sampling: one
    call does 1/32 of a CKKS request (N=2^15, L=15: 2 of its 62 rotations, 32 of its
    1024 NTT-form plain products and additions) and 1/32 of a BFV request (N=8192, L=4:
    4 of its 126 rotations, 128 of its 4096 plain products and additions), i.e. 1/16 of a
    request of the half-and-half mix; the reference evaluates a matvec as rotations,
    plain products and additions (app/matmul.cuh:98-121), with no hoisting."""
    import numpy as np
    import checkers as ck
    rng = np.random.default_rng(31)
    nb, nc = 8192, 1 << 15
    Lb, Lc = len(bfv_moduli) - 1, len(ckks_moduli) - 1
    ct_b = ck.uniform_residues(rng, (2, Lb, nb), bfv_moduli[:Lb])
    key_b = ck.random_key(rng, bfv_moduli, 13)
    pl_b = rng.integers(0, T_BFV, nb, dtype=np.uint64)
    ct_c = ck.uniform_residues(rng, (2, Lc, nc), ckks_moduli[:Lc])
    key_c = ck.random_key(rng, ckks_moduli, 15)
    pl_c = ck.uniform_residues(rng, (Lc, nc), ckks_moduli[:Lc])
    e_b, e_c = ck.galois_elt(13, 1), ck.galois_elt(15, 1)

    def make(kind):
        if kind == "reference":
            sb = ck.RefSession(1, nb, bfv_moduli, T_BFV)
            sc = ck.RefSession(CKKS, nc, ckks_moduli)
            rot_b = lambda: sb.apply_galois(ct_b, e_b, key_b, Lb)
            rot_c = lambda: sc.apply_galois(ct_c, e_c, key_c, Lc)
            mul_b = lambda: sb.plain_op(2, ct_b, pl_b, Lb)
            mul_c = lambda: sc.plain_op(2, ct_c, pl_c, Lc)
            add_b = lambda: sb.ct_binary(0, ct_b, ct_b, Lb)
            add_c = lambda: sc.ct_binary(0, ct_c, ct_c, Lc)
        else:
            rot_b = lambda: ck.apply_galois_keyswitch(1, ct_b, e_b, key_b, bfv_moduli, Lb, 13)
            rot_c = lambda: ck.apply_galois_keyswitch(CKKS, ct_c, e_c, key_c, ckks_moduli, Lc, 15)
            mul_b = lambda: ck.bfv_multiply_plain(ct_b, pl_b, bfv_moduli[:Lb], T_BFV, 13)
            mul_c = lambda: ck.multiply_plain_ntt(ct_c, pl_c, ckks_moduli, 15)
            add_b = lambda: ck.mod_add(ct_b, ct_b, bfv_moduli)
            add_c = lambda: ck.mod_add(ct_c, ct_c, ckks_moduli)

        def run():
            for _ in range(2):
                rot_c()
            for _ in range(32):
                mul_c()
                add_c()
            for _ in range(4):
                rot_b()
            for _ in range(128):
                mul_b()
                add_b()
        return run

    return cpu_baseline(make, "requests/s", 1.0 / 16, budget_s,
                        "1/32 of a CKKS request + 1/32 of a BFV request per call (rotations, plain products and "
                        "additions in the request's proportions)")


def bench_bfv8192(args):
    """The CPU-runnable BFV path: N=8192, BFVDefault(8192) primes {0x7fffffd8001,
    0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001} (4 data + special),
    t = 65537, real keys (device key generation).  Two vectors of 8192 slots uniform mod t
    are batch-encoded and encrypted once; a step is multiply -> relinearize -> decrypt ->
    batch-decode of that pair (one request, latency-bound), timed eagerly and then replayed
    from a captured HIP graph (value: the graph replays).  Self-checks: the decoded
    slots are the slot-wise products mod t, and the multiply + relinearize words equal the
    oracle's on the same ciphertexts and key."""
    import numpy as np
    import torch
    import sealhip
    import checkers as ck
    sealhip.init(0)
    logn, n, t = 13, 8192, T_BFV
    moduli = [0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001]
    K = len(moduli)
    L = K - 1
    ctx = sealhip.Context(sealhip.SCHEME_BFV, n, moduli, t)
    st = torch.cuda.current_stream().cuda_stream
    dev = lambda shape: torch.empty(shape, dtype=torch.int64, device="cuda")
    wsb = lambda nbytes: dev(((max(nbytes, 8) + 7) // 8,))
    sk, pk, rk = dev((K, n)), dev((2, K, n)), dev((K - 1, 2, K, n))
    ctx.keygen_secret(1, sk.data_ptr(), st)
    ctx.keygen_public(sk.data_ptr(), 2, pk.data_ptr(), st)
    kw = ctx.keygen_workspace_size()
    ctx.keygen_relin(sk.data_ptr(), 3, rk.data_ptr(), wsb(kw).data_ptr(), kw, st)
    rng = np.random.default_rng(1)
    x = rng.integers(0, t, n, dtype=np.uint64)
    y = rng.integers(0, t, n, dtype=np.uint64)
    px, py = dev((n,)), dev((n,))
    for v, p_ in ((x, px), (y, py)):
        dv = torch.from_numpy(v.view(np.int64).copy()).cuda()
        ctx.batch_encode(dv.data_ptr(), n, p_.data_ptr(), st)
    cts = dev((2, 2, L, n))
    ew = ctx.encrypt_workspace_size(L)
    eb = wsb(ew)
    t0 = time.perf_counter()
    ctx.encrypt(pk.data_ptr(), px.data_ptr(), L, 7, cts[0].data_ptr(), eb.data_ptr(), ew, st)
    ctx.encrypt(pk.data_ptr(), py.data_ptr(), L, 8, cts[1].data_ptr(), eb.data_ptr(), ew, st)
    torch.cuda.synchronize()
    enc_ms = (time.perf_counter() - t0) * 1e3
    prod = dev((1, 3, L, n))
    mw, kw1, dw = ctx.multiply_workspace_size(1, L), ctx.keyswitch_workspace_size(1, L), ctx.decrypt_workspace_size(L)
    mb, kb, db = wsb(mw), wsb(kw1), wsb(dw)
    out, vals, scratch = dev((n,)), dev((n,)), dev((n,))

    def step(probe=None, stream=None):
        s_ = stream or st
        ctx.multiply(cts[0].data_ptr(), cts[1].data_ptr(), prod.data_ptr(), 1, L, mb.data_ptr(), mw, s_)
        if probe:
            ctx.probe_keyswitch(*probe)
        ctx.relinearize(prod.data_ptr(), rk.data_ptr(), 1, L, kb.data_ptr(), kw1, s_)
        ctx.decrypt(prod.data_ptr(), 2, sk.data_ptr(), L, out.data_ptr(), db.data_ptr(), dw, s_)
        ctx.batch_decode(out.data_ptr(), vals.data_ptr(), scratch.data_ptr(), s_)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    steps = max(args.steps, 20)
    probes = [(sealhip.HipEvent(), sealhip.HipEvent()) for _ in range(steps)]
    with Timer(torch) as tm:
        for i in range(steps):
            step(probes[i])
    ms = tm.ms / steps
    rows_ms = sum(b.elapsed_ms(e) for b, e in probes) / steps
    want = (x.astype(object) * y.astype(object) % t).astype(np.uint64)
    if not np.array_equal(vals.cpu().numpy().view(np.uint64), want):
        raise SystemExit("bench self-check failed: BFV N=8192 pipeline does not decrypt to the slot products")
    c = cts.cpu().numpy().view(np.uint64)
    o2 = ck.relinearize(1, ck.bfv_multiply(c[0], c[1], moduli[:L], t, logn), rk.cpu().numpy().view(np.uint64),
                        moduli, L, logn)
    if not np.array_equal(prod.cpu().numpy().view(np.uint64)[0, :2], o2):
        raise SystemExit("bench self-check failed: BFV N=8192 multiply+relinearize words differ from the oracle")
    # the same pipeline captured in a HIP graph (its ~30 launches replayed as one): every
    # replay, with eager work between, must decode to the slot products
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(graph, stream=cap):
            step(stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(cap)
    want_t = torch.from_numpy(want.view(np.int64).copy()).cuda()
    for _ in range(3):
        vals.zero_()
        graph.replay()
        torch.cuda.synchronize()
        if not torch.equal(vals, want_t):
            raise SystemExit("bench self-check failed: a graph replay of the BFV pipeline lost the slot products")
        step()
        torch.cuda.synchronize()
    with Timer(torch) as tg:
        for _ in range(steps):
            graph.replay()
    graph_ms = tg.ms / steps
    eager_ms = ms
    ms = graph_ms
    limb = n * 8
    # the probed launch: the ModUp rows pass over the L(L+1) limbs of one ciphertext (read
    # and written once) and the L+1 twiddle tables
    rows_bytes = 2 * L * (L + 1) * limb + (L + 1) * n * 16
    out_json = {"metric": "BFV multiply+relinearize+decrypt pipelines/s, N=8192 (latency)",
                "value": 1e3 / ms, "unit": "pipelines/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
                "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "u64", "data": "synthetic (seeded slots mod t; device-generated keys)",
                "config": {"workload": "BFV N=8192, BFVDefault(8192) primes (4 data + special), t=65537: one "
                                       "multiply -> relinearize -> decrypt -> batch decode per step, captured in a "
                                       "HIP graph and replayed",
                           "parallelism": "single GPU"},
                "latency_ms": ms, "eager_latency_ms": eager_ms, "graph_over_eager": eager_ms / ms,
                "encrypt_two_ms": enc_ms,
                "roofline": {"bound": "hbm", "achieved": rows_bytes / (rows_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": rows_bytes / (rows_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "traffic": None,
                             "kernel": "ntt_rows_modup_kernel<13,...>: the key switch's ModUp rows pass "
                                       f"({L * (L + 1)} limbs; a latency-bound launch)",
                             "algorithmic_bytes": rows_bytes, "avg_ms": rows_ms},
                "self_check": "decrypts to the slot products mod t (eager and every graph replay); "
                              "multiply+relinearize words == oracle"}
    if not args.no_cpu:
        ct2 = [np.ascontiguousarray(c[0]), np.ascontiguousarray(c[1])]
        key_h = rk.cpu().numpy().view(np.uint64).copy()
        sk_h = sk.cpu().numpy().view(np.uint64).copy()

        def make(kind):
            if kind == "reference":
                sess = ck.RefSession(1, n, moduli, t)
                return lambda: sess.batch_decode(sess.decrypt(sess.relinearize(sess.multiply(ct2[0], ct2[1], L),
                                                                               key_h, L), sk_h, L))
            return lambda: ck.batch_decode(ck.bfv_decrypt(ck.relinearize(1, ck.bfv_multiply(ct2[0], ct2[1],
                                                                                             moduli[:L], t, logn),
                                                                         key_h, moduli, L, logn),
                                                          sk_h, moduli, t, logn), t, logn)

        out_json["cpu_baseline"] = cpu_baseline(make, "pipelines/s", 1, args.cpu_seconds,
                                                "multiply -> relinearize -> decrypt -> batch decode of the same "
                                                "ciphertexts and keys, N=8192")
    return out_json


def bench_primitives(args):
    """The path's other entry points, each timed alone with its inputs resident in HBM.

    CKKS at the headline shape (N=2^16, CoeffModulus::Create(65536, {60 x 25}), level 24,
    128 ciphertexts): one rotation per ciphertext (shc_apply_galois_keyswitch), rescale to
    level 23, public-key encryption of 128 NTT-form plaintexts (shc_encrypt_many),
    decryption (shc_decrypt_many), CKKSEncoder encode and decode of 128 x 32768 slots
    (shc_ckks_encode_many / shc_ckks_decode_many).  BFV at the CPU-runnable shape (N=8192,
    BFVDefault(8192), t=65537, 128 ciphertexts): encryption, decryption and the
    BatchEncoder (batched; the BatchEncoder also one call per plaintext, as the
    reference's).  Each operation reports
    its rate, the whole call's algorithmic bytes over its time against the HBM peak, and
    the reference CPU path on the same operation (its own randomness for encryption).
    Self-checks: BFV encode -> encrypt -> decrypt -> decode returns the slots; CKKS
    decrypt(encrypt(m)) - m is small in every coefficient of its first limb; CKKS
    decode(encode(x)) is within 2^-20 of x."""
    import numpy as np
    import torch
    import sealhip
    import checkers as ck
    sealhip.init(0)
    st = torch.cuda.current_stream().cuda_stream
    dev = lambda shape: torch.empty(shape, dtype=torch.int64, device="cuda")
    wsb = lambda nbytes: dev(((max(nbytes, 8) + 7) // 8,))
    B = 128
    steps, warmup = max(args.steps, 3), max(args.warmup, 1)
    cpu_s = args.cpu_seconds / 3
    ops = {}

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        with Timer(torch) as tm:
            for _ in range(steps):
                fn()
        return tm.ms / steps

    def record(name, metric, unit, units, ms, nbytes, workload, entry, make=None, sample=None):
        gbs = nbytes / (ms * 1e-3) / 1e9
        d = {"metric": metric, "value": units / (ms * 1e-3), "unit": unit, "ms_per_step": ms, "steps": steps,
             "config": {"workload": workload, "parallelism": "single GPU"},
             "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": gbs / HBM_PEAK_GBS, "traffic": None, "kernel": f"{entry}: the whole call",
                          "algorithmic_bytes": nbytes, "avg_ms": ms}}
        if make is not None and not args.no_cpu:
            d["cpu_baseline"] = cpu_baseline(make, unit, 1, cpu_s, sample)
        ops[name] = d
        print(f"[bench] primitives: {name} {d['value']:.1f} {unit}", file=sys.stderr, flush=True)

    # ---------------------------------------------------------------- CKKS, N = 2^16, L = 24
    logn, n = 16, 1 << 16
    moduli = sealhip.coeff_modulus_create(n, [60] * 25)
    K, L = len(moduli), len(moduli) - 1
    limb = n * 8
    shape = f"N=2^16, 24+1 primes of 60 bits, level 24, batch of {B}"
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    gen = torch.Generator(device="cuda").manual_seed(17)
    rng = np.random.default_rng(17)
    h_ct = ck.uniform_residues(rng, (2, L, n), moduli[:L])  # CPU-baseline inputs
    h_key = ck.random_key(rng, moduli, logn)

    x = dev((B, 2, L, n))
    fill_uniform(torch, x, moduli[:L], 2, gen)
    key = dev((K - 1, 2, K, n))
    fill_uniform(torch, key, moduli, 2, gen)
    elt = ctx.galois_elt_from_step(1)
    gw = ctx.galois_workspace_size(B, L)
    gb = wsb(gw)
    ms = timed(lambda: ctx.apply_galois_keyswitch(x.data_ptr(), key.data_ptr(), elt, B, L, gb.data_ptr(), gw, st))
    sess = {}

    def ref_sess(kind, scheme, nn, mods, t=0):
        if kind != "reference":
            return None
        return ck.RefSession(scheme, nn, mods, t)

    def make_rot(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.apply_galois(h_ct, ck.galois_elt(logn, 1), h_key, L)
        return lambda: ck.apply_galois_keyswitch(CKKS, h_ct, ck.galois_elt(logn, 1), h_key, moduli, L, logn)

    record("ckks_rotate", "CKKS rotations/s (one key switch each), " + shape, "rotations/s", B, ms,
           (B * 4 * L + 2 * L * (L + 1)) * limb, "CKKS rotation by one slot of each ciphertext, " + shape,
           "shc_apply_galois_keyswitch", make_rot, "one rotation (applyGaloisInplace) per call")
    del gb, key
    torch.cuda.empty_cache()

    out = dev((B, 2, L - 1, n))
    rw = ctx.rescale_workspace_size(B, 2, L)
    rb = wsb(rw)
    ms = timed(lambda: ctx.rescale_to_next(x.data_ptr(), out.data_ptr(), B, 2, L, rb.data_ptr(), rw, st))

    def make_rescale(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.rescale(h_ct, L)
        return lambda: ck.rescale(CKKS, h_ct, moduli[:L], logn)

    record("ckks_rescale", "CKKS rescales/s, " + shape, "rescales/s", B, ms, B * (2 * L + 2 * (L - 1)) * limb,
           "CKKS rescale to the next level, " + shape, "shc_rescale_to_next", make_rescale,
           "one rescaleToNext per call")
    del out, rb
    torch.cuda.empty_cache()

    # ciphertext addition and plaintext multiplication (NTT form), HBM-bound
    y = dev((B, 2, L, n))
    fill_uniform(torch, y, moduli[:L], 2, gen)
    pl1 = dev((L, n))
    fill_uniform(torch, pl1, moduli[:L], 0, gen)
    zo = dev((B, 2, L, n))
    ms = timed(lambda: ctx.add(x.data_ptr(), y.data_ptr(), zo.data_ptr(), 2 * B, L, st))
    h_y = ck.uniform_residues(rng, (2, L, n), moduli[:L])
    h_pl = ck.uniform_residues(rng, (L, n), moduli[:L])

    def make_add(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.ct_binary(0, h_ct, h_y, L)
        return lambda: ck.mod_add(h_ct, h_y, moduli)

    record("ckks_add", "CKKS ciphertext additions/s, " + shape, "additions/s", B, ms, B * 6 * L * limb,
           "CKKS ciphertext + ciphertext, " + shape, "shc_add", make_add, "one Evaluator::add per call")
    ms = timed(lambda: ctx.multiply_plain_ntt(x.data_ptr(), pl1.data_ptr(), zo.data_ptr(), 2 * B, L, st))

    def make_mulp(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.plain_op(2, h_ct, h_pl, L)
        return lambda: ck.multiply_plain_ntt(h_ct, h_pl, moduli, logn)

    record("ckks_multiply_plain", "CKKS ciphertext x plaintext products/s, " + shape, "products/s", B, ms,
           (B * 4 * L + L) * limb, "CKKS ciphertext x NTT-form plaintext, " + shape, "shc_multiply_plain_ntt",
           make_mulp, "one Evaluator::multiplyPlain per call")
    del y, zo, pl1
    torch.cuda.empty_cache()

    # real keys for the round trip
    sk, pk = dev((K, n)), dev((2, K, n))
    ctx.keygen_secret(5, sk.data_ptr(), st)
    ctx.keygen_public(sk.data_ptr(), 6, pk.data_ptr(), st)
    plain = dev((B, L, n))
    fill_uniform(torch, plain, moduli[:L], 1, gen)
    ew = ctx.encrypt_many_workspace_size(L, B)
    eb = wsb(ew)
    ms = timed(lambda: ctx.encrypt_many(pk.data_ptr(), plain.data_ptr(), B, L, 99, x.data_ptr(), eb.data_ptr(), ew,
                                        st))
    h_pk = pk.cpu().numpy().view(np.uint64).copy()
    h_plain = plain[0].cpu().numpy().view(np.uint64).copy()

    def make_enc(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.encrypt(h_pk, h_plain, L)
        u = ck.rns_of_small(np.random.default_rng(1).integers(-1, 2, size=n), moduli)
        e = ck.rns_of_small(np.random.default_rng(2).integers(-3, 4, size=n), moduli)

        def run():
            z = ck.encrypt_zero_asymmetric(CKKS, h_pk, u, e, e, moduli, logn)
            ck.mod_add(ck.rescale(CKKS, z, moduli, logn)[0], h_plain, moduli)
        return run

    record("ckks_encrypt", "CKKS public-key encryptions/s, " + shape, "encryptions/s", B, ms,
           (2 * K + B * 3 * L) * limb, "CKKS encryption of NTT-form plaintexts with the public key, " + shape,
           "shc_encrypt_many", make_enc, "one Encryptor::encrypt per call")
    del eb
    torch.cuda.empty_cache()

    dec = dev((B, L, n))
    ms = timed(lambda: ctx.decrypt_many(x.data_ptr(), 2, B, sk.data_ptr(), L, dec.data_ptr(), 0, 0, st))
    # self-check: decrypt(encrypt(m)) - m is a small error (first limb, coefficient form)
    diff = dev((1, L, n))
    ctx.sub(dec[0].data_ptr(), plain[0].data_ptr(), diff.data_ptr(), 1, L, st)
    ctx.ntt_inverse(diff.data_ptr(), 1, L, st)
    d0 = diff[0, 0].cpu().numpy().view(np.uint64).astype(object)
    if max(min(int(v), moduli[0] - int(v)) for v in d0) >= (1 << 30):
        raise SystemExit("bench self-check failed: CKKS decrypt(encrypt(m)) - m is not small")
    h_sk = sk.cpu().numpy().view(np.uint64).copy()

    def make_dec(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.decrypt(h_ct, h_sk, L)
        return lambda: ck.decrypt_phase(CKKS, h_ct, h_sk[:L], moduli, logn)

    record("ckks_decrypt", "CKKS decryptions/s, " + shape, "decryptions/s", B, ms, (B * 3 * L + L) * limb,
           "CKKS decryption to the NTT-form plaintext, " + shape, "shc_decrypt_many", make_dec,
           "one Decryptor::decrypt per call")
    del dec, diff, x, sk, pk
    torch.cuda.empty_cache()

    half = n // 2
    vals = torch.rand((B, n), dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    cw = ctx.ckks_encode_many_workspace_size(B)
    cb = wsb(cw)
    status = torch.zeros((B,), dtype=torch.int32, device="cuda")
    scale = 2.0 ** 40
    ms = timed(lambda: ctx.ckks_encode_many(vals.data_ptr(), half, n, B, scale, L, plain.data_ptr(),
                                            status.data_ptr(), cb.data_ptr(), cw, st))
    h_vals = vals[0].cpu().numpy()
    h_cplx = h_vals[0::2] + 1j * h_vals[1::2]

    def make_enc_ckks(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.ckks_encode(h_cplx, L, scale)
        return lambda: ck.ckks_encode(h_cplx, moduli[:L], scale, logn)

    slot_bytes = B * (half * 16 + L * limb)
    record("ckks_encode", "CKKSEncoder encodes/s, N/2 = 32768 slots each, " + shape, "encodes/s", B, ms,
           slot_bytes, "CKKSEncoder encode of 32768 complex slots to an NTT-form plaintext at level 24, batch of "
           f"{B}", "shc_ckks_encode_many", make_enc_ckks, "one CKKSEncoder::encode per call")
    del cb
    back = torch.empty((B, n), dtype=torch.float64, device="cuda")
    dw = ctx.ckks_decode_many_workspace_size(L, B)
    db = wsb(dw)
    ms = timed(lambda: ctx.ckks_decode_many(plain.data_ptr(), L, B, scale, back.data_ptr(), db.data_ptr(), dw, st))
    if int(status.abs().sum()) or float((back - vals).abs().max()) > 2.0 ** -20:
        raise SystemExit("bench self-check failed: CKKS decode(encode(x)) differs from x")
    h_plain = plain[0].cpu().numpy().view(np.uint64).copy()

    def make_dec_ckks(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return lambda: s_.ckks_decode(h_plain, L, scale)
        return lambda: ck.ckks_decode(h_plain, moduli[:L], scale, logn)

    record("ckks_decode", "CKKSEncoder decodes/s, N/2 = 32768 slots each, " + shape, "decodes/s", B, ms,
           slot_bytes, "CKKSEncoder decode of an NTT-form plaintext at level 24 to 32768 complex slots, batch of "
           f"{B}", "shc_ckks_decode_many", make_dec_ckks, "one CKKSEncoder::decode per call")
    del db, back, vals, plain, status
    torch.cuda.empty_cache()

    # relinearization key generation at the key level: K - 1 digits x 2 x K limbs
    sk = dev((K, n))
    ctx.keygen_secret(11, sk.data_ptr(), st)
    rk = dev((K - 1, 2, K, n))
    kw = ctx.keygen_workspace_size()
    kb = wsb(kw)
    ms = timed(lambda: ctx.keygen_relin(sk.data_ptr(), 12, rk.data_ptr(), kb.data_ptr(), kw, st))
    h_sk = sk.cpu().numpy().view(np.uint64).copy()

    def make_keygen(kind):
        s_ = ref_sess(kind, CKKS, n, moduli)
        if s_ is not None:
            return s_.keygen_relin
        r_ = np.random.default_rng(3)
        a = ck.uniform_residues(r_, (K - 1, K, n), moduli)
        e = ck.rns_of_small(r_.integers(-3, 4, size=n), moduli)
        e = np.ascontiguousarray(np.broadcast_to(e, (K - 1, K, n)))
        return lambda: ck.generate_kswitch_key(h_sk, h_sk, a, e, moduli, logn)

    record("ckks_relin_keygen", "relinearization keys/s (KeyGenerator::createRelinKeys), " + shape.split(", batch")[0],
           "keys/s", 1, ms, (2 * K * (K - 1) + K) * limb,
           "one relinearization key at N=2^16, 24+1 primes of 60 bits (24 digits x 2 x 25 limbs, 629 MB)",
           "shc_keygen_relin", make_keygen, "one KeyGenerator + createRelinKeys per call (its secret key included)")
    del rk, kb, sk
    torch.cuda.empty_cache()
    io = facade_io_leg()
    if "error" not in io:
        for name, what, key, nbytes in (("ckks_ciphertext_save", "Ciphertext::save", "ct_save_ms", io["ct_bytes"]),
                                        ("ckks_ciphertext_load", "Ciphertext::load", "ct_load_ms", io["ct_bytes"])):
            t_ms = io[key]
            ops[name] = {"metric": f"{what} of one level-24 ciphertext (N=2^16, 2 x 24 limbs) per second, through a "
                                   "std::stringstream (troyn:: API, device words in and out)",
                         "value": 1e3 / t_ms, "unit": "ciphertexts/s", "ms_per_step": t_ms, "steps": io["reps"],
                         "bytes": nbytes, "host_gb_s": nbytes / (t_ms * 1e-3) / 1e9,
                         "bound": "the PCIe copy of the words between device and host and the host stream copy",
                         "cpu_baseline": None,
                         "cpu_baseline_note": "the reference's CPU Ciphertext has save / load commented out "
                                              "(src/ciphertext.h:453-520); its serialization is in the CUDA classes "
                                              "only, so no CPU path exists to time",
                         "relin_key": {"bytes": io["relin_key_bytes"], "save_ms": io["relin_key_save_ms"],
                                       "load_ms": io["relin_key_load_ms"]},
                         "self_check": "save -> load -> save gives the same bytes (ciphertext and relinearization key)"}
        ops["ckks_relin_keygen"]["facade"] = {"relin_keygen_ms": io["relin_keygen_ms"],
                                              "galois_keygen_ms": io["galois_keygen_ms"],
                                              "path": "troyn::KeyGenerator createRelinKeys / createGaloisKeys({1})"}
    else:
        ops["ckks_ciphertext_save"] = io

    # ---------------------------------------------------------------- BFV, N = 8192
    logn, n, t = 13, 8192, T_BFV
    moduli = [0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001]
    K, L = len(moduli), len(moduli) - 1
    limb = n * 8
    shape = f"N=8192, BFVDefault(8192) (4 data + special), t=65537, batch of {B}"
    ctx = sealhip.Context(sealhip.SCHEME_BFV, n, moduli, t)
    sk, pk = dev((K, n)), dev((2, K, n))
    ctx.keygen_secret(7, sk.data_ptr(), st)
    ctx.keygen_public(sk.data_ptr(), 8, pk.data_ptr(), st)
    values = torch.randint(0, t, (B, n), generator=gen, device="cuda", dtype=torch.int64)
    plain = dev((B, n))

    def encode_each():
        for g in range(B):
            ctx.batch_encode(values[g].data_ptr(), n, plain[g].data_ptr(), st)

    per_call_ms = timed(encode_each)
    ms = timed(lambda: ctx.batch_encode_many(values.data_ptr(), n, n, B, plain.data_ptr(), st))
    h_v = values[0].cpu().numpy().view(np.uint64).copy()

    def make_benc(kind):
        s_ = ref_sess(kind, BFV_SCHEME, n, moduli, t)
        if s_ is not None:
            return lambda: s_.batch_encode(h_v)
        return lambda: ck.batch_encode(h_v, t, logn)

    record("bfv_batch_encode", "BatchEncoder encodes/s, " + shape, "encodes/s", B, ms, B * 2 * n * 8,
           "BatchEncoder encode of 8192 slots mod t, " + shape, "shc_batch_encode_many", make_benc,
           "one BatchEncoder::encode per call")
    ops["bfv_batch_encode"]["one_call_per_plaintext"] = {"value": B / (per_call_ms * 1e-3), "ms": per_call_ms,
                                                         "entry": "shc_batch_encode"}

    cts = dev((B, 2, L, n))
    ew = ctx.encrypt_many_workspace_size(L, B)
    eb = wsb(ew)
    ms = timed(lambda: ctx.encrypt_many(pk.data_ptr(), plain.data_ptr(), B, L, 123, cts.data_ptr(), eb.data_ptr(),
                                        ew, st))
    h_pk = pk.cpu().numpy().view(np.uint64).copy()
    h_plain = plain[0].cpu().numpy().view(np.uint64).copy()

    def make_benc2(kind):
        s_ = ref_sess(kind, BFV_SCHEME, n, moduli, t)
        if s_ is not None:
            return lambda: s_.encrypt(h_pk, h_plain, L)
        u = ck.rns_of_small(np.random.default_rng(1).integers(-1, 2, size=n), moduli)
        e = ck.rns_of_small(np.random.default_rng(2).integers(-3, 4, size=n), moduli)

        def run():
            z = ck.rescale(BFV_SCHEME, ck.encrypt_zero_asymmetric(BFV_SCHEME, h_pk, u, e, e, moduli, logn), moduli,
                           logn)
            ck.bfv_add_plain_scaled(z[0], h_plain, moduli, t, logn)
        return run

    record("bfv_encrypt", "BFV public-key encryptions/s, " + shape, "encryptions/s", B, ms,
           (2 * K + B * 2 * L) * limb + B * n * 8, "BFV encryption of batch-encoded plaintexts, " + shape,
           "shc_encrypt_many", make_benc2, "one Encryptor::encrypt per call")
    del eb
    out = dev((B, n))
    dw = ctx.decrypt_many_workspace_size(L, B)
    db = wsb(dw)
    ms = timed(lambda: ctx.decrypt_many(cts.data_ptr(), 2, B, sk.data_ptr(), L, out.data_ptr(), db.data_ptr(), dw,
                                        st))
    if not torch.equal(out, plain):
        raise SystemExit("bench self-check failed: BFV decrypt(encrypt(m)) != m")
    h_ct = cts[0].cpu().numpy().view(np.uint64).copy()
    h_sk = sk.cpu().numpy().view(np.uint64).copy()

    def make_bdec(kind):
        s_ = ref_sess(kind, BFV_SCHEME, n, moduli, t)
        if s_ is not None:
            return lambda: s_.decrypt(h_ct, h_sk, L)
        return lambda: ck.bfv_decrypt(h_ct, h_sk[:L], moduli, t, logn)

    record("bfv_decrypt", "BFV decryptions/s, " + shape, "decryptions/s", B, ms, (B * 2 * L + L) * limb + B * n * 8,
           "BFV decryption (scale and round) to plaintexts mod t, " + shape, "shc_decrypt_many", make_bdec,
           "one Decryptor::decrypt per call")
    vals_out, scratch = dev((B, n)), dev((B, n))

    def decode_each():
        for g in range(B):
            ctx.batch_decode(out[g].data_ptr(), vals_out[g].data_ptr(), scratch.data_ptr(), st)

    per_call_ms = timed(decode_each)
    ms = timed(lambda: ctx.batch_decode_many(out.data_ptr(), B, vals_out.data_ptr(), scratch.data_ptr(), st))
    if not torch.equal(vals_out, values):
        raise SystemExit("bench self-check failed: BFV encode -> encrypt -> decrypt -> decode lost slots")
    h_p = out[0].cpu().numpy().view(np.uint64).copy()

    def make_bdec2(kind):
        s_ = ref_sess(kind, BFV_SCHEME, n, moduli, t)
        if s_ is not None:
            return lambda: s_.batch_decode(h_p)
        return lambda: ck.batch_decode(h_p, t, logn)

    record("bfv_batch_decode", "BatchEncoder decodes/s, " + shape, "decodes/s", B, ms, B * 2 * n * 8,
           "BatchEncoder decode of 8192 slots mod t, " + shape, "shc_batch_decode_many", make_bdec2,
           "one BatchEncoder::decode per call")
    ops["bfv_batch_decode"]["one_call_per_plaintext"] = {"value": B / (per_call_ms * 1e-3), "ms": per_call_ms,
                                                         "entry": "shc_batch_decode"}
    del cts, out, db, vals_out, plain, values
    torch.cuda.empty_cache()
    # context set-up: moduli, NTT / RNS / BEHZ / encoder tables built and uploaded
    def setup_entry(name, scheme, nn, mods, t, what):
        reps = 3
        ctx_ = sealhip.Context(scheme, nn, mods, t)  # warm (first-use costs)
        del ctx_
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx_ = sealhip.Context(scheme, nn, mods, t)
            del ctx_
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        d = {"metric": f"context set-ups/s ({what})", "value": 1e3 / ms, "unit": "contexts/s", "ms_per_step": ms,
             "steps": reps, "config": {"workload": what, "parallelism": "single GPU"},
             "bound": "host table construction and upload (one-time per parameter set)"}
        if not args.no_cpu and ck.ref() is not None:
            def make(kind):
                return lambda: ck.RefSession(scheme, nn, mods, t)
            d["cpu_baseline"] = cpu_baseline(make, "contexts/s", 1, cpu_s,
                                             "SEALContext(parms, expand_mod_chain = true) + Evaluator per call")
        ops[name] = d
        print(f"[bench] primitives: {name} {d['value']:.1f} contexts/s", file=sys.stderr, flush=True)

    setup_entry("ckks_context_setup", sealhip.SCHEME_CKKS, 1 << 16, sealhip.coeff_modulus_create(1 << 16, [60] * 25),
                0, "CKKS N=2^16, CoeffModulus::Create(65536, {60 x 25}): NTT tables of 25 primes, rescale / key-switch "
                   "constants and CKKS encoder tables of every level")
    setup_entry("bfv_context_setup", sealhip.SCHEME_BFV, n, moduli, t,
                "BFV N=8192, BFVDefault(8192), t=65537: NTT tables, BEHZ tables of every level, batching tables")
    first = ops["ckks_encrypt"]
    return {"metric": "per-operation throughput of the path's other entry points (rotation, rescale, encryption, "
                      "decryption, encoders)", "value": first["value"], "unit": first["unit"],
            "ms_per_step": first["ms_per_step"], "steps": steps, "n_gpus": 1, "warmup": warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": "synthetic (seeded uniform residues and slots; device-generated keys for the round trips)",
            "config": {"workload": "per-operation rates: CKKS N=2^16 L=24 and BFV N=8192, batches of 128",
                       "parallelism": "single GPU"},
            "ops": ops}


SUB_KEYS = ("metric", "value", "unit", "ms_per_step", "steps", "config", "roofline", "step_roofline", "cpu_baseline")


def bench_default(args):
    """The headline line (CKKS mul+relin at N=2^16, L=24, batch 128) with every other
    workload as a sub-object under "configs": the CPU-runnable BFV path (N=8192), the
    batched NTT, the CKKS chain in a HIP graph, the hoisted rotation sums, the served
    matvec mix and one object per other entry point (rotation, rescale, encryption,
    decryption, encoders) -- each with its value, step time, roofline and CPU
    baseline.  A sub-workload whose self-check fails stops the run; one that fails
    otherwise is reported with its error."""
    import copy
    import torch
    line = bench_mulrelin(args)
    if args.headline_only:
        return line
    sub = copy.copy(args)
    sub.steps, sub.warmup = args.sub_steps, min(args.warmup, 2)
    configs = {}
    for name, fn in (("bfv_n8192", bench_bfv8192), ("ntt", bench_ntt), ("chain", bench_chain),
                     ("rotsum", bench_rotsum), ("serve", bench_serve), ("primitives", bench_primitives)):
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        try:
            r = fn(sub)
            if "ops" in r:  # one sub-object per operation
                for k, v in r["ops"].items():
                    configs[k] = v
                print(f"[bench] {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
                continue
            d = {k: r[k] for k in SUB_KEYS if k in r}
            for k in ("batched", "plain", "capacity_per_s", "admission", "hoisted_over_separate", "graph_over_eager",
                      "eager_latency_ms",
                      "latency_ms", "self_check", "per_level_graph_ms", "grouped_graph_ms", "path"):
                if k in r:
                    d[k] = r[k]
        except (RuntimeError, MemoryError, OSError) as e:
            d = {"error": f"{type(e).__name__}: {e}"}
        d["wall_s"] = time.perf_counter() - t0
        configs[name] = d
        print(f"[bench] {name} done in {d['wall_s']:.1f} s", file=sys.stderr, flush=True)
    line["configs"] = configs
    return line


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
    out = {"mulrelin": bench_default, "ntt": bench_ntt, "copy": bench_copy, "chain": bench_chain,
           "rotsum": bench_rotsum, "matvec": bench_matvec, "serve": bench_serve,
           "bfv8192": bench_bfv8192, "bfv": bench_bfv, "primitives": bench_primitives}[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
