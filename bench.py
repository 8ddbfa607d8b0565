#!/usr/bin/env python3
"""Benchmark of the MI355X RNS polynomial core.

Default workload (`--workload ntt`): batched negacyclic NTT, N = 2^16, 24 primes of
60 bits (= 1 mod 2^17), 128 ciphertexts x 2 polys x 24 limbs, forward then inverse.
One step = one forward + one inverse transform of the whole batch, inputs resident
in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "seal-cuda_amd"))
sys.path.insert(0, os.path.join(HERE, "tests"))

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="ntt", choices=["ntt"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    return p.parse_args()


def cpu_baseline_ntt(moduli, logn, budget_s):
    """Reference CPU path (oracle/_ref) when present, else the oracle restatement,
    timed on this host: forward+inverse of single limbs until the budget is spent."""
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
    import checkers as ck  # noqa: F401  (cpu baseline only)
    sealhip.init(0)
    logn, n, L, polys = 16, 1 << 16, 24, 256
    moduli = sealhip.coeff_modulus_create(n, [60] * L)
    ctx = sealhip.Context(sealhip.SCHEME_CKKS, n, moduli)
    gen = torch.Generator(device="cuda").manual_seed(1)
    d = torch.empty((polys, L, n), dtype=torch.int64, device="cuda")
    for i, q in enumerate(moduli):
        d[:, i, :] = torch.randint(0, q, (polys, n), generator=gen, device="cuda", dtype=torch.int64)
    st = torch.cuda.current_stream().cuda_stream

    def step():
        ctx.ntt_forward(d.data_ptr(), polys, L, st)
        ctx.ntt_inverse(d.data_ptr(), polys, L, st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # forward-only timing of the whole forward transform (both passes) for the roofline
    fw0, fw1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    fw0.record()
    ctx.ntt_forward(d.data_ptr(), polys, L, st)
    fw1.record()
    torch.cuda.synchronize()
    fwd_ms = fw0.elapsed_time(fw1)
    ctx.ntt_inverse(d.data_ptr(), polys, L, st)
    torch.cuda.synchronize()

    limbs = polys * L
    limb_bytes = n * 8
    # compulsory traffic of one forward transform: read + write every limb once,
    # plus one read of each prime's forward twiddle table (N x 16 B)
    fwd_bytes = 2 * limbs * limb_bytes + L * n * 16
    step_bytes = 2 * fwd_bytes
    achieved = fwd_bytes / (fwd_ms * 1e-3) / 1e9
    out = {
        "metric": "batched NTT limbs/s (forward+inverse), N=2^16, 24x60-bit primes",
        "value": limbs / (ms * 1e-3),
        "unit": "limb fwd+inv/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (uniform residues, seeded)",
        "config": {"workload": "batched NTT: N=2^16, 24x60-bit primes, 128 ct x 2 polys x 24 limbs, fwd+inv",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "forward NTT (columns + rows pass)", "algorithmic_bytes": fwd_bytes,
                     "avg_ms": fwd_ms},
        "wall_s": wall,
        "step_algorithmic_bytes": step_bytes,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ntt(moduli, logn, args.cpu_seconds)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus N > 1 is not supported "
                                   "(independent ciphertexts shard across GPUs with no exchange)"}))
        sys.exit(2)
    out = bench_ntt(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
