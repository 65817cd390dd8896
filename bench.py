"""Throughput of the exact 868-bit negacyclic product on one MI355X.

Workload (one step): one persistent, graph-captured launch over a batch of B
independent products c = a*b mod (X^N+1, q), N = 32768, q a seeded random
odd 868-bit modulus, coefficients uniform in [0,q) as 28 u32 limbs, inputs
resident in HBM.  `value` is batched products/s over K timed steps; the plain
single-product path (three launches per product, inputs in HBM) is timed
beside it, with p50/p99 latencies, and the oracle CPU port on a bounded
sample of the same workload (which also checks the GPU results bit for bit).

    python bench.py [--steps K] [--warmup W] [--batch B] [--n N] [--cpu-seconds S]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "fherma-polymul-baseline_amd")
METRIC = "N=2^15 868-bit negacyclic mults/s, batched vs plain (bit-exact); p99 ms"
HBM_PEAK_GBS = 8000.0


def package():
    spec = importlib.util.spec_from_file_location("polymul_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["polymul_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_bytes(n, k):
    """Compulsory HBM bytes of one product: conversion reads 2 coefficient
    arrays (112 B each) and writes 2k residue arrays; the fused NTT tile reads
    2k and writes k residue arrays; the CRT reads k residue arrays and writes
    the 112 B coefficients."""
    conv = 2 * n * 112 + 2 * k * n * 4
    ntt = 3 * k * n * 4
    crt = k * n * 4 + n * 112
    return conv + ntt + crt


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--plain", type=int, default=200, help="plain single-product launches to time")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", default="phased", choices=["phased", "persistent"],
                    help="batched launch: one launch per phase, or one persistent kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "products are independent, so more GPUs would each run their own batch", file=sys.stderr)
        return 2

    import torch
    pm = package()
    wl = pm.workload
    n, B = args.n, args.batch
    q = wl.modulus(args.seed)
    ctx = pm.Context(q, n)
    k = ctx.info()["k"]
    # a dedicated stream (default priority): handle 0 would mean "context stream"
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream

    # inputs resident in HBM: product i uses seed (seed, i)
    batch = pm.Batch(ctx, [n] * B, mode=args.mode)
    inputs = []
    for i in range(B):
        s = wl.mix64(np.uint64(args.seed * 1000003 + i)).item()
        a, b = wl.poly(s, 0, n, q), wl.poly(s, 1, n, q)
        batch.upload(i, a, b)
        if i < 64:
            inputs.append((a, b))
    print(f"[bench] inputs ready: {B} products, N={n}, k={k}", file=sys.stderr, flush=True)

    # ---- batched path -------------------------------------------------------
    for _ in range(args.warmup):
        batch.run(sptr)
    batch.wait(sptr)
    step_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for e0, e1 in ev:
        e0.record(stream)
        batch.run(sptr)
        e1.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    batch.wait(sptr)
    step_ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    batched_rate = B * args.steps / elapsed
    print(f"[bench] batched: {batched_rate:.0f} products/s, step p50 {percentile(step_ms, 50):.3f} ms",
          file=sys.stderr, flush=True)

    # ---- plain path (one product per launch sequence, inputs in HBM) --------
    da = torch.from_numpy(inputs[0][0].view(np.int32)).cuda()
    db = torch.from_numpy(inputs[0][1].view(np.int32)).cuda()
    dc = torch.empty_like(da)
    for _ in range(5):
        ctx.multiply_device(n, da, db, dc, sptr)
    torch.cuda.synchronize()
    plain_ms = []
    for _ in range(args.plain):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.multiply_device(n, da, db, dc, sptr)
        e1.record(stream)
        e1.synchronize()
        plain_ms.append(e0.elapsed_time(e1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain):
        ctx.multiply_device(n, da, db, dc, sptr)
    torch.cuda.synchronize()
    plain_rate = args.plain / (time.perf_counter() - t0)
    plain_c = dc.cpu().numpy().view(np.uint32).reshape(n, 28)
    # drop-in path with host buffers (PCIe-inclusive), as fherma_run times it
    host_ms = []
    for _ in range(20):
        t1 = time.perf_counter()
        host_c = ctx.multiply(inputs[0][0], inputs[0][1])
        host_ms.append((time.perf_counter() - t1) * 1e3)
    print(f"[bench] plain: {plain_rate:.0f} products/s, p50 {percentile(plain_ms, 50):.3f} ms, "
          f"host drop-in p50 {percentile(host_ms, 50):.3f} ms", file=sys.stderr, flush=True)

    batch_c = [batch.download(i) for i in range(min(len(inputs), B))]
    gpu_consistent = bool(np.array_equal(batch_c[0], plain_c) and np.array_equal(host_c, plain_c))

    # ---- roofline of the batched launch -------------------------------------
    per_product = algorithmic_bytes(n, k)
    launch_ms = percentile(step_ms, 50)
    achieved = per_product * B / (launch_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_batched_n{n}_b{B}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")

    # ---- CPU baseline: the oracle port on a bounded sample -------------------
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # the checker; timed here as the CPU port of the reference path
    orc = oracle_lib.Oracle()
    primes = orc.reference_primes_for(n, q)
    done, exact = 0, True
    t0 = time.perf_counter()
    while done < len(batch_c) and (time.perf_counter() - t0) < args.cpu_seconds:
        a, b = inputs[done]
        c = orc.rns_multiply(q, a, b, primes)
        exact = exact and bool(np.array_equal(c, batch_c[done]))
        done += 1
    cpu_elapsed = time.perf_counter() - t0
    cpu_rate = done / cpu_elapsed

    line = {
        "metric": METRIC,
        "value": round(batched_rate, 1),
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"batched N={n} 868-bit negacyclic products, {B} per graph launch ({args.mode}), inputs in HBM",
            "n": n, "batch": B, "k_primes": k, "modulus": f"seeded random odd 868-bit (seed {args.seed})",
            "step_p50_ms": round(percentile(step_ms, 50), 4), "step_p99_ms": round(percentile(step_ms, 99), 4),
            "plain_products_per_s": round(plain_rate, 1),
            "plain_p50_ms": round(percentile(plain_ms, 50), 4), "plain_p99_ms": round(percentile(plain_ms, 99), 4),
            "batched_over_plain": round(batched_rate / plain_rate, 3),
            "dropin_host_buffers_p50_ms": round(percentile(host_ms, 50), 4),
            "dropin_host_buffers_p99_ms": round(percentile(host_ms, 99), 4),
            "gpu_paths_agree": gpu_consistent,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"batched graph ({args.mode}), event-timed on its stream",
            "algorithmic_bytes_per_product": per_product,
        },
        "cpu_baseline": {
            "value": round(cpu_rate, 4), "unit": "products/s", "cores": 1, "kind": "port",
            "sample": f"{done} products of the same workload (N={n}), oracle RNS port, single thread",
        },
        "bit_exact": bool(exact and gpu_consistent and done > 0),
    }
    print(json.dumps(line), flush=True)
    batch.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
