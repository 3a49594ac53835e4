"""Benchmark of the MI355X Collatz range verifier.

Workload (one step): verify every n in [2^40, 2^40 + 2^36) with the k=32 residue sieve
(25,005,944 survivors, resident in HBM), the mod 9 sieve, two u64 limbs (three on demand)
and a record per 2^24-wide chunk: one call of collatz_verify_range.  Inputs (the sieve) are
built once and stay resident before the timed region.

Prints one JSON line (see DESIGN.md for the fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("verified n/sec on 1 GPU (batched/served vs plain, bit-exact); p50/p99 latency")
LO = 1 << 40
WIDTH = 1 << 36
K = 32
CHUNK = 24
# Algorithmic integer operations (32-bit lane ops) per check and per candidate for L
# 32-bit words: OPS_DR = 13 L + 7, OPS_SETUP = 4 L + 8 (L = 4 for two u64 limbs).
L_WORDS = 4
OPS_DR = 13 * L_WORDS + 7
OPS_SETUP = 4 * L_WORDS + 8
# gfx950 int32 VALU issue peak: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz.
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def cpu_baseline(threads: int, sample_chunks: int):
    """The oracle (CPU restatement) on a bounded sample of the same workload."""
    from oracle import bind as B
    t0 = time.time()
    sv = B.sieve(K)
    build_s = time.time() - t0
    lo = LO + (WIDTH // 2)
    hi = lo + sample_chunks * (1 << CHUNK)
    t0 = time.time()
    _, tot = B.verify_range(sv, lo, hi, CHUNK, threads=threads)
    dt = time.time() - t0
    return {"value": (hi - lo) / dt, "unit": "n/s", "cores": threads, "kind": "port",
            "sample": f"{sample_chunks} chunks of 2^{CHUNK} from the middle of the range "
                      f"({int(tot['count'])} candidates, {dt:.1f} s on {threads} threads; "
                      f"sieve build {build_s:.1f} s excluded as on the GPU)"}, tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", choices=["plain", "persistent", "graph"], default="plain")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-chunks", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU sharding is described in DESIGN.md)"}))
        sys.exit(2)
    import collatz_amd as cz
    form = cz.FORM_PLAIN if args.form == "plain" else cz.FORM_PERSISTENT
    sv = cz.Sieve(K)
    info = sv.info()

    def step():
        if args.form == "graph":
            return sv.verify_batch([(LO, LO + WIDTH)], chunk_log2=CHUNK, use_graph=True)[0]
        return sv.verify_range(LO, LO + WIDTH, CHUNK, form=form)

    for _ in range(args.warmup):
        step()
    dev_ms = []
    t0 = time.perf_counter()
    out = None
    for _ in range(args.steps):
        out = step()
        dev_ms.append(out.device_ms)
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    tot = out.total
    n_per_s = WIDTH / (dt / args.steps)
    kern_ms = sum(dev_ms) / len(dev_ms)
    ops = out.checks * OPS_DR + int(tot["count"]) * OPS_SETUP
    achieved = ops / (kern_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": n_per_s,
        "unit": "n/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "verify [2^40, 2^40+2^36) with a k=32 sieve, mod 9 on, "
                               "2x u64 limbs, 2^24-wide chunk records",
                   "form": args.form, "k": K, "chunk_log2": CHUNK, "width": WIDTH,
                   "survivors": int(info.survivors)},
        "survivors_per_s": int(tot["count"]) / (dt / args.steps),
        "candidates": int(tot["count"]),
        "checks": int(out.checks),
        "kernel_ms": kern_ms,
        "total_hash": "%08X" % int(tot["hash"]),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS,
                     "unit": "TOP/s", "frac": achieved / VALU_PEAK_TOPS, "traffic": None,
                     "kernel": "k_verify_plain" if form == cz.FORM_PLAIN else "k_verify_persistent"},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"], _ = cpu_baseline(threads, args.cpu_chunks)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
