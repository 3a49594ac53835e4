"""Benchmark of the MI355X Collatz range verifier: batched and served forms against the
plain kernel over the same ranges, bit-exact.

Headline workload (one step): the 4096 intervals of the batch workload (SplitMix64 seed 0;
starts uniform in [2^48, 2^60), widths log-uniform in [2^20, 2^34] rounded up to a multiple
of 2^32; 2.17e13 integers in all), verified with the k=32 residue sieve (25,005,944
survivors, resident in HBM), the mod 9 sieve, two u64 limbs (three on demand) and one
record per 2^32-wide chunk, in ONE collatz_verify_batch call (persistent launch replayed
from a hipGraph, device work stealing, autotuned launch shape).  `value` is that form's
integers per second.

Beside it, on the same GPU and sieve:
  * plain: the same 4096 intervals, one collatz_verify_range call (plain kernel) each; the
    per-interval records of both forms are compared and must be identical;
  * range: [2^40, 2^40 + 2^36) with 2^24-wide chunk records, plain and persistent;
  * served: the request stream (SplitMix64 seed 1; widths log-uniform in [2^16, 2^36];
    80% of starts in [2^40, 2^62), 20% in [2^62, 2^64)) replayed at 50..500 requests/s
    through collatz_server_*, with p50/p99 latency, and once all at once (saturated);
  * large sieve: k = 37 on [2^60, 2^60 + 2^44) with 2^30-wide chunks;
  * reference program: the reference's own tree-mode run for all n < 2^50 (reference mode),
    beside the paper's published RTX 3060 time for the same run;
  * cpu_baseline: the CPU restatement (oracle) on a bounded sample of the same workload.

Every call returns after its stream has synchronised; the timed region is also bracketed
by hipDeviceSynchronize.  Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "verified n/sec on 1 GPU (batched/served vs plain, bit-exact); p50/p99 latency"
K = 32
BATCH_CHUNK = 32
LOOKAHEAD_BITS = 16
RANGE_LO, RANGE_WIDTH, RANGE_CHUNK = 1 << 40, 1 << 36, 24
# Algorithmic integer work, in 32-bit lane operations, for L 32-bit words per value: per
# check (two odd and two even runs, hash and compare) OPS_DR = 13 L + 7, and per candidate
# (mod 9, start value, base) OPS_SETUP = 4 L + 8.  Two-limb work whose starts stay below
# 2^68 runs on L = 3 words (96 bits; a candidate whose check ends at or above 2^86 is
# rerun on three limbs); otherwise L = 4.
def ops_dr(L):
    return 13 * L + 7


def ops_setup(L):
    return 4 * L + 8


def words_for(max_hi):
    return 3 if max_hi <= (1 << 68) else 4


# 32-bit VALU peak in lane-ops/s: MI355X's published FP32 vector rate, 157.3 TFLOP/s, is
# 78.6 T fused multiply-adds (two flops each) per second, which is 256 CUs x 2.4 GHz x two
# wave64 instructions per CU per cycle.  The probe collatz_amd/valu_peak measures what the
# integer instructions reach (profiles/valu_peak_gfx950.md): VOP2 add, xor, or and shifts
# 1.6-1.7 wave instructions per CU per cycle (55-62 T), VOP3 ones (multiplies, funnel
# shifts, bit-field extract, three-input adds) 0.95 (37-38 T), and the schedule's own check
# loop on 96-bit values in registers 40.5 T algorithmic lane-ops/s.  bench.py runs the
# probe beside the workload and reports the achieved rate against both.
VALU_PEAK_TOPS = 157.3 / 2
VALU_PROBE = os.path.join(ROOT, "collatz_amd", "valu_peak")
RATES = (50, 100, 200, 500)
FULL_RATES = (200, 500)
# HBM traffic of the batched launch, measured with rocprofv3 on this workload
# (profiles/k_verify_persistent_batch4096_k32.md): FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0,
# the factors a 1 GiB copy kernel gives.  Reads are the 16-byte survivor records (13.19 B per
# candidate measured: the records of the rows that share a survivor window, taken together
# by the tile order, come partly from L2); writes are the per-tile records.
READ_BYTES_PER_CANDIDATE = 13.19
WRITE_BYTES_PER_LAUNCH = 0.078e9


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_hip = None


def device_sync():
    """hipDeviceSynchronize from the HIP runtime the library is linked against (the timed
    region is bracketed by it; the library's own calls also return after their streams have
    synchronised)."""
    global _hip
    if _hip is None:
        import ctypes
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
    if _hip is not None and _hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def records_equal(a, b) -> bool:
    return a.shape == b.shape and a.tobytes() == b.tobytes()


CPU_CHUNK = 26


def cpu_baseline(intervals, threads: int, n_intervals: int, width_log2: int):
    """The oracle (CPU restatement) on the first `n_intervals` intervals, each cut to its
    first 2^width_log2 integers.  The oracle runs chunk records in parallel, so it is given
    2^26-wide records (the work is the same; only the record granularity differs)."""
    from oracle import bind as B
    t0 = time.time()
    sv = B.sieve(K)
    build_s = time.time() - t0
    n = 0
    t0 = time.time()
    for lo, hi in intervals[:n_intervals]:
        hi = min(hi, lo + (1 << width_log2))
        B.verify_range(sv, lo, hi, CPU_CHUNK, threads=threads)
        n += hi - lo
    dt = time.time() - t0
    # The per-core figure: the first four of the same intervals cut to 2^30, one thread.
    n1 = 0
    t1 = time.time()
    for lo, hi in intervals[:4]:
        hi = min(hi, lo + (1 << 30))
        B.verify_range(sv, lo, hi, CPU_CHUNK, threads=1)
        n1 += hi - lo
    dt1 = time.time() - t1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n / dt, "unit": "n/s", "cores": threads, "kind": "port",
            "sample": f"first {n_intervals} intervals of the batch workload, each cut to "
                      f"2^{width_log2} integers, 2^{CPU_CHUNK}-wide records ({n:.3e} n in "
                      f"{dt:.1f} s on {threads} threads; oracle sieve build {build_s:.1f} s "
                      f"excluded, as on the GPU)",
            "single_thread": {"value": n1 / dt1, "unit": "n/s",
                              "sample": f"first 4 intervals cut to 2^30 ({n1:.3e} n in "
                                        f"{dt1:.1f} s)"},
            "cpu_model": model, "host_cpus": os.cpu_count()}


C4_LO, C4_WIDTH_LOG2, C4_CHUNK, C4_K = 1 << 60, 44, 30, 37


def total_fields_equal(cpu_total, gpu_chunk) -> bool:
    """A chunk record's results (the CPU total of its 2^26-wide records against the GPU's
    record of the whole chunk): everything but the index."""
    import numpy as np
    keys = ("n_checked", "count", "stop_sum", "stop_xor", "max_exc", "max_start",
            "overflow_count", "hash", "flags")
    return all(np.array_equal(cpu_total[k], gpu_chunk[k]) for k in keys)


def cpu_configs(threads, gpu):
    """The oracle on the host cores for the other configurations, each beside the GPU figure
    of the same workload, with the records compared:
      c0  [1, 2^32), k = 20, 2^24-wide records: in full;
      c1  [2^40, 2^40 + 2^36), k = 32, 2^24-wide records: in full;
      c3  the first requests of the served stream (seed 1) up to 2^36 integers, k = 32: a
          prefix, so the rate is an extrapolation to the stream;
      c4  32 of the 16384 2^30-wide chunks of [2^60, 2^60 + 2^44), k = 37 (8 chunks in each
          of 4 seeded residue windows, from the oracle's windowed sieve): extrapolated.
    `gpu` holds the GPU side: c0 and c1 chunk records, c4 chunk records, the sieve handle
    for c3 and the GPU rates."""
    import random

    from oracle import bind as B
    from collatz_amd import workloads as W
    out = {}

    def entry(n, dt, sample, total_n, same, gpu_rate):
        e = {"value": n / dt, "unit": "n/s", "cores": threads, "kind": "port",
             "sample": sample, "integers": n, "seconds": dt, "bit_exact_vs_gpu": same,
             "gpu_n_per_s": gpu_rate}
        if total_n != n:
            e["extrapolated_full_run_s"] = total_n / (n / dt)
        return e

    s20 = B.sieve(20)
    t0 = time.time()
    ch, _ = B.verify_range(s20, 1, 1 << 32, 24, threads=threads)
    dt = time.time() - t0
    out["c0"] = entry((1 << 32) - 1, dt, "in full: [1, 2^32), k = 20, 256 records of 2^24",
                      (1 << 32) - 1, records_equal(ch, gpu["c0_chunks"]), gpu["c0_rate"])
    del s20
    s32 = B.sieve(K)
    t0 = time.time()
    ch, _ = B.verify_range(s32, RANGE_LO, RANGE_LO + RANGE_WIDTH, RANGE_CHUNK, threads=threads)
    dt = time.time() - t0
    out["c1"] = entry(RANGE_WIDTH, dt, "in full: [2^40, 2^40+2^36), k = 32, 4096 records of "
                      "2^24", RANGE_WIDTH, records_equal(ch, gpu["c1_chunks"]), gpu["c1_rate"])
    reqs = W.request_stream(1, 20000)
    sample, n = [], 0
    for r in reqs:
        if n >= 1 << 36:
            break
        sample.append(r)
        n += r[2] - r[1]
    t0 = time.time()
    cpu = [B.verify_range(s32, lo, hi, CPU_CHUNK, threads=threads)[1] for _, lo, hi in sample]
    dt = time.time() - t0
    same = True
    for (_, lo, hi), tot in zip(sample, cpu):
        g = gpu["sieve32"].verify_range(lo, hi, 40).total
        same = same and total_fields_equal(tot, g)
    out["c3"] = entry(n, dt, f"the first {len(sample)} requests of the served stream (seed 1), "
                      "k = 32, 2^26-wide records", sum(hi - lo for _, lo, hi in reqs), same,
                      gpu["c3_rate"])
    del s32
    rng = random.Random(4)
    windows = rng.sample(range(1 << (C4_K - C4_CHUNK)), 4)
    per_q = 1 << (C4_WIDTH_LOG2 - C4_K)
    chunks = [(q << (C4_K - C4_CHUNK)) + w for w in windows for q in rng.sample(range(per_q), 8)]
    t0 = time.time()
    sw = B.sieve_windows(C4_K, C4_CHUNK, windows, threads=threads)
    build_s = time.time() - t0
    t0 = time.time()
    cpu = [B.verify_range(sw, C4_LO + (j << C4_CHUNK), C4_LO + ((j + 1) << C4_CHUNK), CPU_CHUNK,
                          threads=threads)[1] for j in chunks]
    dt = time.time() - t0
    same = all(total_fields_equal(t, gpu["c4_chunks"][j]) for t, j in zip(cpu, chunks))
    out["c4"] = entry(len(chunks) << C4_CHUNK, dt,
                      f"32 of the 16384 2^30-wide chunks of [2^60, 2^60+2^44), k = 37: 8 in "
                      f"each of the residue windows {sorted(windows)} (sieve of those windows "
                      f"{build_s:.1f} s, excluded)", 1 << C4_WIDTH_LOG2, same, gpu["c4_rate"])
    return out


def served_leg(sv, reqs, rate, chunk_log2, check_n=0):
    """Replays `reqs` (gap, lo, hi) at `rate` requests/s (rate None: all at once).  The
    results of the first `check_n` requests are compared with the plain kernel's records of
    the same ranges afterwards (outside the timed replay)."""
    import collatz_amd as cz
    srv = cz.Server(sv, chunk_log2=chunk_log2)
    tickets = []
    t0 = time.perf_counter()

    def producer():
        t = 0.0
        for gap, lo, hi in reqs:
            if rate:
                t += gap / rate
                d = t0 + t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            tickets.append((srv.submit(lo, hi), hi - lo))

    th = threading.Thread(target=producer)
    th.start()
    lat, n, first_submit, last_done = [], 0, None, 0.0
    kept = []
    i = 0
    t_log = time.perf_counter()
    while True:
        if i < len(tickets):
            t, w = tickets[i]
            total, chunks, (ts, ta, td) = srv.wait(t)
            if i < check_n:
                kept.append((total, chunks))
            lat.append(td - ts)
            n += w
            first_submit = ts if first_submit is None else min(first_submit, ts)
            last_done = max(last_done, td)
            i += 1
            if time.perf_counter() - t_log > 10:
                t_log = time.perf_counter()
                log(f"served at {rate or 'saturation'}: {i}/{len(reqs)} requests done")
        elif not th.is_alive() and i >= len(tickets):
            break
        else:
            time.sleep(0.0005)
    th.join()
    srv.close()
    same = True
    for (total, chunks), (_, lo, hi) in zip(kept, reqs):
        p = sv.verify_range(lo, hi, chunk_log2, form=cz.FORM_PLAIN)
        if not (records_equal(chunks, p.chunks) and total.tobytes() == p.total.tobytes()):
            same = False
    lat.sort()
    pct = lambda p: lat[min(len(lat) - 1, int(p * len(lat)))]
    span_s = (last_done - first_submit) / 1e3
    return {"requests": len(lat), "p50_ms": pct(0.50), "p99_ms": pct(0.99),
            "max_ms": lat[-1], "n_per_s": n / span_s, "integers": n,
            "offered_req_per_s": rate, "checked_vs_plain": len(kept),
            "bit_exact_vs_plain": same}


def lookahead_leg(cz, ivs, total_n, steps):
    """The batch workload with the Step 2 look-ahead of B = LOOKAHEAD_BITS steps at depth 32:
    fewer candidates iterated (different counts and checksums by definition), the same
    integers verified.  n/s by device time, as the headline; the per-call rate beside it."""
    sla = cz.Sieve(K, lookahead_bits=LOOKAHEAD_BITS)
    sla.verify_batch(ivs, chunk_log2=BATCH_CHUNK, use_graph=True)
    ms = []
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        o = sla.verify_batch(ivs, chunk_log2=BATCH_CHUNK, use_graph=True)
        ms.append(o[0].device_ms)
    device_sync()
    d = (time.perf_counter() - t0) / steps
    lah = 0
    for x in o:
        lah ^= int(x.total["hash"])
    lms = sum(ms) / len(ms)
    sla.close()
    return {"bits": LOOKAHEAD_BITS, "semantics": "cuda", "n_per_s": total_n / (lms * 1e-3),
            "host_inclusive_n_per_s": total_n / d, "ms_per_step": d * 1e3, "kernel_ms": lms,
            "candidates": sum(int(x.total["count"]) for x in o),
            "checks": sum(int(x.checks) for x in o), "total_hash": "%08X" % lah}


# main.rs's configuration (LASTBITS 6, eight bit vectors, its own search3 in 16-step blocks):
# the paper's one-core Rust timings and counts of integers iterated (p.3 section 1.1, p.10
# Fig 5.1).
RUST_PUBLISHED = {40: {"seconds": 21.0, "candidates": 756583624},
                  50: {"seconds": 176 * 60.0, "candidates": 434413363553}}


def rust_leg(cz, Ns):
    """The whole tree below 2^N under main.rs semantics from the root "3 2 2 9 1 0": Step 1
    with main.rs's prunes, its Step 2 filters and its search3 blocks; numcases is what
    main.rs prints."""
    out = {}
    for N in Ns:
        t = time.perf_counter()
        r = cz.reference_run((3, 2, 2, 9, 1, 0), N, 6, num_bv=8, semantics=cz.SEM_RUST,
                             cutoff_log2=120)
        d = time.perf_counter() - t
        rec = {"seconds": d, "n_per_s": (1 << N) / d, "candidates": int(r.candidates),
               "numcases": int(r.numcases), "leaves": int(r.leaves),
               "overflow_count": int(r.overflow_count)}
        pub = RUST_PUBLISHED.get(N)
        if pub:
            rec["published"] = dict(pub, hardware="1 AMD Zen 3 core at 4.6 GHz (main.rs)",
                                    source="paper p.3 section 1.1, p.10 Fig 5.1")
            rec["candidates_match_published"] = rec["candidates"] == pub["candidates"]
            rec["speedup_vs_published"] = pub["seconds"] / d
        out[str(N)] = rec
        log(f"main.rs configuration, all n < 2^{N}: {rec}")
    return out


# The reference's n < 2^60 run: 24.4 h on an RTX 3060 (paper p.3-4 section 1.1), done as
# the 278,699 lines of cases_big (paper p.4).
CASES_BIG_LINES = 278699
PAPER_N60_SECONDS = 24.4 * 3600


def deep_lines_leg(cz, N=60):
    """The reference program (CUDA semantics, A = 10, 16 bit vectors, CUTOFFPOWER 100) for
    the 240 sampled lines of cases_big at MAXBITS = N (leaves at depth N - 10, 128 bits for
    N > 50), one collatz_reference_run per line; the mean time per line times the 278,699
    lines projects the whole n < 2^60 run."""
    path = os.path.join(ROOT, "tests", "golden", "cases_big_sample.txt")
    with open(path) as f:
        lines = [tuple(int(x) for x in ln.split()) for ln in f if ln.strip()]
    cz.reference_run(lines[0], N, 10, semantics=cz.SEM_CUDA)  # tables, untimed
    t = time.perf_counter()
    cand = leaves = 0
    h = 0
    for ln in lines:
        r = cz.reference_run(ln, N, 10, semantics=cz.SEM_CUDA)
        cand += int(r.candidates)
        leaves += int(r.leaves)
        h ^= r.complete_hash
    d = time.perf_counter() - t
    per_line = d / len(lines)
    return {"N": N, "lines": len(lines), "seconds": d, "seconds_per_line": per_line,
            "candidates": cand, "leaves": leaves, "xor_of_hashes": "%08X" % h,
            "projected_full_run_hours": per_line * CASES_BIG_LINES / 3600,
            "published_full_run_hours": PAPER_N60_SECONDS / 3600 if N == 60 else None,
            "published_hardware": "RTX 3060 (paper p.3-4)" if N == 60 else None,
            "projected_speedup_vs_published":
                PAPER_N60_SECONDS / (per_line * CASES_BIG_LINES) if N == 60 else None,
            "note": "a projection from a 240-line sample of cases_big (the sample is the "
                    "fixture tests/golden/cases_big_sample.txt), not a full run"}


def valu_probe():
    """The int32 VALU ceiling measured on this GPU by collatz_amd/valu_peak (a child process,
    about a second): per-instruction lane-op rates at full occupancy and the rate of the
    schedule's check loop in registers.  None if the probe is not built."""
    import subprocess
    if not os.path.exists(VALU_PROBE):
        return None
    out = subprocess.run([VALU_PROBE], capture_output=True, text=True, timeout=120, check=True)
    d = json.loads(out.stdout)
    ops = d["ops"]
    vop2 = ("v_add_u32", "v_xor_b32", "v_or_b32", "v_lshrrev_b32")
    vop3 = ("v_mul_lo_u32", "v_alignbit_b32", "v_mad_u32_u24", "v_bfe_u32", "v_add3_u32")
    return {"vop2_simple_tops": max(ops[k]["tops"] for k in vop2 if k in ops),
            "vop3_tops": max(ops[k]["tops"] for k in vop3 if k in ops),
            "checks_loop_tops": d["checks_loop"]["algorithmic_tops"],
            "checks_loop_checks_per_s": d["checks_loop"]["checks_per_s"],
            "per_instruction_tops": {k: round(v["tops"], 2) for k, v in ops.items()},
            "source": "collatz_amd/valu_peak (profiles/valu_peak_gfx950.md)"}


def tree_leg(cz, N):
    """The reference's own program (tree mode) for all n < 2^N from the root case line
    "3 2 2 9 1 0", its default knobs (LASTBITS 10, STEPSIZEINITIAL 24, 16 bit vectors,
    CUTOFFPOWER 100, collatz_cudav7.cu semantics).  The paper times this run at N = 50 on an
    RTX 3060: 166 s (p.3, section 1.1)."""
    t = time.perf_counter()
    r = cz.reference_run((3, 2, 2, 9, 1, 0), N, 10, semantics=cz.SEM_CUDA)
    d = time.perf_counter() - t
    # Algorithmic integer work at L = 3 (the reference's candidates start below 2^96 and run
    # on 96 bits) over the whole call's wall time: sieve, tables and both iteration passes.
    ach = (int(r.checks) * ops_dr(3) + int(r.candidates) * ops_setup(3)) / d / 1e12
    tree = {"N": N, "A": 10, "B": 24, "semantics": "cuda", "seconds": d,
            "n_per_s": (1 << N) / d, "complete_hash": "%X" % r.complete_hash,
            "launches": int(r.launches), "leaves": int(r.leaves),
            "candidates": int(r.candidates), "checks": int(r.checks),
            "roofline_frac": ach / VALU_PEAK_TOPS}
    if N == 50:
        tree["published"] = {"hardware": "RTX 3060", "seconds": 166.0, "n_per_s": 6.8e12,
                             "source": "paper p.3 section 1.1"}
        tree["speedup_vs_published"] = 166.0 / d
    return tree


def range_leg(cz, sv, steps, forms=("plain", "persistent"), keep=None):
    """The single range [2^40, 2^40 + 2^36) with the k = 32 sieve and 2^24-wide chunk
    records, in the plain drop-in form and the persistent form of collatz_verify_range
    (keep, a dict, receives the plain form's chunk records)."""
    rng = {}
    by_name = {"plain": cz.FORM_PLAIN, "persistent": cz.FORM_PERSISTENT}
    for name in forms:
        form = by_name[name]
        sv.verify_range(RANGE_LO, RANGE_LO + RANGE_WIDTH, RANGE_CHUNK, form=form)
        ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            o = sv.verify_range(RANGE_LO, RANGE_LO + RANGE_WIDTH, RANGE_CHUNK, form=form)
            ms.append(o.device_ms)
        d = (time.perf_counter() - t0) / steps
        dms = sum(ms) / len(ms)
        Lr = words_for(RANGE_LO + RANGE_WIDTH)
        ach = (int(o.checks) * ops_dr(Lr) + int(o.total["count"]) * ops_setup(Lr)) / \
            (dms * 1e-3) / 1e12
        rng[name] = {"n_per_s": RANGE_WIDTH / d, "ms_per_call": d * 1e3,
                     "device_ms": dms, "device_n_per_s": RANGE_WIDTH / (dms * 1e-3),
                     "candidates": int(o.total["count"]), "checks": int(o.checks),
                     "roofline_frac": ach / VALU_PEAK_TOPS,
                     "hash": "%08X" % int(o.total["hash"])}
        log(f"range, {name}: {rng[name]}")
        if keep is not None and name == "plain":
            keep["c1_chunks"] = o.chunks.copy()
            keep["c1_rate"] = rng[name]["device_n_per_s"]
    return rng


def large_leg(cz, steps, keep=None):
    """Large sieve: [2^60, 2^60 + 2^44) with k = 37 (|S_37| records resident), 2^30-wide
    chunks, batched form (keep, a dict, receives its chunk records)."""
    t = time.time()
    s37 = cz.Sieve(37)
    build_s = time.time() - t
    lo, hi = 1 << 60, (1 << 60) + (1 << 44)
    s37.verify_batch([(lo, hi)], chunk_log2=30, use_graph=True)
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        o = s37.verify_batch([(lo, hi)], chunk_log2=30, use_graph=True)[0]
        ms.append(o.device_ms)
    d = (time.perf_counter() - t0) / steps
    kms = sum(ms) / len(ms)
    L37 = words_for(hi)
    ach = (int(o.checks) * ops_dr(L37) + int(o.total["count"]) * ops_setup(L37)) / \
        (kms * 1e-3) / 1e12
    out = {"range": "[2^60, 2^60+2^44)", "chunk_log2": 30, "n_per_s": (hi - lo) / (kms * 1e-3),
           "host_inclusive_n_per_s": (hi - lo) / d,
           "ms_per_step": d * 1e3, "kernel_ms": kms, "candidates": int(o.total["count"]),
           "checks": int(o.checks), "survivors": int(s37.info().survivors),
           "sieve_build_s": build_s, "escalated": int(o.escalated),
           "hash": "%08X" % int(o.total["hash"]), "roofline_frac": ach / VALU_PEAK_TOPS}
    log(f"large sieve k=37: {out}")
    if keep is not None:
        keep["c4_chunks"] = o.chunks.copy()
        keep["c4_rate"] = out["n_per_s"]
    return out


def c0_leg(cz, steps, keep=None):
    """[1, 2^32) with the k = 20 sieve and 2^24-wide records (the CPU-runnable
    configuration) in the plain drop-in form (keep, a dict, receives its records)."""
    s20 = cz.Sieve(20)
    o = s20.verify_range(1, 1 << 32, 24)
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        o = s20.verify_range(1, 1 << 32, 24)
        ms.append(o.device_ms)
    d = (time.perf_counter() - t0) / steps
    dms = sum(ms) / len(ms)
    ach = (int(o.checks) * ops_dr(3) + int(o.total["count"]) * ops_setup(3)) / \
        (dms * 1e-3) / 1e12
    out = {"range": "[1, 2^32)", "k": 20, "chunk_log2": 24, "form": "plain",
           "device_ms": dms, "ms_per_call": d * 1e3, "n_per_s": ((1 << 32) - 1) / (dms * 1e-3),
           "host_inclusive_n_per_s": ((1 << 32) - 1) / d, "candidates": int(o.total["count"]),
           "checks": int(o.checks), "roofline_frac": ach / VALU_PEAK_TOPS,
           "hash": "%08X" % int(o.total["hash"])}
    if keep is not None:
        keep["c0_chunks"] = o.chunks.copy()
        keep["c0_rate"] = out["n_per_s"]
    log(f"c0: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-passes", type=int, default=1)
    ap.add_argument("--range-steps", type=int, default=20)
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of the served replays at the prefix rates")
    ap.add_argument("--serve-rates", default=",".join(str(r) for r in RATES),
                    help="comma-separated request rates (req/s) of the served replays")
    ap.add_argument("--serve-full-rates", default=",".join(str(r) for r in FULL_RATES),
                    help="rates at which the whole 20000-request stream is replayed; the "
                         "others replay the first serve-seconds x rate requests")
    ap.add_argument("--serve-check", type=int, default=500,
                    help="served results compared with the plain kernel, per replay")
    ap.add_argument("--serve-all", action="store_true",
                    help="replay the whole 20000-request stream at every rate")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-range", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-lookahead", action="store_true")
    ap.add_argument("--no-tree", action="store_true")
    ap.add_argument("--tree-n", type=int, default=50,
                    help="MAXBITS of the reference-program leg (the whole tree below 2^N)")
    ap.add_argument("--only-tree", action="store_true",
                    help="run only the reference-program leg and print its record")
    ap.add_argument("--no-rust", action="store_true")
    ap.add_argument("--no-deep", action="store_true")
    ap.add_argument("--only-deep", action="store_true",
                    help="run only the n < 2^60 case-line leg and print it")
    ap.add_argument("--rust-n", default="40,50",
                    help="MAXBITS of the main.rs-configuration legs (comma-separated)")
    ap.add_argument("--only-rust", action="store_true",
                    help="run only the main.rs-configuration legs and print them")
    ap.add_argument("--only-lookahead", action="store_true",
                    help="run only the look-ahead leg (batch workload, B = 16) and print it")
    ap.add_argument("--only-range", default=None, metavar="FORMS",
                    help="run only the [2^40, 2^40+2^36) range leg in the given forms "
                         "(comma-separated: plain, persistent) and print it")
    ap.add_argument("--only-c0", action="store_true",
                    help="run only the [1, 2^32) leg at k = 20 (plain form) and print it")
    ap.add_argument("--only-large", action="store_true",
                    help="run only the k = 37 leg ([2^60, 2^60+2^44)) and print it")
    ap.add_argument("--large-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-intervals", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device tensor copy three times (the "
                         "reference reading for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "ok": bool((b == 1).all())}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this verifier runs on one GPU; --gpus must be 1. Ranges "
                          "shard across GPUs with no exchange during compute."}))
        sys.exit(2)
    import collatz_amd as cz
    from collatz_amd import workloads as W

    if args.only_tree:
        print(json.dumps(tree_leg(cz, args.tree_n)), flush=True)
        return
    if args.only_deep:
        print(json.dumps(deep_lines_leg(cz)), flush=True)
        return
    if args.only_rust:
        print(json.dumps(rust_leg(cz, [int(x) for x in args.rust_n.split(",") if x])),
              flush=True)
        return

    if args.only_large:
        print(json.dumps(large_leg(cz, args.large_steps)), flush=True)
        return
    if args.only_c0:
        print(json.dumps(c0_leg(cz, args.range_steps)), flush=True)
        return
    if args.only_range:
        s32 = cz.Sieve(K)
        print(json.dumps(range_leg(cz, s32, args.range_steps,
                                   tuple(x for x in args.only_range.split(",") if x))),
              flush=True)
        return

    ivs = W.batch_intervals(0, 4096, K)
    total_n = sum(hi - lo for lo, hi in ivs)
    if args.only_lookahead:
        print(json.dumps(lookahead_leg(cz, ivs, total_n, args.steps)), flush=True)
        return
    t = time.time()
    sv = cz.Sieve(K)
    info = sv.info()
    log(f"sieve k={K}: {int(info.survivors)} survivors in {time.time() - t:.1f} s")

    # Headline: the batched form.
    def batch_step():
        return sv.verify_batch(ivs, chunk_log2=BATCH_CHUNK, use_graph=True)

    for _ in range(args.warmup):
        batch_step()
    log(f"batched warmup done ({args.warmup} steps)")
    dev_ms = []
    device_sync()
    t0 = time.perf_counter()
    outs = None
    for _ in range(args.steps):
        outs = batch_step()
        dev_ms.append(outs[0].device_ms)
    device_sync()
    dt = time.perf_counter() - t0
    log(f"batched: {args.steps} steps in {dt:.2f} s")
    wall_ms_step = dt / args.steps * 1e3
    count = sum(int(o.total["count"]) for o in outs)
    checks = sum(int(o.checks) for o in outs)
    # Device time of a step (HIP events on the engine's stream): the set-up, verification
    # and reduction kernels of the launch and the three-limb pass, with every input already
    # in device memory.  The wall time of a call also covers host planning of the 4096
    # intervals and their hand-over and the records' return over PCIe: reported beside.
    kern_ms = sum(dev_ms) / len(dev_ms)
    ms_step = kern_ms
    Lw = words_for(max(hi for _, hi in ivs))
    ops = checks * ops_dr(Lw) + count * ops_setup(Lw)
    achieved = ops / (kern_ms * 1e-3) / 1e12
    value = total_n / (kern_ms * 1e-3)
    value_host = total_n / (dt / args.steps)
    h = 0
    for o in outs:
        h ^= int(o.total["hash"])

    line = {
        "metric": METRIC, "value": value, "unit": "n/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "batched: 4096 intervals (SplitMix64 seed 0), starts U[2^48,2^60),"
                               " widths log-U[2^20,2^34] rounded up to 2^32; k=32 sieve, mod 9 on,"
                               " 2x u64 limbs, 2^32-wide chunk records, one graph-replayed "
                               "persistent launch per step",
                   "form": "batched", "k": K, "chunk_log2": BATCH_CHUNK, "intervals": len(ivs),
                   "integers_per_step": total_n, "survivors": int(info.survivors)},
        "survivors_per_s": count / (kern_ms * 1e-3),
        "host_inclusive": {"n_per_s": value_host, "ms_per_step": wall_ms_step,
                           "note": "wall time of each collatz_verify_batch call: host planning,"
                                   " interval hand-over and record return over PCIe included"},
        "candidates": count, "checks": checks, "kernel_ms": kern_ms,
        "escalated": sum(int(o.escalated) for o in outs),
        "total_hash": "%08X" % h,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS,
                     "unit": "TOP/s", "frac": achieved / VALU_PEAK_TOPS,
                     "traffic": int(count * READ_BYTES_PER_CANDIDATE + WRITE_BYTES_PER_LAUNCH),
                     "hbm_GBps": (count * READ_BYTES_PER_CANDIDATE + WRITE_BYTES_PER_LAUNCH)
                                 / (kern_ms * 1e-3) / 1e9,
                     "kernel": f"k_verify_persistent<{Lw}>",
                     "ops": f"checks x (13L+7) + candidates x (4L+8), L={Lw}",
                     "peak_source": "published FP32 vector 157.3 TFLOP/s as lane-ops (FMA = 2 "
                                    "flop): two wave64 instructions per CU per cycle"},
        "cpu_baseline": None,
    }
    probe = valu_probe()
    if probe:
        line["roofline"]["measured"] = probe
        line["roofline"]["frac_of_checks_loop"] = achieved / probe["checks_loop_tops"]
        log(f"valu probe: {probe['vop2_simple_tops']:.1f} T (VOP2), {probe['vop3_tops']:.1f} T "
            f"(VOP3), check loop {probe['checks_loop_tops']:.1f} T")

    if not args.no_plain:
        plain_s = 0.0
        plain_dev_ms = 0.0
        same = True
        for p in range(args.plain_passes):
            t0 = time.perf_counter()
            pouts = [sv.verify_range(lo, hi, BATCH_CHUNK, form=cz.FORM_PLAIN) for lo, hi in ivs]
            plain_s += time.perf_counter() - t0
            plain_dev_ms += sum(o.device_ms for o in pouts)
            if p == 0:
                for a, b in zip(outs, pouts):
                    if not (records_equal(a.chunks, b.chunks) and
                            a.total.tobytes() == b.total.tobytes()):
                        same = False
        plain_rate = total_n / (plain_s / args.plain_passes)
        plain_dev_rate = total_n / (plain_dev_ms / args.plain_passes * 1e-3)
        line["plain"] = {"n_per_s": plain_rate, "ms_per_pass": plain_s / args.plain_passes * 1e3,
                         "device_n_per_s": plain_dev_rate,
                         "device_ms_per_pass": plain_dev_ms / args.plain_passes,
                         "calls_per_pass": len(ivs)}
        # Both forms by the wall time of their calls (4096 plain calls against one batch),
        # and by device time alone (the kernels of the 4096 calls against the batch's).
        line["batched_vs_plain"] = value_host / plain_rate
        line["batched_vs_plain_device"] = value / plain_dev_rate
        line["bit_exact_batched_vs_plain"] = same
        log(f"plain: {plain_rate:.3e} n/s ({plain_dev_rate:.3e} by device time), batched/plain "
            f"{value_host / plain_rate:.2f} ({value / plain_dev_rate:.2f} by device time), "
            f"identical records {same}")

    keep = {"sieve32": sv}
    if not args.no_range:
        line["range_2p40_w2p36"] = range_leg(cz, sv, args.range_steps, keep=keep)
        line["range_1_2p32_k20"] = c0_leg(cz, args.range_steps, keep=keep)

    if not args.no_large:
        line["large_sieve_k37"] = large_leg(cz, args.large_steps, keep=keep)

    if not args.no_lookahead:
        line["lookahead"] = lookahead_leg(cz, ivs, total_n, args.steps)
        line["lookahead"]["vs_headline"] = line["lookahead"]["n_per_s"] / value
        log(f"look-ahead B={LOOKAHEAD_BITS}: {line['lookahead']}")

    if not args.no_tree:
        line["reference_tree_mode"] = tree_leg(cz, args.tree_n)
        log(f"reference program, all n < 2^{args.tree_n}: {line['reference_tree_mode']}")

    if not args.no_rust:
        line["rust_mode"] = rust_leg(cz, [int(x) for x in args.rust_n.split(",") if x])

    if not args.no_deep:
        line["reference_n60_lines"] = deep_lines_leg(cz)
        log(f"reference program, n < 2^60 case lines: {line['reference_n60_lines']}")

    if not args.no_serve:
        reqs = W.request_stream(1, 20000)
        full = {int(x) for x in args.serve_full_rates.split(",") if x}
        # Untimed: the autotuner's sweeps for the (word width, magnitude) buckets the launches
        # of the stream fall in, once per sieve.  A server never sweeps inside a launch (it
        # runs the default shape and sweeps while idle, one timing launch at a time), so the
        # buckets are tuned here by one batched call each over a 2^32-wide range of that
        # magnitude (bit length / 8 of the end), which sweeps before it launches; then the
        # first 300 requests run one at a time (a launch per request, as at low rates) and
        # all at once (launches that mix many requests, as at high rates).
        t_warm = time.perf_counter()
        for c in sorted({hi.bit_length() // 8 for _, _, hi in reqs}):
            lo = max(1 << (8 * c), 1 << 40)
            sv.verify_batch([(lo, lo + (1 << 32))], chunk_log2=BATCH_CHUNK)
        srv = cz.Server(sv, chunk_log2=BATCH_CHUNK)
        for _, lo, hi in reqs[:300]:
            srv.wait(srv.submit(lo, hi))
        srv.close()
        served_leg(sv, reqs[:300], None, BATCH_CHUNK)
        served = {"warmup_s": time.perf_counter() - t_warm}
        log(f"served warm-up (autotuner sweeps) {served['warmup_s']:.1f} s")
        for rate in (int(x) for x in args.serve_rates.split(",") if x):
            n = len(reqs) if (args.serve_all or rate in full) else \
                min(len(reqs), max(50, int(rate * args.serve_seconds)))
            served[str(rate)] = served_leg(sv, reqs[:n], rate, BATCH_CHUNK, args.serve_check)
            log(f"served at {rate} req/s: {served[str(rate)]}")
        served["saturated"] = served_leg(sv, reqs, None, BATCH_CHUNK, args.serve_check)
        log(f"served saturated: {served['saturated']}")
        served["stream"] = ("SplitMix64 seed 1, 20000 requests: widths log-U[2^16,2^36], 80% of "
                            "starts U[2^40,2^62), 20% U[2^62,2^64); k=32 sieve resident, "
                            "2^32-wide records; each rate replays the first `requests` of the "
                            "stream, saturated submits all of them at once")
        line["served"] = served
        keep["c3_rate"] = served["saturated"]["n_per_s"]
        line["bit_exact_served_vs_plain"] = all(
            v["bit_exact_vs_plain"] for v in served.values() if isinstance(v, dict))

    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(ivs, threads, args.cpu_intervals, 32)
        log(f"cpu baseline: {line['cpu_baseline']['value']:.3e} n/s")
        if all(x in keep for x in ("c0_chunks", "c1_chunks", "c3_rate", "c4_chunks")):
            line["cpu_baseline"]["per_config"] = cpu_configs(threads, keep)
            log(f"cpu baselines per configuration: {line['cpu_baseline']['per_config']}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
