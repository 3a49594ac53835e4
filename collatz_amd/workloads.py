"""Synthetic workloads of the benchmark, drawn from a frozen generator.

The generator is SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): state += 0x9E3779B97F4A7C15, then two xor-shift-multiply rounds.
Every draw below takes whole 64-bit outputs in a fixed order, so the lists are the same on
every machine; tests/golden holds the interval list and a digest of the request stream.
"""
from __future__ import annotations

import hashlib
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def unit(self) -> float:
        """Uniform in [0, 1) with 53 bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, n: int) -> int:
        """Uniform integer in [0, n) for n <= 2^128 (two draws, then a reduction whose bias
        is below 2^-64 for the spans used here)."""
        x = (self.next() << 64) | self.next()
        return x % n


def log_uniform_width(rng: SplitMix64, lo_log2: int, hi_log2: int) -> int:
    """2^(lo + (hi - lo) u), u uniform in [0, 1), as an integer in [2^lo, 2^hi]."""
    w = int(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * rng.unit()))
    return max(1 << lo_log2, min(w, 1 << hi_log2))


def batch_intervals(seed: int = 0, count: int = 4096, k: int = 32):
    """The batched workload: `count` intervals; starts uniform in [2^48, 2^60), widths
    log-uniform in [2^20, 2^34] rounded up to a multiple of 2^k."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        lo = (1 << 48) + rng.below((1 << 60) - (1 << 48))
        w = log_uniform_width(rng, 20, 34)
        w = -(-w >> k) << k
        out.append((lo, lo + w))
    return out


def request_stream(seed: int = 1, count: int = 20000):
    """The served workload: `count` requests as (gap, lo, hi), where gap is the unit-rate
    exponential inter-arrival time (divide by the rate in requests/s for seconds).  Widths
    are log-uniform in [2^16, 2^36]; 80% of starts are uniform in [2^40, 2^62) and 20% in
    [2^62, 2^64)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        gap = -math.log(1.0 - rng.unit())
        w = log_uniform_width(rng, 16, 36)
        if rng.unit() < 0.8:
            lo = (1 << 40) + rng.below((1 << 62) - (1 << 40))
        else:
            lo = (1 << 62) + rng.below((1 << 64) - (1 << 62))
        out.append((gap, lo, lo + w))
    return out


def digest(items) -> str:
    h = hashlib.sha256()
    for it in items:
        h.update((" ".join(repr(x) for x in it) + "\n").encode())
    return h.hexdigest()
