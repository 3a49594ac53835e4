"""Oracle fixtures for the large-sieve workload (k = 37 on [2^60, 2^60 + 2^44), 2^30-wide
chunks): |S_37|, an order-independent checksum of the survivor records, and the oracle's
chunk records for three slices of the range.  The CPU oracle needs about a minute and
~25 GB of host memory for the k = 37 sieve, too much for a test run, so its results are
committed here.

    make -C oracle && python tests/golden/make_k37_fixtures.py
Writes k37_fixtures.json.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import bind as B  # noqa: E402

K = 37
BASE = 1 << 60
SLICES = [
    (BASE, BASE + (1 << 30)),                                   # first chunk
    (BASE + (1 << 37) - (1 << 29), BASE + (1 << 37) + (1 << 29)),  # across a q boundary
    (BASE + (8197 << 30) + 12345, BASE + (8198 << 30) - 777),   # ragged, mid-range
]


def checksum(sv):
    """Sum and xor of a 64-bit mix of (r, T^k(r)+1, f) over all survivors."""
    r = sv.r.astype(np.uint64)
    m = sv.number.astype(np.uint64)
    f = sv.f.astype(np.uint64)
    with np.errstate(over="ignore"):
        h = r * np.uint64(0x9E3779B97F4A7C15) ^ m * np.uint64(0xBF58476D1CE4E5B9) ^ (f << np.uint64(56))
    return int(np.bitwise_xor.reduce(h)), int(h.sum(dtype=np.uint64))


def main():
    t = time.time()
    sv = B.sieve(K)
    print(f"sieve k={K}: {len(sv.r)} survivors in {time.time() - t:.1f} s", flush=True)
    x, s = checksum(sv)
    out = {"k": K, "survivors": len(sv.r), "checksum_xor": x, "checksum_sum": s, "slices": []}
    for lo, hi in SLICES:
        t = time.time()
        chunks, total = B.verify_range(sv, lo, hi, 30, threads=8)
        out["slices"].append({"lo": lo, "hi": hi, "chunk_log2": 30,
                              "chunks": [{k: (v.tolist() if hasattr(v, "tolist") else v)
                                          for k, v in zip(chunks.dtype.names, c)} for c in chunks]})
        print(f"slice [{lo:#x}, {hi:#x}): {int(total['count'])} candidates in "
              f"{time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, "k37_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
