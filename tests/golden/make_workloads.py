"""Writes the benchmark's synthetic workload fixtures from collatz_amd.workloads:
batch_intervals_seed0.txt (the 4096 intervals, "lo hi" in decimal) and
request_stream_seed1.json (a digest of the 20,000 requests and the first 64 of them).

    python tests/golden/make_workloads.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from collatz_amd import workloads as W  # noqa: E402


def main():
    ivs = W.batch_intervals(0, 4096, 32)
    with open(os.path.join(HERE, "batch_intervals_seed0.txt"), "w") as f:
        for lo, hi in ivs:
            f.write(f"{lo} {hi}\n")
    reqs = W.request_stream(1, 20000)
    out = {"seed": 1, "count": len(reqs), "sha256": W.digest(reqs),
           "first": [[g, lo, hi] for g, lo, hi in reqs[:64]]}
    with open(os.path.join(HERE, "request_stream_seed1.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
