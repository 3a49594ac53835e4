"""Complete hash of the whole N=45 run of the reference (collatz_openclv2.cpp, compiled
as is into oracle/_ref/libref.so), as the XOR over all 278,699 lines of cases_big, which
partition the search1 tree (every line runs independently, as the paper's processes do).

Run in the build container (needs /root/reference):
    make -C oracle && python tests/golden/make_ref_full45.py
Writes ref_full45.json: the complete hash, the candidate-free totals (leaves per bin) and the
number of lines.
"""
import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def work(lines):
    from oracle import bind as B
    h = 0
    bins = [0] * 16
    bad = []
    for line in lines:
        t = [int(x) for x in line.split()]
        r = B.ref_case(t, cap=0)
        if r is None:
            bad.append(t)
            continue
        h ^= r["hash"]
        for i in range(16):
            bins[i] += r["bins"][i]
    return h, bins, bad


def main():
    lines = open("/root/reference/cases_big").read().split("\n")
    lines = [l for l in lines if l.strip()]
    nproc = min(int(os.environ.get("NPROC", "8")), os.cpu_count() or 1)
    chunks = [lines[i::64] for i in range(64)]
    with mp.Pool(nproc) as pool:
        res = pool.map(work, chunks)
    h = 0
    bins = [0] * 16
    bad = []
    for hh, bb, bd in res:
        h ^= hh
        bins = [a + b for a, b in zip(bins, bb)]
        bad += bd
    out = {"N": 45, "A": 10, "B": 24, "semantics": "opencl", "lines": len(lines),
           "complete_hash": h, "bin_leaves": bins, "lines_too_large": bad}
    with open(os.path.join(HERE, "ref_full45.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "lines_too_large"}),
          "too large:", len(bad))


if __name__ == "__main__":
    main()
