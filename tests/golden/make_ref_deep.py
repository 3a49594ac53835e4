"""Writes tests/golden/ref_deep.json: reference-mode results of case lines whose leaves sit
deeper than 40 bits (N - A = 44, 45, 50), computed by the oracle (u128 throughout), for the
GPU tests of the 128-bit leaf path.

Lines: the sampled cases_big lines at sigfig 23 with threepower 15, the sample's smallest
subtrees (112,311 leaves at depth 45, 2,435,885 at depth 50).  Runs:
  * N = 55 and N = 60, A = 10, 16 bit vectors, the CUDA program's compares, CUTOFFPOWER 100
    (the reference's defaults apart from MAXBITS; paper p.3-4 runs n < 2^55 and n < 2^60);
  * N = 60 with CUTOFFPOWER 76, where trajectories of starts near 2^60 overflow;
  * N = 50, A = 6, 8 bit vectors, main.rs semantics (its own search3: numcases and its
    2^120 overflow test, and at 2^72 to make overflows happen).
Usage: python tests/golden/make_ref_deep.py [threads]
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import bind as B  # noqa: E402

SEM_CUDA, SEM_RUST = 0, 1


def lines():
    out = []
    with open(os.path.join(HERE, "cases_big_sample.txt")) as f:
        for ln in f:
            t = [int(x) for x in ln.split()]
            if t and t[1] == 23 and t[2] == 15:
                out.append(t)
    return out


def starts_digest(starts):
    return hashlib.sha256("\n".join(str(x) for x in sorted(starts)).encode()).hexdigest()


def run(t, N, A, num_bv, sem, cutoff, threads):
    kn = B.knobs(N, A, 24, num_bv, sem, cutoff)
    bv = B.precompute_bv(kn)
    m9 = B.mod9_rows(N - A, A)
    out, starts = B.case_run_starts(kn, bv, m9, B.Node.make(*t), cap=1 << 22, threads=threads)
    assert len(starts) == out.overflow_count
    return {"line": t, "N": N, "A": A, "B": 24, "num_bv": num_bv, "semantics": sem,
            "cutoff": cutoff, "complete_hash": out.complete_hash, "candidates": out.candidates,
            "leaves": out.leaves, "bin_leaves": list(out.bin_leaves), "launches": out.launches,
            "overflow_count": out.overflow_count, "numcases": out.numcases,
            "overflow_starts_sha256": starts_digest(starts)}


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count() or 1
    ls = lines()
    runs = []
    for t in ls[:3]:
        runs.append(run(t, 55, 10, 16, SEM_CUDA, 100, threads))
    for t in ls[3:5]:
        runs.append(run(t, 60, 10, 16, SEM_CUDA, 100, threads))
    runs.append(run(ls[5], 60, 10, 16, SEM_CUDA, 76, threads))
    for t, cut in ((ls[6], 120), (ls[7], 72)):
        runs.append(run(t, 50, 6, 8, SEM_RUST, cut, threads))
    with open(os.path.join(HERE, "ref_deep.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_ref_deep.py", "runs": runs}, f, indent=1)
        f.write("\n")
    for r in runs:
        print(r["N"], r["A"], r["semantics"], r["cutoff"], r["leaves"], r["candidates"],
              r["overflow_count"], hex(r["complete_hash"]), r["numcases"])


if __name__ == "__main__":
    main()
