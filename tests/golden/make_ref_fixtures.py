"""Writes golden vectors from the reference's own CPU code (oracle/_ref/libref.so, built by
oracle/Makefile from /root/reference/collatz_openclv2.cpp) and copies the reference's case
data used by the tests.

Run in the build container, where /root/reference exists:
    make -C oracle && python tests/golden/make_ref_fixtures.py

Outputs (data only): ref_vectors.json, cases_tiny.txt, cases_big_sample.txt.
"""
import hashlib
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import bind as B  # noqa: E402

REF = "/root/reference"


def main():
    if not B.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built")
    consts = B.ref_constants()
    out = {"constants": consts}
    # Bit vectors and mod 9 rows (N=45, A=10, B=24), as sha256 of the little-endian u32 words.
    out["bv_sha256"] = [hashlib.sha256(B.ref_bv(i).tobytes()).hexdigest() for i in range(16)]
    out["bv_popcount"] = [int(np.unpackbits(B.ref_bv(i)[: 1 << 19].view(np.uint8)).sum())
                          for i in range(16)]
    out["mod9"] = B.ref_mod9().tolist()
    out["magic"] = [int(x) for x in B.ref_magic()]
    # search3 on random and on crafted overflowing inputs (CUTOFFPOWER = 100).
    rng = random.Random(12345)
    s3 = []
    for _ in range(300):
        base = rng.randrange(3, 1 << 62)
        num = rng.randrange(2, 1 << rng.choice([40, 64, 90, 97]))
        s3.append([base, num, B.ref_search3(base, num)])
    for _ in range(40):
        base = rng.randrange(1 << 90, 1 << 96)
        num = rng.randrange(1 << 96, 1 << 99) & ~0xFF
        s3.append([base, num, B.ref_search3(base, num)])
    out["search3"] = [[str(a), str(b), c] for a, b, c in s3]
    # Whole case lines.
    tiny = open(os.path.join(REF, "cases_tiny")).read().splitlines()
    big = open(os.path.join(REF, "cases_big")).read().splitlines()
    rng = random.Random(2024)
    sample = sorted(rng.sample(range(len(big)), 240))
    lines = [big[i] for i in sample]
    cases = []
    for line in lines:
        t = [int(x) for x in line.split()]
        r = B.ref_case(t)
        if r is None:
            continue
        h = hashlib.sha256()
        for row in sorted(tuple(int(x) for x in rr) for rr in r["leaves"]):
            h.update(("%d %d %d\n" % row).encode())
        cases.append({"case": t, "hash": r["hash"], "launches": r["launches"],
                      "bins": r["bins"], "leaves": r["nleaves"], "leaves_sha256": h.hexdigest()})
    out["cases"] = cases
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    with open(os.path.join(HERE, "cases_tiny.txt"), "w") as f:
        f.write("\n".join(tiny) + "\n")
    with open(os.path.join(HERE, "cases_big_sample.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {len(cases)} case vectors, {len(s3)} search3 vectors")


if __name__ == "__main__":
    main()
