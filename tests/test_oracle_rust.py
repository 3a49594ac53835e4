"""main.rs's own Step 3 and outputs (main.rs:321-341, :97-98): the oracle's table-driven
search3 against the shifted map iterated one step at a time with Python integers, and the
numcases / overflow outputs of whole case runs under main.rs semantics against the same
brute force over every candidate."""
import random

import pytest

from conftest import read_cases

ROOT_NODE = (3, 2, 2, 9, 1, 0)


def shifted_step(m):
    """One step of T on number = T-value + 1 (main.rs:115-120)."""
    return m + m // 2 if m % 2 == 0 else m // 2 + 1


def search3_brute(base, number, cutoff=120):
    """(overflowed, tally): the first multiple of 16 steps at which number < base, or
    number >= 2^cutoff (main.rs:333-339)."""
    tally = 0
    while True:
        for _ in range(16):
            number = shifted_step(number)
        tally += 1
        if number < base:
            return False, tally
        if number >= 1 << cutoff:
            return True, tally


def test_search3_matches_step_by_step(oracle):
    rng = random.Random(5)
    cases = []
    for _ in range(3000):
        n = rng.randrange(3, 1 << rng.choice([20, 40, 60, 64, 72, 90]))
        # A start value above n, as after search1's sigfig steps: n + something.
        v = n + rng.randrange(1, 1 << rng.choice([4, 20, 40]))
        cases.append((n, v, rng.choice([64, 80, 100, 120, 150])))
    # Start values just below the cutoff (main.rs's u128 would wrap above 2^118.6 there).
    for c in (100, 119, 120, 121, 150):
        for d in range(1, 40):
            cases.append(((1 << (c - 8)) + d, (1 << c) - (d << (c - 12)), c))
    for n, v, c in cases:
        assert oracle.rust_search3(n, v, c) == search3_brute(n, v, c), (n, v, c)


def test_rust_case_run_numcases_by_brute_force(oracle):
    """numcases and the overflow reports of a whole run (N = 26, A = 6, 8 bit vectors,
    main.rs semantics) equal the brute force over the candidates the run iterates, at
    main.rs's cutoff and at a low one that makes overflows happen."""
    for cutoff in (120, 40):
        kn = oracle.knobs(26, 6, num_bv=8, semantics=oracle.SEM_RUST, cutoff=cutoff)
        bv = oracle.precompute_bv(kn)
        m9 = oracle.mod9_rows(20, 6)
        out, starts = oracle.case_run_starts(kn, bv, m9, oracle.Node.make(*ROOT_NODE))
        assert out.complete_hash == 0 and out.candidates > 1000
        # The candidates: leaves of search1 at depth 20, filtered as search2 does.
        leaves = oracle.search1(oracle.Node.make(*ROOT_NODE), 20, oracle.SEM_RUST)
        mbi, tpo = 20, oracle.tpo(20)
        want_cases, want_starts, ncand = 0, [], 0
        bvw = [[int(w) for w in row] for row in bv]
        m9w = [[int(w) for w in row] for row in m9]
        for lf in leaves:
            f, number, base = lf.threepower, lf.number, lf.base
            b = f - tpo
            idx = (oracle.magic(f) * (number & 0xFFFFFFFF)) & ((1 << 24) - 1) if b < 8 else 0
            for a in range(64):
                if not (m9w[base % 9][a // 64] >> (a % 64)) & 1:
                    continue
                if b < 8:
                    i = (idx + a) & ((1 << 24) - 1)
                    if not (bvw[b][i // 64] >> (i % 64)) & 1:
                        continue
                nl = number & ((1 << 64) - 1)
                if lf.doubleeven == 2 and (nl + a) % 2 == 1:
                    continue
                if lf.doubleeven == 1:
                    if f % 2 == 0 and (nl + a) % 4 == 1:
                        continue
                    if f % 2 == 1 and (nl + 3 * a) % 4 == 1:
                        continue
                ncand += 1
                n = base + (a << mbi)
                ovf, t = search3_brute(n, number + a * 3 ** f, cutoff)
                want_cases += t
                if ovf:
                    want_starts.append(n)
        assert out.candidates == ncand
        assert out.numcases == want_cases
        assert out.overflow_count == len(want_starts) == len(starts)
        assert sorted(starts) == sorted(want_starts)
        if cutoff == 40:
            assert want_starts


def test_deep_fixtures_match_the_oracle(oracle):
    """tests/golden/ref_deep.json (leaves at depth 45 and 44) is what the oracle computes
    now: an N = 55 line (reference semantics) and both main.rs-semantics N = 50 lines."""
    import hashlib
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_deep.json")
    runs = json.load(open(path))["runs"]
    for f in [runs[0]] + [r for r in runs if r["semantics"] == oracle.SEM_RUST]:
        kn = oracle.knobs(f["N"], f["A"], f["B"], f["num_bv"], f["semantics"], f["cutoff"])
        out, starts = oracle.case_run_starts(kn, oracle.precompute_bv(kn),
                                             oracle.mod9_rows(f["N"] - f["A"], f["A"]),
                                             oracle.Node.make(*f["line"]), cap=1 << 20,
                                             threads=8)
        for k in ("complete_hash", "candidates", "leaves", "launches", "overflow_count",
                  "numcases"):
            assert getattr(out, k) == f[k], (f["line"], k)
        assert list(out.bin_leaves) == f["bin_leaves"]
        d = hashlib.sha256("\n".join(str(x) for x in sorted(starts)).encode()).hexdigest()
        assert d == f["overflow_starts_sha256"]
