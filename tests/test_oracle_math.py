"""The oracle against mathematics that fixes its answers: Lemma 2.3, the residue sieve's
prunes, the mod 9 classes, and naive step-by-step T iteration with Python integers."""
import random

import numpy as np
import pytest


def T(x):
    return x // 2 if x % 2 == 0 else (3 * x + 1) // 2


def T_iter(x, k):
    f = 0
    for _ in range(k):
        if x % 2:
            f += 1
        x = T(x)
    return x, f


@pytest.fixture(scope="module")
def s20(oracle):
    return oracle.sieve(20)


def test_sieve_numbers_are_T_k_plus_one(s20):
    """number = T^k(r) + 1 and f = number of odd steps, by naive iteration."""
    rng = random.Random(1)
    for i in rng.sample(range(len(s20.r)), 2000):
        x, f = T_iter(int(s20.r[i]), 20)
        assert int(s20.number[i]) == x + 1
        assert int(s20.f[i]) == f


def test_sieve_sorted_and_three_mod_four(s20):
    r = s20.r.astype(object)
    assert np.all(np.diff(s20.r.astype(np.int64)) > 0)
    assert all(int(x) % 4 == 3 for x in r[:5000])


@pytest.mark.parametrize("k,w", [(16, 10), (20, 14), (20, 4), (20, 20)])
def test_sieve_window_is_a_window_of_the_sieve(oracle, k, w):
    """Residue windows [top 2^w, (top + 1) 2^w) of the windowed walk (every one, or 200
    seeded ones where there are more) hold exactly the full sieve's survivors there, with
    the same numbers and odd-step counts."""
    full = oracle.sieve(k)
    tops = full.r >> np.uint64(w)
    n = 1 << (k - w)
    for top in (range(n) if n <= 256 else random.Random(k * 100 + w).sample(range(n), 200)):
        win = oracle.sieve_windows(k, w, [top])
        sel = tops == np.uint64(top)
        assert np.array_equal(win.r, full.r[sel]), (k, w, top)
        assert np.array_equal(win.number, full.number[sel])
        assert np.array_equal(win.f, full.f[sel])
    # Several windows at once: their union.
    rng = random.Random(k + w)
    for _ in range(5):
        ts = rng.sample(range(n), min(n, rng.choice([2, 5, 17])))
        win = oracle.sieve_windows(k, w, ts, threads=4)
        sel = np.isin(tops, np.array(ts, np.uint64))
        assert np.array_equal(win.r, full.r[sel]), (k, w, ts)
        assert np.array_equal(win.number, full.number[sel])


def test_lemma_2_3(s20):
    """T^k(q 2^k + r) = T^k(r) + q 3^f for q up to 2^90 (paper p.4 Lemma 2.3)."""
    rng = random.Random(2)
    for i in rng.sample(range(len(s20.r)), 300):
        r, f = int(s20.r[i]), int(s20.f[i])
        q = rng.randrange(1, 1 << rng.choice([8, 40, 90]))
        x, f2 = T_iter((q << 20) + r, 20)
        assert f2 == f
        assert x == int(s20.number[i]) - 1 + q * 3 ** f


def test_induction_prune_descends(oracle):
    """Residues cut by 485 f <= 306 k (collatz_cudav7.cu:507) descend below n after k steps
    once n is large: T^k(n) < n for n = q 2^k + r, q >= 2^16."""
    rng = random.Random(3)
    k = 16
    for _ in range(400):
        r = rng.randrange(3, 1 << k, 4)
        x, f = T_iter(r, k)
        if 485 * f <= 306 * k:
            n = (rng.randrange(1 << 16, 1 << 40) << k) + r
            assert T_iter(n, k)[0] < n


def test_mod9_rows_closed_form(oracle):
    for mbi in (20, 29, 35, 41):
        rows = oracle.mod9_rows(mbi, 10)
        bits = np.unpackbits(rows.view(np.uint8), bitorder="little").reshape(9, -1)
        for a in range(9):
            for i in range(0, 1024, 7):
                assert bits[a, i] == (((a + (i << mbi)) % 9) in (0, 1, 3, 6, 7))


@pytest.mark.parametrize("seed", [4, 5])
def test_trajectory_against_naive_iteration(oracle, seed):
    """or_traj from T^k(n)+1: the final value is T^J(n)+1 below n, and the max is the max of
    T^j(n)+1 over k <= j <= J, both computed step by step."""
    rng = random.Random(seed)
    k = 20
    for _ in range(300):
        n = rng.randrange(3, 1 << rng.choice([24, 40, 64, 90]))
        x, _ = T_iter(n, k)
        h, ovf, steps, mx, fin = oracle.traj(n, x + 1, cutoff=187, final=True)
        assert ovf == 0
        J = k + steps
        y = n
        best = 0
        for j in range(J + 1):
            if j >= k:
                best = max(best, y + 1)
            if j < J:
                y = T(y)
        assert fin == y + 1
        assert fin < n or n < 4
        assert mx == best


def test_range_against_brute_force(oracle, s20):
    """Chunk records of a small range: candidate set, counts, sums and maxima from naive
    per-n computation."""
    lo, hi, C = (5 << 20) + 12345, (5 << 20) + 12345 + 600000, 18
    chunks, total = oracle.verify_range(s20, lo, hi, C)
    members = set(int(x) for x in s20.r)
    exp = {}
    for n in range(lo, hi):
        j = n >> C
        e = exp.setdefault(j, dict(n=0, count=0, ssum=0, sxor=0, hash=0, mx=(0, 0)))
        e["n"] += 1
        if (n % (1 << 20)) not in members or (n % 9) not in (0, 1, 3, 6, 7):
            continue
        x, _ = T_iter(n, 20)
        h, _, steps, mx = oracle.traj(n, x + 1)
        J = 20 + steps
        e["count"] += 1
        e["ssum"] += J
        e["sxor"] ^= J
        e["hash"] ^= h
        cand = (mx - 1, -n)
        if e["count"] == 1 or cand > e["mx"]:
            e["mx"] = cand
    assert len(chunks) == len(exp)
    for ch in chunks:
        e = exp[int(ch["chunk_index"])]
        assert int(ch["n_checked"]) == e["n"]
        assert int(ch["count"]) == e["count"]
        assert int(ch["stop_sum"]) == e["ssum"]
        assert int(ch["stop_xor"]) == e["sxor"]
        assert int(ch["hash"]) == e["hash"]
        mexc = int(ch["max_exc"][0]) | (int(ch["max_exc"][1]) << 64) | (int(ch["max_exc"][2]) << 128)
        mst = int(ch["max_start"][0]) | (int(ch["max_start"][1]) << 64)
        assert (mexc, -mst) == e["mx"]
