"""The oracle's range-mode look-ahead (Step 2 at depth k) against the reference-mode run of
the whole tree, and the facts its index rests on."""
import random

import numpy as np
import pytest

ROOT_NODE = (3, 2, 2, 9, 1, 0)


@pytest.mark.parametrize("sem", [0, 2])
def test_range_lookahead_is_the_reference_candidate_set(oracle, sem):
    """Over [1, 2^N) with k = N - A, B = 24, the range form with the look-ahead iterates
    exactly the candidates of the reference's program run from the root: the same count and
    the same complete hash (collatz_cudav7.cu:95-183; or_case_run is checked against the
    compiled reference)."""
    N, A = 32, 8
    kn = oracle.knobs(N, A, semantics=sem)
    bv = oracle.precompute_bv(kn)
    m9 = oracle.mod9_rows(N - A, A)
    ref = oracle.case_run(kn, bv, m9, oracle.Node.make(*ROOT_NODE))
    sv = oracle.sieve(N - A)
    chunks, tot = oracle.verify_range(sv, 1, 1 << N, 28, cutoff=100, threads=8,
                                      lookahead_bits=24, lookahead_semantics=sem)
    assert int(tot["count"]) == ref.candidates
    assert int(tot["hash"]) == ref.complete_hash
    # Without the look-ahead the range form iterates a strict superset.
    _, plain = oracle.verify_range(sv, 1, 1 << N, 28, cutoff=100, threads=8)
    assert int(plain["count"]) > int(tot["count"])


def test_low_bits_decide_the_first_steps(oracle):
    """B steps of the shifted map (even: 3x/2, odd: floor(x/2) + 1) on any number x take the
    same parities as on x mod 2^B, so the look-ahead indexes by number mod 2^B."""
    rng = random.Random(5)
    for _ in range(300):
        B = rng.choice([8, 16, 24])
        x = rng.getrandbits(rng.choice([40, 64, 100, 190])) | (1 << 30)
        y = x % (1 << B)
        for _ in range(B):
            assert x % 2 == y % 2
            if x % 2 == 0:
                x, y = 3 * x // 2, 3 * y // 2
            else:
                x, y = x // 2 + 1, y // 2 + 1


def test_dip_bounds(oracle):
    """The dip is at most 306 B + 179 (all B steps halvings, then the path-merge term of an
    even-length run), which the residue 1 (a fixed point of floor(x/2) + 1) reaches for even
    B; an odd residue dips at least 306 (its first step halves)."""
    for B in (8, 16, 24):
        assert oracle.la_dip(1, B, oracle.SEM_CXX) == 306 * B + 179
        assert oracle.la_dip(0, B, oracle.SEM_CXX) == 0
        rng = random.Random(B)
        for _ in range(2000):
            n0 = rng.getrandbits(B)
            d = oracle.la_dip(n0, B, oracle.SEM_CXX)
            assert 0 <= d <= 306 * B + 179
            if n0 % 2 == 1:
                assert d >= 306


def test_chunks_are_subsets(oracle):
    """Per chunk, the look-ahead keeps a subset of the candidates: same integers checked,
    fewer or equal candidates."""
    sv = oracle.sieve(20)
    lo, hi = (1 << 40) + 12345, (1 << 40) + (1 << 26)
    a, _ = oracle.verify_range(sv, lo, hi, 22, threads=8, lookahead_bits=16)
    b, _ = oracle.verify_range(sv, lo, hi, 22, threads=8)
    assert np.array_equal(a["n_checked"], b["n_checked"])
    assert np.all(a["count"] <= b["count"]) and a["count"].sum() < b["count"].sum()
