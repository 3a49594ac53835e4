"""Reference mode under seeded random knobs (N, A, B, number of bit vectors, compare
semantics, cutoff): the whole tree from the root against the oracle's restatement of the
reference program (complete hash, candidates, leaves per bin, overflow count, launches)."""
import random

import pytest

pytestmark = pytest.mark.gpu
ROOT_NODE = (3, 2, 2, 9, 1, 0)


def configs(seed, n):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        sem = rng.choice([0, 1, 2])
        N = rng.randrange(24, 33)
        A = rng.randrange(1, 13)
        B = rng.choice([b for b in (8, 12, 16, 20) if b >= A])   # the ABI requires A <= B
        num_bv = 8 if sem == 1 else rng.choice([4, 16, 24])
        cutoff = rng.choice([64, 66, 72, 90, 95, 96, 100, 123, 128, 187])
        out.append((N, A, B, num_bv, sem, cutoff))
    return out


@pytest.mark.parametrize("seed", [11, 12])
def test_random_reference_runs(oracle, seed):
    import collatz_amd as cz
    root = oracle.Node.make(*ROOT_NODE)
    for N, A, B, num_bv, sem, cutoff in configs(seed, 12):
        kn = oracle.knobs(N, A, B, num_bv, sem, cutoff)
        want = oracle.case_run(kn, oracle.precompute_bv(kn), oracle.mod9_rows(N - A, A), root)
        got = cz.reference_run(ROOT_NODE, N, A, B, num_bv=num_bv, semantics=sem,
                               cutoff_log2=cutoff, overflow_cap=1 << 16)
        tag = f"N={N} A={A} B={B} bv={num_bv} sem={sem} cutoff={cutoff}"
        assert got.complete_hash == want.complete_hash, tag
        assert got.numcases == want.numcases, tag
        assert got.candidates == want.candidates, tag
        assert list(got.bin_leaves) == list(want.bin_leaves), tag
        assert got.overflow_count == want.overflow_count, tag
        assert got.launches == want.launches, tag
        assert got.leaves == want.leaves, tag
