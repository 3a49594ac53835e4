"""The benchmark workloads at their real sizes against the oracle (which runs on the host's
cores): the whole config-1 range, and the 4096-interval batch with a seeded sample of its
intervals checked at full width."""
import random

import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def s32():
    import collatz_amd as cz
    return cz.Sieve(32)


@pytest.fixture(scope="module")
def o32(oracle):
    return oracle.sieve(32)


def test_config1_whole_range_against_oracle(s32, o32, oracle):
    """[2^40, 2^40 + 2^36), k = 32, 2^24-wide records (4096 of them), both forms."""
    lo, hi = 1 << 40, (1 << 40) + (1 << 36)
    chunks, total = oracle.verify_range(o32, lo, hi, 24, threads=16)
    for form in (0, 1):
        g = s32.verify_range(lo, hi, 24, form=form)
        assert_chunks_equal(g.chunks, chunks, f"form {form}")
        assert_total_equal(g.total, total, "total")


def test_batch_workload_sample_against_oracle(s32, o32, oracle):
    """The whole batch workload (4096 intervals, 2^32-wide records, one graph-replayed
    launch); 64 seeded intervals at full width: their totals against the oracle's, and
    their records with 2^26-wide chunks (a second batch of just those intervals)."""
    from collatz_amd import workloads as W
    ivs = W.batch_intervals(0, 4096, 32)
    outs = s32.verify_batch(ivs, chunk_log2=32, use_graph=True)
    idx = sorted(random.Random(3).sample(range(len(ivs)), 64))
    sample = [ivs[i] for i in idx]
    fine = s32.verify_batch(sample, chunk_log2=26)
    for i, (lo, hi), f in zip(idx, sample, fine):
        chunks, total = oracle.verify_range(o32, lo, hi, 26, threads=16)
        assert_chunks_equal(f.chunks, chunks, f"interval {i}")
        for field in ("count", "stop_sum", "stop_xor", "max_exc", "max_start",
                      "overflow_count", "hash", "flags", "n_checked"):
            assert (outs[i].total[field] == total[field]).all(), (i, field)
