"""collatz_verify_batch: many intervals in one persistent launch; each interval's records
equal the oracle's (and so the plain form's)."""
import random

import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def s20():
    import collatz_amd as cz
    return cz.Sieve(20)


@pytest.fixture(scope="module")
def o20(oracle):
    return oracle.sieve(20)


def intervals(seed, n):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        lo = rng.randrange(1, 1 << rng.choice([30, 40, 62, 64, 70]))
        out.append((lo, lo + rng.randrange(1, 1 << rng.choice([10, 20, 24]))))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_oracle(s20, o20, oracle, graph):
    ivs = intervals(7, 80)
    res = s20.verify_batch(ivs, chunk_log2=22, use_graph=graph)
    for (lo, hi), r in zip(ivs, res):
        chunks, total = oracle.verify_range(o20, lo, hi, 22)
        assert_chunks_equal(r.chunks, chunks, f"[{lo:#x},{hi:#x})")
        assert_total_equal(r.total, total)


def test_graph_replay_repeats(s20):
    """Every replay of the captured launch sequence gives the records of the eager launch,
    with other eager work and synchronisations between the replays."""
    ivs = intervals(8, 30)
    eager = s20.verify_batch(ivs, chunk_log2=20, use_graph=False)
    for i in range(4):
        again = s20.verify_batch(ivs, chunk_log2=20, use_graph=True)
        for a, b in zip(eager, again):
            assert_chunks_equal(b.chunks, a.chunks, f"replay {i}")
            assert_total_equal(b.total, a.total, f"replay {i}")
        s20.verify_range(1, 1 << 26)
        s20.verify_batch(ivs[:3], chunk_log2=20, use_graph=False)


@pytest.mark.parametrize("tile,waves,U", [(64, 2, 8), (256, 4, 1), (4096, 16, 2), (65536, 28, 3)])
def test_launch_shapes(s20, o20, oracle, tile, waves, U):
    lo, hi = 1 << 34, (1 << 34) + (1 << 27)
    chunks, _ = oracle.verify_range(o20, lo, hi, 24)
    r = s20.verify_batch([(lo, hi)], chunk_log2=24, tile_residues=tile, waves_per_cu=waves,
                         checks_per_refill=U)[0]
    assert_chunks_equal(r.chunks, chunks, f"shape {tile},{waves},{U}")


def test_overlapping_and_duplicate_intervals(s20, o20, oracle):
    ivs = [(1000, 5000000), (1000, 5000000), (4000000, 9000000), ((1 << 62) + 7, (1 << 62) + 99999)]
    res = s20.verify_batch(ivs, chunk_log2=21, cutoff_log2=70, overflow_cap=4096)
    for (lo, hi), r in zip(ivs, res):
        chunks, total = oracle.verify_range(o20, lo, hi, 21, 70)
        assert_chunks_equal(r.chunks, chunks)
        assert len(r.overflow_starts) == int(total["overflow_count"])
        assert all(lo <= n < hi for n in r.overflow_starts)
        single = s20.verify_range(lo, hi, 21, cutoff_log2=70)
        assert sorted(single.overflow_starts) == sorted(r.overflow_starts)


def test_overflow_count_exact_beyond_the_list(s20):
    """More overflow reports than the device list keeps (cutoff 2^64 just above starts near
    2^62): every form reports the exact count in overflow_written, with no start buffer and
    with one that is too short (then the call fails with ECAPACITY and still reports it)."""
    import ctypes as C

    import collatz_amd as cz
    lo, hi = 1 << 62, (1 << 62) + (1 << 24)
    plain = s20.verify_range(lo, hi, 22, cutoff_log2=64, overflow_cap=1 << 20)
    n = int(plain.total["overflow_count"])
    assert n > 4096 and plain.overflow_written == n == len(plain.overflow_starts)
    assert all(lo <= x < hi for x in plain.overflow_starts)
    pers = s20.verify_range(lo, hi, 22, cutoff_log2=64, form=cz.FORM_PERSISTENT,
                            overflow_cap=1 << 20)
    assert pers.overflow_written == n
    assert sorted(pers.overflow_starts) == sorted(plain.overflow_starts)
    # Batched, no start buffers (the list of the launch keeps 1024 reports).
    other = (lo + (1 << 30), lo + (1 << 30) + (1 << 20))
    b = s20.verify_batch([(lo, hi), other], chunk_log2=22, cutoff_log2=64)
    assert b[0].overflow_written == n
    assert b[1].overflow_written == int(b[1].total["overflow_count"]) > 0
    for form in (cz.FORM_PLAIN, cz.FORM_PERSISTENT):
        res = cz.RangeResult()
        buf = (C.c_uint64 * 64)()
        res.overflow_starts = C.cast(buf, C.POINTER(C.c_uint64))
        res.overflow_cap = 32
        opts = cz.VerifyOpts(22, 2, 64, form)
        rc = cz.load().collatz_verify_range(s20.handle, cz.u128_pair(lo), cz.u128_pair(hi),
                                            C.byref(opts), None, C.byref(res))
        assert rc == cz.ECAPACITY and res.overflow_written == n, (form, rc)
        assert all(lo <= (buf[2 * i] | (buf[2 * i + 1] << 64)) < hi for i in range(32))
