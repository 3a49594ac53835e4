"""Served form: requests submitted while others run; every result equals the same range
verified alone (collatz_verify_range) and the oracle."""
import random
import threading
import time

import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def s20():
    import collatz_amd as cz
    return cz.Sieve(20)


def requests(seed, n):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        # Magnitudes that run on 96 bits, on two limbs and on three limbs, so launches
        # change word width as the mix changes.
        lo = rng.randrange(1, 1 << rng.choice([40, 62, 64, 70, 99, 112]))
        out.append((lo, lo + rng.randrange(1 << 10, 1 << rng.choice([16, 22, 26]))))
    return out


def test_results_alone_equal_in_the_mix(s20, oracle):
    import collatz_amd as cz
    reqs = requests(3, 60)
    srv = cz.Server(s20, chunk_log2=22, slice_candidates=1 << 16)
    tickets = []
    # Submit from another thread in bursts, so requests arrive while launches run.
    def producer():
        for i, (lo, hi) in enumerate(reqs):
            tickets.append(srv.submit(lo, hi))
            if i % 7 == 0:
                time.sleep(0.002)
    th = threading.Thread(target=producer)
    th.start()
    th.join()
    o20 = oracle.sieve(20)
    for t, (lo, hi) in zip(tickets, reqs):
        total, chunks, times = srv.wait(t, 60000)
        alone = s20.verify_range(lo, hi, 22)
        assert_chunks_equal(chunks, alone.chunks, f"request [{lo:#x},{hi:#x})")
        ochunks, ototal = oracle.verify_range(o20, lo, hi, 22)
        assert_chunks_equal(chunks, ochunks)
        assert_total_equal(total, ototal)
        assert times[0] <= times[1] <= times[2]
    srv.close()


def test_server_overlaps_requests(s20):
    """Short requests submitted behind a long one share its launches and finish first."""
    import collatz_amd as cz
    srv = cz.Server(s20, chunk_log2=20, slice_candidates=1 << 16)
    big = srv.submit(1 << 40, (1 << 40) + (1 << 28))
    small = [srv.submit((1 << 50) + i * (1 << 20), (1 << 50) + (i + 1) * (1 << 20)) for i in range(5)]
    tb = srv.wait(big, 60000)[2]
    ts = [srv.wait(t, 60000)[2] for t in small]
    assert all(s[2] < tb[2] for s in ts)
    srv.close()


def test_server_bad_request(s20):
    import collatz_amd as cz
    srv = cz.Server(s20)
    with pytest.raises(cz.CollatzError):
        srv.submit(10, 10)
    srv.close()


def test_admission_budget_serialises(s20):
    """With a device budget smaller than any request, requests are admitted one at a time
    (a request is always admitted when nothing else is in flight) and still all finish with
    the records they have alone."""
    import collatz_amd as cz
    srv = cz.Server(s20, chunk_log2=20, max_device_bytes=1)
    reqs = requests(5, 12)
    tickets = [srv.submit(lo, hi) for lo, hi in reqs]
    prev_done = -1.0
    for t, (lo, hi) in zip(tickets, reqs):
        total, chunks, (ts, ta, td) = srv.wait(t, 60000)
        assert ta >= prev_done - 1e-6  # admitted only after the previous one finished
        prev_done = td
        assert_chunks_equal(chunks, s20.verify_range(lo, hi, 20).chunks, f"[{lo:#x},{hi:#x})")
    srv.close()


def test_two_servers_and_direct_calls_share_a_sieve(s20):
    """Two servers and direct calls on one sieve handle at once (each server owns its own
    engine; direct persistent calls serialise on the sieve's engine)."""
    import collatz_amd as cz
    a, b = cz.Server(s20, chunk_log2=22), cz.Server(s20, chunk_log2=22)
    reqs = requests(9, 16)
    ta = [a.submit(lo, hi) for lo, hi in reqs[:8]]
    tb = [b.submit(lo, hi) for lo, hi in reqs[8:]]
    errors = []

    def direct():
        try:
            for lo, hi in reqs[:4]:
                g = s20.verify_range(lo, hi, 22, form=cz.FORM_PERSISTENT)
                assert_chunks_equal(g.chunks, s20.verify_range(lo, hi, 22).chunks)
        except Exception as e:  # reported below
            errors.append(e)

    th = threading.Thread(target=direct)
    th.start()
    for srv, ts, rs in ((a, ta, reqs[:8]), (b, tb, reqs[8:])):
        for t, (lo, hi) in zip(ts, rs):
            _, chunks, _ = srv.wait(t, 60000)
            assert_chunks_equal(chunks, s20.verify_range(lo, hi, 22).chunks)
    th.join()
    assert not errors, errors
    a.close()
    b.close()


def test_server_tunes_in_idle_time_only():
    """A launch whose (word width, magnitude) bucket has no tuned shape runs the default
    shape and queues the bucket's sweep, which the server runs while idle, one timing launch
    at a time: results stay those of the range alone, and a request that arrives during the
    sweep waits at most for one timing launch (a few ms), not for the whole sweep (some 80
    launches)."""
    import collatz_amd as cz
    sv = cz.Sieve(20)  # fresh: nothing tuned yet
    srv = cz.Server(sv, chunk_log2=24)
    # About 7.8e6 candidates: above the size the tuner tunes for (2^22).
    big = ((1 << 50) + 7, (1 << 50) + (1 << 29))
    total, chunks, _ = srv.wait(srv.submit(*big), 60000)
    assert_chunks_equal(chunks, sv.verify_range(*big, 24).chunks, "big request")
    lat = []
    for i in range(12):
        time.sleep(0.02)  # idle: the queued sweep advances
        lo = (1 << 44) + i * (1 << 22)
        total, chunks, (ts, ta, td) = srv.wait(srv.submit(lo, lo + (1 << 22)), 60000)
        assert_chunks_equal(chunks, sv.verify_range(lo, lo + (1 << 22), 24).chunks,
                            f"request {i}")
        lat.append(td - ts)
    # Once more in the same bucket, after the sweep has had time to finish.
    time.sleep(1.0)
    total, chunks, _ = srv.wait(srv.submit(*big), 60000)
    assert_chunks_equal(chunks, sv.verify_range(*big, 24).chunks, "big request, tuned")
    srv.close()
    assert max(lat) < 150.0, lat
    sv.close()
