"""The served workload itself: the k = 32 sieve resident on the device and the first 400
requests of the request stream of seed 1 (widths log-uniform in [2^16, 2^36]; 80% of starts
in [2^40, 2^62), 20% in [2^62, 2^64), so ends reach past 2^64), submitted in bursts while
launches run.  Every result equals the same range verified alone by the plain kernel, and a
seeded sample of 24 (at least 6 from [2^62, 2^64), and the widest request of the prefix)
equals the oracle."""
import random
import threading
import time

import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu

N_REQ = 400
CHUNK = 28


@pytest.fixture(scope="module")
def s32():
    import collatz_amd as cz
    return cz.Sieve(32)


@pytest.fixture(scope="module")
def stream():
    from collatz_amd import workloads as W
    return [(lo, hi) for _, lo, hi in W.request_stream(1, 20000)[:N_REQ]]


@pytest.fixture(scope="module")
def served(s32, stream):
    import collatz_amd as cz
    srv = cz.Server(s32, chunk_log2=CHUNK)
    tickets = []

    def producer():
        # Bursts of 1..16 requests with short pauses, so that requests arrive while
        # launches of earlier ones run and join later launches.
        rng = random.Random(11)
        i = 0
        while i < len(stream):
            burst = rng.randint(1, 16)
            for lo, hi in stream[i:i + burst]:
                tickets.append(srv.submit(lo, hi))
            i += burst
            time.sleep(rng.choice([0.0, 0.0005, 0.002]))

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    out = [srv.wait(t, 120000) for t in tickets]
    srv.close()
    return out


def test_served_stream_equals_alone(s32, stream, served):
    import collatz_amd as cz
    assert len(served) == N_REQ
    for (lo, hi), (total, chunks, times) in zip(stream, served):
        alone = s32.verify_range(lo, hi, CHUNK, form=cz.FORM_PLAIN)
        assert_chunks_equal(chunks, alone.chunks, f"request [{lo:#x},{hi:#x})")
        assert_total_equal(total, alone.total, f"request [{lo:#x},{hi:#x})")
        assert times[0] <= times[1] <= times[2]


def test_served_stream_sample_against_oracle(stream, served, oracle):
    o32 = oracle.sieve(32)
    high = [i for i, (lo, _) in enumerate(stream) if lo >= 1 << 62]
    low = [i for i, (lo, _) in enumerate(stream) if lo < 1 << 62]
    rng = random.Random(7)
    widest = max(range(len(stream)), key=lambda i: stream[i][1] - stream[i][0])
    idx = set(rng.sample(high, 6)) | set(rng.sample([i for i in low if i != widest], 17))
    idx.add(widest)
    assert len(idx) >= 24 and sum(1 for i in idx if i in high) >= 6
    assert (stream[widest][1] - stream[widest][0]).bit_length() == 36
    for i in sorted(idx):
        lo, hi = stream[i]
        chunks, total = oracle.verify_range(o32, lo, hi, CHUNK, threads=16)
        assert_chunks_equal(served[i][1], chunks, f"request {i} [{lo:#x},{hi:#x})")
        assert_total_equal(served[i][0], total, f"request {i}")
