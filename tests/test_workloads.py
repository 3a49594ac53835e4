"""The benchmark's synthetic workloads: the frozen generator and the committed fixtures."""
import json
import os

from collatz_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_splitmix64_known_answers():
    # Published first outputs of SplitMix64 from seed 0.
    r = W.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                           0x06C45D188009454F]


def test_batch_intervals_fixture():
    ivs = W.batch_intervals(0, 4096, 32)
    with open(os.path.join(GOLD, "batch_intervals_seed0.txt")) as f:
        want = [tuple(int(x) for x in l.split()) for l in f if l.strip()]
    assert ivs == want
    for lo, hi in ivs:
        assert (1 << 48) <= lo < (1 << 60)
        w = hi - lo
        assert w % (1 << 32) == 0 and (1 << 32) <= w <= (1 << 34)


def test_request_stream_fixture():
    reqs = W.request_stream(1, 20000)
    with open(os.path.join(GOLD, "request_stream_seed1.json")) as f:
        want = json.load(f)
    assert W.digest(reqs) == want["sha256"]
    assert [list(r) for r in reqs[:64]] == want["first"]
    hi_frac = sum(lo >= (1 << 62) for _, lo, _ in reqs) / len(reqs)
    assert 0.19 < hi_frac < 0.21
    for _, lo, hi in reqs:
        assert (1 << 16) <= hi - lo <= (1 << 36) and (1 << 40) <= lo < (1 << 64)
