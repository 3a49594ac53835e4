"""GPU sieve (search1 per residue + bit array + compaction) against the oracle."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [8, 12, 16, 20, 24, 28])
def test_sieve_equals_oracle(oracle, k):
    import collatz_amd as cz
    sv = cz.Sieve(k)
    r, num, f = sv.survivors()
    o = oracle.sieve(k)
    assert np.array_equal(r, o.r)
    assert np.array_equal(num, o.number)
    assert np.array_equal(f, o.f)
    inf = sv.info()
    assert inf.survivors == len(o.r)
    assert inf.bitarray_bits == (1 << (k - 2) if k >= 9 else 0)


def test_sieve_k32(oracle):
    """k = 32: 25,005,944 survivors (the 2^32 residue classes), identical lists."""
    import collatz_amd as cz
    sv = cz.Sieve(32)
    r, num, f = sv.survivors()
    assert len(r) == 25005944
    o = oracle.sieve(32)
    assert np.array_equal(r, o.r)
    assert np.array_equal(num, o.number)
    assert np.array_equal(f, o.f)
    # number = T^32(r) + 1 by naive iteration, on a sample.
    rng = random.Random(0)
    for i in rng.sample(range(len(r)), 200):
        x, odd = int(r[i]), 0
        for _ in range(32):
            if x % 2:
                odd += 1
                x = (3 * x + 1) // 2
            else:
                x //= 2
        assert int(num[i]) == x + 1 and int(f[i]) == odd


def test_sieve_rejects_bad_k():
    import collatz_amd as cz
    for k in (0, 7, 39, 64):
        with pytest.raises(cz.CollatzError):
            cz.Sieve(k)
