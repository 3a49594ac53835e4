"""collatz_verify_range_noalloc (the plain form on the library's fixed device array; the call
allocates no device memory) against collatz_verify_range's plain form, bit for bit, through
the C ABI: two calls in a row, the second with a smaller input than the first (stale
contents of the array from the first call would show), then the limits."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu


def run(fn, sv, lo, hi, chunk_log2, cutoff, cap):
    import collatz_amd as cz
    n = ((hi - 1) >> chunk_log2) - (lo >> chunk_log2) + 1
    chunks = np.zeros(n, cz.CHUNK_DTYPE)
    ovf = np.zeros(2 * cap, np.uint64)
    res = cz.RangeResult()
    res.chunks = chunks.ctypes.data_as(C.POINTER(cz.Chunk))
    res.chunks_cap = n
    res.overflow_starts = ovf.ctypes.data_as(C.POINTER(C.c_uint64))
    res.overflow_cap = cap
    opts = cz.VerifyOpts(chunk_log2, 2, cutoff, cz.FORM_PLAIN)
    rc = fn(sv.handle, cz.u128_pair(lo), cz.u128_pair(hi), C.byref(opts), None, C.byref(res))
    total = np.frombuffer(bytearray(res.total), cz.CHUNK_DTYPE)[0]
    k = min(res.overflow_written, cap)
    starts = sorted(int(ovf[2 * i]) | (int(ovf[2 * i + 1]) << 64) for i in range(k))
    return rc, res, total, chunks, starts


CASES = [
    # 2^26 integers just below 2^68 (96-bit words; some candidates climb past 2^86 and are
    # rerun on three limbs), 256 records.
    ((1 << 67) + 3, (1 << 67) + (1 << 26), 18, 187),
    # A smaller range with overflows at a cutoff of 2^64 (four words, overflow path).
    ((1 << 60) + 5, (1 << 60) + (1 << 20), 16, 64),
]


def test_noalloc_equals_plain_two_calls():
    import collatz_amd as cz
    L = cz.load()
    sv = cz.Sieve(20)
    seen = []
    for lo, hi, c, cut in CASES:
        print(f"noalloc case [{lo:#x}, {hi:#x}) chunk 2^{c} cutoff {cut}", flush=True)
        rc0, want, wtot, wchunks, wstarts = run(L.collatz_verify_range, sv, lo, hi, c, cut,
                                                1 << 16)
        rc1, got, gtot, gchunks, gstarts = run(L.collatz_verify_range_noalloc, sv, lo, hi, c,
                                               cut, 1 << 16)
        assert rc0 == cz.OK, L.collatz_last_error()
        assert_chunks_equal(gchunks, wchunks, "noalloc chunks")
        assert_total_equal(gtot, wtot, "noalloc total")
        assert int(wtot["count"]) > 0
        assert got.checks == want.checks and got.escalated == want.escalated
        assert got.overflow_written == want.overflow_written == int(wtot["overflow_count"])
        if want.overflow_written <= 4096:
            assert rc1 == cz.OK, L.collatz_last_error()
            assert gstarts == wstarts
        else:
            assert rc1 == cz.ECAPACITY
        seen.append((want.escalated, want.overflow_written))
    # The first case took the three-limb pass, the second the overflow path.
    assert seen[0][0] > 0 and seen[1][1] > 0


def test_noalloc_limits():
    """Ranges past the limits fail with EINVAL before any kernel runs; the persistent form
    is refused; a call after them still matches the plain form."""
    import collatz_amd as cz
    L = cz.load()
    sv = cz.Sieve(16)
    lo = 1 << 40
    rc, *_ = run(L.collatz_verify_range_noalloc, sv, lo, lo + (4097 << 16), 16, 187, 16)
    assert rc == cz.EINVAL
    res = cz.RangeResult()
    opts = cz.VerifyOpts(24, 2, 187, cz.FORM_PERSISTENT)
    assert L.collatz_verify_range_noalloc(sv.handle, cz.u128_pair(lo), cz.u128_pair(lo + 99),
                                          C.byref(opts), None, C.byref(res)) == cz.EINVAL
    a = run(L.collatz_verify_range_noalloc, sv, lo + 7, lo + (1 << 22), 20, 187, 16)
    b = run(L.collatz_verify_range, sv, lo + 7, lo + (1 << 22), 20, 187, 16)
    assert a[0] == b[0] == cz.OK
    assert_chunks_equal(a[3], b[3], "after the limits")


def test_noalloc_at_the_chunk_limit(oracle):
    """4096 chunk records of 2^16 at k = 16 (24,576 rows, one per chunk and residue class):
    the most rows a call of the limit holds, whose shares the plain form sizes to fit the
    array's share records; equal to the plain form, and to the oracle on its first and last
    chunks."""
    import collatz_amd as cz
    L = cz.load()
    sv = cz.Sieve(16)
    lo = (1 << 40) + 1
    hi = (1 << 40) + (4096 << 16)
    print(f"noalloc case [{lo:#x}, {hi:#x}) chunk 2^16, 4096 records", flush=True)
    rc1, got, gtot, gchunks, _ = run(L.collatz_verify_range_noalloc, sv, lo, hi, 16, 187, 16)
    rc0, want, wtot, wchunks, _ = run(L.collatz_verify_range, sv, lo, hi, 16, 187, 16)
    assert rc0 == rc1 == cz.OK, L.collatz_last_error()
    assert len(wchunks) == 4096
    assert_chunks_equal(gchunks, wchunks, "noalloc at the chunk limit")
    assert_total_equal(gtot, wtot, "noalloc total at the chunk limit")
    o16 = oracle.sieve(16)
    for a, b, idx in ((lo, (1 << 40) + (1 << 16), 0), ((1 << 40) + (4095 << 16), hi, 4095)):
        ochunks, _ = oracle.verify_range(o16, a, b, 16)
        assert_chunks_equal(gchunks[idx:idx + 1], ochunks, f"chunk {idx} against the oracle")
