"""The versioned per-chunk result file ("collatz-chunks 1"): round trips and refusals."""
import numpy as np
import pytest

from collatz_amd import CHUNK_DTYPE
from collatz_amd import chunkfile as CF


def records(n, seed=0):
    rng = np.random.default_rng(seed)
    ch = np.zeros(n, CHUNK_DTYPE)
    for f in ("chunk_index", "n_checked", "count", "stop_sum", "stop_xor", "overflow_count"):
        ch[f] = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    ch["max_exc"] = rng.integers(0, 1 << 63, (n, 3), dtype=np.uint64)
    ch["max_start"] = rng.integers(0, 1 << 63, (n, 2), dtype=np.uint64)
    ch["hash"] = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    ch["flags"] = rng.integers(0, 4, n)
    return ch


def test_round_trip():
    ch = records(50)
    b = CF.format_chunks(k=32, chunk_log2=24, cutoff_log2=100, lo=3, hi=(1 << 100) + 7,
                         chunks=ch, total=ch[0], overflow_starts=[1 << 90, 5],
                         lookahead_bits=16, lookahead_semantics=2)
    r = CF.parse_chunks(b)
    assert r["chunks"].tobytes() == ch.tobytes()
    assert r["total"].tobytes() == ch[0].tobytes()
    assert (r["k"], r["chunk_log2"], r["cutoff_log2"], r["lo"], r["hi"]) == \
        (32, 24, 100, 3, (1 << 100) + 7)
    assert (r["lookahead_bits"], r["lookahead_semantics"]) == (16, 2)
    assert r["overflow_starts"] == [5, 1 << 90]
    assert CF.format_chunks(k=32, chunk_log2=24, cutoff_log2=100, lo=3, hi=(1 << 100) + 7,
                            chunks=r["chunks"], total=r["total"],
                            overflow_starts=r["overflow_starts"], lookahead_bits=16,
                            lookahead_semantics=2) == b


def test_empty_and_refusals():
    z = np.zeros(0, CHUNK_DTYPE)
    b = CF.format_chunks(k=20, chunk_log2=24, cutoff_log2=187, lo=1, hi=2, chunks=z,
                         total=np.zeros((), CHUNK_DTYPE))
    assert len(CF.parse_chunks(b)["chunks"]) == 0
    with pytest.raises(ValueError):
        CF.parse_chunks(b.replace(b"collatz-chunks 1", b"collatz-chunks 2"))
    flipped = bytearray(b)
    flipped[len(b) // 2] ^= 1
    with pytest.raises(ValueError):
        CF.parse_chunks(bytes(flipped))
    with pytest.raises(ValueError):
        CF.parse_chunks(b[:-3])


def test_field_order_is_the_header_struct():
    """The file's field list is collatz_chunk's member order (include/collatz.h)."""
    assert CF.FIELDS == CHUNK_DTYPE.names
