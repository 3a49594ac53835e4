"""Sieve depths above 32 bits: residues of more than 32 bits in the packed survivor records,
64-bit residue windows, and the large-sieve workload (k = 37 on [2^60, 2^60 + 2^44) in
2^30-wide chunks) against oracle fixtures."""
import json
import os

import numpy as np
import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_sieve_k34_equals_oracle(oracle):
    """k = 34: every survivor (r up to 2^34, so the records carry r's high bits)."""
    import collatz_amd as cz
    sv = cz.Sieve(34)
    o = oracle.sieve(34)
    r, num, f = sv.survivors()
    assert len(r) == len(o.r)
    assert np.array_equal(r, o.r)
    assert np.array_equal(num, o.number)
    assert np.array_equal(f, o.f)
    # Ranges whose residue windows straddle 2^32 and 2^33 inside one q block.
    o_ = oracle
    for lo, hi, C in [((5 << 34) + (1 << 32) - (1 << 22), (5 << 34) + (1 << 32) + (1 << 22), 20),
                      ((1 << 60) + (3 << 32) - 12345, (1 << 60) + (3 << 32) + (1 << 23), 24)]:
        chunks, total = o_.verify_range(o, lo, hi, C, threads=16)
        for form in (cz.FORM_PLAIN, cz.FORM_PERSISTENT):
            g = sv.verify_range(lo, hi, C, form=form)
            assert_chunks_equal(g.chunks, chunks, f"k=34 [{lo:#x},{hi:#x}) form {form}")
            assert_total_equal(g.total, total)


def _fixture_chunks(sl):
    import collatz_amd as cz
    rows = sl["chunks"]
    a = np.zeros(len(rows), cz.CHUNK_DTYPE)
    for i, c in enumerate(rows):
        for k, v in c.items():
            a[i][k] = v
    return a


def test_k37_sieve_and_workload_slices():
    """k = 37: |S_37| and three slices of the large-sieve range, every form, against the
    oracle's records (tests/golden/make_k37_fixtures.py)."""
    path = os.path.join(HERE, "golden", "k37_fixtures.json")
    fx = json.load(open(path))
    import collatz_amd as cz
    sv = cz.Sieve(37)
    assert sv.info().survivors == fx["survivors"]
    for sl in fx["slices"]:
        want = _fixture_chunks(sl)
        lo, hi, C = sl["lo"], sl["hi"], sl["chunk_log2"]
        for form in (cz.FORM_PLAIN, cz.FORM_PERSISTENT):
            g = sv.verify_range(lo, hi, C, form=form)
            assert_chunks_equal(g.chunks, want, f"k=37 [{lo:#x},{hi:#x}) form {form}")
        b = sv.verify_batch([(lo, hi)], chunk_log2=C, use_graph=True)[0]
        assert_chunks_equal(b.chunks, want, f"k=37 [{lo:#x},{hi:#x}) batch")
    # The whole large-sieve range: forms agree and records add up.
    lo, hi = 1 << 60, (1 << 60) + (1 << 44)
    a = sv.verify_batch([(lo, hi)], chunk_log2=30, use_graph=True)[0]
    assert len(a.chunks) == 1 << 14
    assert int(a.chunks["n_checked"].sum()) == 1 << 44
    assert int(a.chunks["count"].sum()) == int(a.total["count"])
    first = _fixture_chunks(fx["slices"][0])
    assert_chunks_equal(a.chunks[:1], first, "first chunk of the whole run")


@pytest.mark.parametrize("k", [35, 38])
def test_range_sieve_counts_equal_reference_mode_leaves(k):
    """The range sieve of depth k and reference mode at N = k + 10 walk the same search1
    tree from the root "3 2 2 9 1 0" by different routes (one bit array over all residues;
    frontier roots expanded one after another): |S_k| equals the reference-mode leaf count.
    At k = 35 the leaves per bit vector are the reference's own (ref_full45.json)."""
    import json
    import os
    import collatz_amd as cz
    sv = cz.Sieve(k)
    n = int(sv.info().survivors)
    sv.close()
    r = cz.reference_run((3, 2, 2, 9, 1, 0), k + 10, 10, semantics=cz.SEM_CUDA)
    assert n == r.leaves == sum(r.bin_leaves)
    if k == 35:
        here = os.path.dirname(os.path.abspath(__file__))
        g = json.load(open(os.path.join(here, "golden", "ref_full45.json")))
        assert list(r.bin_leaves)[:16] == g["bin_leaves"]
