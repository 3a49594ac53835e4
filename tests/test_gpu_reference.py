"""Reference mode on the GPU (collatz_reference_run): the reference program for a case line
-- search1 leaves at depth N - A, BV_i / mod 9 Step 2, schedule and complete hash -- against
the golden vectors written by the reference's own code, the oracle, and the paper."""
import hashlib
import os

import numpy as np
import pytest

from conftest import read_cases

pytestmark = pytest.mark.gpu

ROOT_NODE = (3, 2, 2, 9, 1, 0)
HERE = os.path.dirname(os.path.abspath(__file__))


def test_tables_opencl_match_reference(golden):
    import collatz_amd as cz
    bv, m9 = cz.reference_tables(45, 10, semantics=cz.SEM_OPENCL)
    for i in range(16):
        full = np.concatenate([bv[i], bv[i][:32]])
        assert hashlib.sha256(full.tobytes()).hexdigest() == golden["bv_sha256"][i], i
    assert m9.tolist() == golden["mod9"]


def test_tables_cuda_and_rust_match_oracle(oracle):
    import collatz_amd as cz
    bv, _ = cz.reference_tables(45, 10, semantics=cz.SEM_CUDA)
    o = oracle.precompute_bv(oracle.knobs(45, 10, semantics=oracle.SEM_CXX))
    assert np.array_equal(bv, o.view(np.uint32).reshape(16, -1))
    bv, _ = cz.reference_tables(72, 6, num_bv=8, semantics=cz.SEM_RUST)
    pop = [int(np.unpackbits(bv[i].view(np.uint8)).sum()) for i in range(8)]
    assert pop == [333343, 1590275, 3837929, 6750160, 9809263, 12495505, 14479396, 15712298]


def test_case_lines_match_reference(golden):
    """collatz_openclv2.cpp's complete hash, launch count and bins, per case line."""
    import collatz_amd as cz
    for c in golden["cases"]:
        r = cz.reference_run(c["case"], 45, 10, semantics=cz.SEM_OPENCL)
        assert r.complete_hash == c["hash"], c["case"]
        assert r.launches == c["launches"]
        assert list(r.bin_leaves)[:16] == c["bins"]
        assert r.leaves == c["leaves"] + r.bin_leaves[16]


def test_case_lines_cuda_semantics_match_oracle(oracle, bv_cuda_n45):
    import collatz_amd as cz
    kn, bv, m9 = bv_cuda_n45
    for t in read_cases("cases_big_sample.txt")[:60]:
        r = cz.reference_run(t, 45, 10, semantics=cz.SEM_CUDA)
        o = oracle.case_run(kn, bv, m9, oracle.Node.make(*t))
        assert (r.complete_hash, r.launches, r.candidates, r.leaves) == (
            o.complete_hash, o.launches, o.candidates, o.leaves), t


def test_calls_from_threads_and_reused_buffers():
    """Case lines of different sizes one after another (the device's kept buffers grow and
    are reused), then the same lines from eight threads at once (one call holds the kept
    buffers, the others take their own): every result equals the first pass's (which
    test_case_lines_cuda_semantics_match_oracle and the deep-line tests pin to the oracle)."""
    import threading
    import collatz_amd as cz
    lines = read_cases("cases_big_sample.txt")[:8]
    # a small line after the largest, a deep (128-bit leaf) line between two 64-bit ones
    runs = [(t, 45) for t in lines[:4]] + [(lines[4], 52), (lines[5], 45), (lines[6], 52),
                                           (lines[7], 45)]
    key = lambda r: (r.complete_hash, r.candidates, r.leaves, r.checks, r.launches,
                     r.overflow_count)
    want = [key(cz.reference_run(t, n, 10, semantics=cz.SEM_CUDA)) for t, n in runs]
    again = [key(cz.reference_run(t, n, 10, semantics=cz.SEM_CUDA)) for t, n in reversed(runs)]
    assert again[::-1] == want
    got = [None] * len(runs)
    errs = []

    def run(i):
        try:
            t, n = runs[i]
            got[i] = key(cz.reference_run(t, n, 10, semantics=cz.SEM_CUDA))
        except Exception as e:  # reported below, on the test's thread
            errs.append((i, e))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(runs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    assert not errs, errs
    assert got == want


@pytest.mark.parametrize("N,expected", [(35, 31109345), (36, 56414851), (37, 106551583),
                                        (38, 205942265), (39, 390055737), (40, 756583624),
                                        (41, 1408862059), (42, 2658801769), (43, 5143671444),
                                        (44, 9392592637), (45, 17807287297)])
def test_fig51_counts(N, expected):
    """paper p.10 Fig 5.1: integers to iterate for N = 35..40 (main.rs semantics)."""
    import collatz_amd as cz
    r = cz.reference_run(ROOT_NODE, N, 6, num_bv=8, semantics=cz.SEM_RUST)
    assert r.candidates == expected


def test_cases_tiny_partition_xor():
    """The hash is an XOR over candidates and the case lines partition the tree, so the
    lines' hashes XOR to the hash of the whole run (N = 40)."""
    import collatz_amd as cz
    full = cz.reference_run(ROOT_NODE, 40, 10, semantics=cz.SEM_CUDA)
    h, cand = 0, 0
    for t in read_cases("cases_tiny.txt"):
        r = cz.reference_run(t, 40, 10, semantics=cz.SEM_CUDA)
        h ^= r.complete_hash
        cand += r.candidates
    assert h == full.complete_hash and cand == full.candidates


def test_full_run_n45_matches_reference():
    """The whole run below 2^45 (root 3 2 2 9 1 0): its complete hash equals the XOR of the
    reference's own hashes over all 278,699 lines of cases_big."""
    import json
    path = os.path.join(HERE, "golden", "ref_full45.json")
    if not os.path.exists(path):
        pytest.skip("ref_full45.json not generated")
    g = json.load(open(path))
    assert not g["lines_too_large"]
    import collatz_amd as cz
    r = cz.reference_run(ROOT_NODE, 45, 10, semantics=cz.SEM_OPENCL)
    assert r.complete_hash == g["complete_hash"]
    assert list(r.bin_leaves)[:16] == g["bin_leaves"]


def test_overflow_reports_match_oracle(oracle, bv_opencl_n45):
    """A case line run with CUTOFFPOWER 64: overflow reports (start n) as the oracle counts."""
    import collatz_amd as cz
    kn, bv, m9 = bv_opencl_n45
    kn64 = oracle.knobs(45, 10, semantics=oracle.SEM_OPENCL, cutoff=64)
    t = read_cases("cases_big_sample.txt")[3]
    r = cz.reference_run(t, 45, 10, semantics=cz.SEM_OPENCL, cutoff_log2=64, overflow_cap=1 << 16)
    o = oracle.case_run(kn64, bv, m9, oracle.Node.make(*t))
    assert r.overflow_count == o.overflow_count > 0
    assert r.complete_hash == o.complete_hash
