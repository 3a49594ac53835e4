"""Reference mode with leaves deeper than 40 bits (their numbers take 128 bits: case lines
of the reference's own n < 2^55 and n < 2^60 runs, paper p.3-4) and main.rs's own Step 3
(its search3 blocks, numcases and overflow test, main.rs:321-341), against the oracle: the
fixtures of tests/golden/make_ref_deep.py, small whole trees run live, and the paper's
published count for main.rs's configuration at N = 50."""
import hashlib
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT_NODE = (3, 2, 2, 9, 1, 0)


def fixtures():
    with open(os.path.join(HERE, "golden", "ref_deep.json")) as f:
        return json.load(f)["runs"]


def starts_of(r):
    k = min(int(r.overflow_written), int(r.overflow_cap))
    return [int(r._ovf[2 * i]) | (int(r._ovf[2 * i + 1]) << 64) for i in range(k)]


@pytest.mark.parametrize("i", range(8))
def test_deep_case_lines_equal_oracle(i):
    import collatz_amd as cz
    f = fixtures()[i]
    tag = f"line {f['line']} N={f['N']} A={f['A']} sem={f['semantics']} cutoff={f['cutoff']}"
    print(tag, flush=True)
    r = cz.reference_run(tuple(f["line"]), f["N"], f["A"], f["B"], num_bv=f["num_bv"],
                         semantics=f["semantics"], cutoff_log2=f["cutoff"],
                         overflow_cap=1 << 22)
    assert r.leaves == f["leaves"], tag
    assert list(r.bin_leaves) == f["bin_leaves"], tag
    assert r.launches == f["launches"], tag
    assert r.candidates == f["candidates"], tag
    assert r.complete_hash == f["complete_hash"], tag
    assert r.numcases == f["numcases"], tag
    assert r.overflow_count == f["overflow_count"], tag
    digest = hashlib.sha256("\n".join(str(x) for x in sorted(starts_of(r))).encode()).hexdigest()
    assert digest == f["overflow_starts_sha256"], tag


@pytest.mark.parametrize("N,cutoff", [(30, 120), (33, 40), (35, 64)])
def test_rust_whole_tree_equals_oracle(oracle, N, cutoff):
    """main.rs semantics over the whole tree below 2^N: numcases, candidates, leaves and the
    overflow starts (cutoff 2^40 and 2^64 make them happen) equal the oracle's."""
    import collatz_amd as cz
    kn = oracle.knobs(N, 6, num_bv=8, semantics=oracle.SEM_RUST, cutoff=cutoff)
    want, starts = oracle.case_run_starts(kn, oracle.precompute_bv(kn), oracle.mod9_rows(N - 6, 6),
                                          oracle.Node.make(*ROOT_NODE), cap=1 << 22, threads=16)
    r = cz.reference_run(ROOT_NODE, N, 6, num_bv=8, semantics=cz.SEM_RUST, cutoff_log2=cutoff,
                         overflow_cap=1 << 22)
    assert r.complete_hash == 0
    assert (r.candidates, r.leaves, r.numcases) == (want.candidates, want.leaves, want.numcases)
    assert r.overflow_count == want.overflow_count
    assert sorted(starts_of(r)) == sorted(starts)
    if cutoff == 40:
        assert want.overflow_count > 0


def test_paper_n50_main_rs_count():
    """paper p.3 section 1.1: for all n < 2^50 main.rs's method iterates 434,413,363,553
    integers (A = 6, eight bit vectors, its search2 filters; the leaves at depth 44 take
    128 bits)."""
    import collatz_amd as cz
    r = cz.reference_run(ROOT_NODE, 50, 6, num_bv=8, semantics=cz.SEM_RUST)
    assert r.candidates == 434413363553
    assert r.numcases > r.candidates and r.overflow_count == 0
