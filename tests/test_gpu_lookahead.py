"""Sieves with the Step 2 look-ahead (lookahead_bits): every form against the oracle, bit for
bit, and the whole tree below 2^45 against the reference's own complete hash."""
import json
import os
import random

import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT_NODE = (3, 2, 2, 9, 1, 0)


@pytest.fixture(scope="module")
def o24(oracle):
    return oracle.sieve(24)


@pytest.fixture(scope="module", params=[0, 2], ids=["cuda", "opencl"])
def la24(request):
    import collatz_amd as cz
    return cz.Sieve(24, lookahead_bits=24, lookahead_semantics=request.param)


def ranges():
    rng = random.Random(11)
    out = [(1, 1 << 30), ((1 << 40) + 77, (1 << 40) + (1 << 29))]
    for bits in (50, 63, 66, 100):       # one, two and three-limb word widths
        lo = rng.getrandbits(bits) | (1 << (bits - 1))
        out.append((lo, lo + rng.randrange(1 << 27, 1 << 28)))
    return out


def test_forms_match_oracle(la24, o24, oracle):
    sem = la24.lookahead_semantics
    for lo, hi in ranges():
        chunks, total = oracle.verify_range(o24, lo, hi, 24, threads=16, lookahead_bits=24,
                                            lookahead_semantics=sem)
        for form in (0, 1):
            g = la24.verify_range(lo, hi, 24, form=form)
            assert_chunks_equal(g.chunks, chunks, f"[{lo:#x},{hi:#x}) form {form} sem {sem}")
            assert_total_equal(g.total, total, "total")


def test_small_lookahead_bits(o24, oracle):
    import collatz_amd as cz
    for B in (8, 13, 20):
        sv = cz.Sieve(24, lookahead_bits=B)
        assert sv.info().lookahead_bits == B
        lo = (1 << 52) + 999
        chunks, total = oracle.verify_range(o24, lo, lo + (1 << 28), 24, threads=16,
                                            lookahead_bits=B)
        for form in (0, 1):
            g = sv.verify_range(lo, lo + (1 << 28), 24, form=form)
            assert_chunks_equal(g.chunks, chunks, f"B={B} form {form}")
        sv.close()


def test_batch_and_server_match_range(la24):
    import collatz_amd as cz
    rng = random.Random(3)
    ivs = []
    for _ in range(40):
        lo = rng.getrandbits(rng.choice([40, 60, 64, 70])) + 1
        ivs.append((lo, lo + rng.randrange(1 << 20, 1 << 27)))
    want = [la24.verify_range(lo, hi, 22) for lo, hi in ivs]
    for graph in (False, True):
        got = la24.verify_batch(ivs, chunk_log2=22, use_graph=graph)
        for (lo, hi), g, w in zip(ivs, got, want):
            assert_chunks_equal(g.chunks, w.chunks, f"batch [{lo:#x},{hi:#x})")
    srv = cz.Server(la24, chunk_log2=22)
    tickets = [srv.submit(lo, hi) for lo, hi in ivs]
    for t, w in zip(tickets, want):
        total, chunks, _ = srv.wait(t)
        assert_chunks_equal(chunks, w.chunks, "server")
    srv.close()


def test_full_tree_n45_matches_reference():
    """[1, 2^45) with k = 35 and the 24-step look-ahead (OpenCL compares) is the reference's
    whole N = 45 run (A = 10, CUTOFFPOWER 100): its hash is the XOR of the reference's own
    complete hashes over all 278,699 lines of cases_big, and its count is the reference-mode
    run's."""
    import collatz_amd as cz
    g = json.load(open(os.path.join(HERE, "golden", "ref_full45.json")))
    sv = cz.Sieve(35, lookahead_bits=24, lookahead_semantics=cz.SEM_OPENCL)
    ref = cz.reference_run(ROOT_NODE, 45, 10, semantics=cz.SEM_OPENCL)
    for form in (1, 0):
        out = sv.verify_range(1, 1 << 45, 40, cutoff_log2=100, form=form)
        assert int(out.total["hash"]) == g["complete_hash"] == ref.complete_hash, form
        assert int(out.total["count"]) == ref.candidates
        assert out.checks == ref.checks
        assert int(out.total["n_checked"]) == (1 << 45) - 1


def test_full_tree_n40_cuda_matches_reference_mode():
    """The same identity under the CUDA program's compares (collatz_cudav7.cu:449,461), at
    N = 40: range form with the look-ahead against reference mode."""
    import collatz_amd as cz
    sv = cz.Sieve(30, lookahead_bits=24, lookahead_semantics=cz.SEM_CUDA)
    ref = cz.reference_run(ROOT_NODE, 40, 10, semantics=cz.SEM_CUDA)
    out = sv.verify_range(1, 1 << 40, 36, cutoff_log2=100, form=1)
    assert int(out.total["hash"]) == ref.complete_hash
    assert int(out.total["count"]) == ref.candidates
    # The same candidates run the same checks, counted per lane by the range kernels and
    # per wave by reference mode.
    assert out.checks == ref.checks
