"""Spot checks (the reference re-runs a random launch's candidates on the CPU and compares
its hash, collatz_cudav7.cu:619-635): sampled tiles of reference mode re-run by the list
kernel, sampled chunk records of the batched form re-run by the plain kernel.  Every check
here must pass, and a run with spot checks must report exactly what the same run without
them reports."""
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT_NODE = (3, 2, 2, 9, 1, 0)
CLI = os.path.join(os.path.dirname(HERE), "collatz_amd", "collatz_run")


def same_result(a, b):
    return (a.complete_hash, a.candidates, a.numcases, a.overflow_count, a.leaves,
            a.launches, list(a.bin_leaves)) == (b.complete_hash, b.candidates, b.numcases,
                                                b.overflow_count, b.leaves, b.launches,
                                                list(b.bin_leaves))


# (N, A, num_bv, semantics, cutoff): the narrow pass with the reference schedule, OpenCL
# compares, a cutoff below 2^96 (the four-word pass with overflow resets), main.rs's blocks
# at 2^120 and with its overflow stops at 2^40.
CONFIGS = [(40, 10, 16, 0, 100), (38, 10, 16, 2, 100), (36, 10, 16, 0, 64),
           (36, 6, 8, 1, 120), (33, 6, 8, 1, 40)]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_reference_spot_checks_pass(cfg):
    import collatz_amd as cz
    N, A, num_bv, sem, cutoff = cfg
    print("spot", cfg, flush=True)
    plain = cz.reference_run(ROOT_NODE, N, A, num_bv=num_bv, semantics=sem, cutoff_log2=cutoff,
                             overflow_cap=1 << 22)
    assert plain.spot_checks == 0
    r = cz.reference_run(ROOT_NODE, N, A, num_bv=num_bv, semantics=sem, cutoff_log2=cutoff,
                         overflow_cap=1 << 22, spot_check_freq=3)
    assert r.spot_checks > 0 and r.spot_check_failures == 0
    assert same_result(r, plain), cfg


def test_reference_spot_checks_deep_line():
    """A case line of the n < 2^60 run (128-bit leaf numbers), every tile sampled up to the
    per-launch limit (2^23 candidates: 32 tiles at A = 10), against its fixture."""
    import collatz_amd as cz
    with open(os.path.join(HERE, "golden", "ref_deep.json")) as f:
        fx = [x for x in json.load(f)["runs"] if x["N"] == 60 and x["cutoff"] == 100][0]
    r = cz.reference_run(tuple(fx["line"]), 60, 10, semantics=fx["semantics"],
                         cutoff_log2=100, spot_check_freq=1)
    assert r.spot_checks == 32 and r.spot_check_failures == 0
    assert (r.complete_hash, r.candidates, r.leaves) == (fx["complete_hash"], fx["candidates"],
                                                         fx["leaves"])


def test_batch_spot_checks_pass():
    """Mixed intervals (two and three limbs, a low cutoff with overflows, widths from one
    chunk to many): 24 sampled chunk records re-run by the plain kernel."""
    import collatz_amd as cz
    sv = cz.Sieve(20)
    ivs = [((1 << 40) + 12345, (1 << 40) + (1 << 26)), ((1 << 62) + 5, (1 << 62) + (1 << 23)),
           ((1 << 100) + 7, (1 << 100) + (1 << 22)), (3, 1 << 24),
           ((1 << 127) - (1 << 22), (1 << 127) + (1 << 20))]
    for cutoff in (187, 80):
        want = sv.verify_batch(ivs, chunk_log2=20, cutoff_log2=cutoff, use_graph=True)
        got = sv.verify_batch(ivs, chunk_log2=20, cutoff_log2=cutoff, use_graph=True,
                              spot_checks=24)
        assert sum(o.spot_checks for o in got) == 24
        assert all(o.spot_checks == 0 for o in want)
        for a, b in zip(want, got):
            assert a.total.tobytes() == b.total.tobytes()
            assert a.chunks.tobytes() == b.chunks.tobytes()
    # More checks asked for than records stored: every record is checked once.
    small = [((1 << 50) + 1, (1 << 50) + (1 << 21))]
    got = sv.verify_batch(small, chunk_log2=20, spot_checks=100)
    assert got[0].spot_checks == len(got[0].chunks) == 2


def test_cli_reports_spot_checks():
    """The CLI's summary line counts the spot checks it ran (--spot-freq 2: every second
    tile), and the launch count and hash are those of the run without them."""
    line = "3 2 2 9 1 0"
    outs = []
    for freq in ("0", "2"):
        p = subprocess.run([CLI, "--N", "36", "--A", "10", "--spot-freq", freq], input=line,
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        m = re.search(r"Number of kernel launches = (\d+)\. Spot checks = (\d+)\. "
                      r"Complete hash = ([0-9A-F]+)\.", p.stderr)
        assert m, p.stderr
        assert "Spot check failed" not in p.stderr
        outs.append((int(m.group(1)), int(m.group(2)), m.group(3)))
    assert outs[0][1] == 0 and outs[1][1] > 0
    assert (outs[0][0], outs[0][2]) == (outs[1][0], outs[1][2])
