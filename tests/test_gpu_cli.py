"""collatz_run, the command-line front end: the reference program's stdin/stderr protocol
(collatz_openclv2.cpp:279-280, collatz_cudav7.cu:181-182) and the range mode."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "collatz_amd", "collatz_run")


def run(args, stdin=""):
    if not os.path.exists(CLI):
        pytest.fail("collatz_run is not built (python -c 'import __graft_entry__ as g; g.build()')")
    return subprocess.run([CLI] + args, input=stdin, capture_output=True, text=True, timeout=120)


def test_usage_without_gpu():
    p = run(["--bogus"])
    assert p.returncode == 2 and "usage" in p.stderr


@pytest.mark.gpu
def test_case_lines_protocol(golden):
    for c in golden["cases"][:8]:
        p = run(["--N", "45", "--A", "10", "--semantics", "opencl"], " ".join(map(str, c["case"])))
        assert p.returncode == 0, p.stderr
        m = re.search(r"Number of kernel launches = (\d+) \nSpot checks = \d+\. Complete hash = ([0-9A-F]+)\.",
                      p.stderr)
        assert m, p.stderr
        assert int(m.group(1)) == c["launches"] and int(m.group(2), 16) == c["hash"], c["case"]


@pytest.mark.gpu
def test_range_mode(oracle):
    o = oracle.sieve(20)
    lo, hi = (1 << 40) + 5, (1 << 40) + (1 << 22) + 77
    chunks, total = oracle.verify_range(o, lo, hi, 20)
    p = run(["--range", str(lo), str(hi), "--k", "20", "--chunk", "20"])
    assert p.returncode == 0, p.stderr
    lines = [l for l in p.stdout.splitlines() if l.startswith("chunk ")]
    assert len(lines) == len(chunks)
    for l, c in zip(lines, chunks):
        f = l.split()
        assert int(f[1]) == int(c["chunk_index"]) and int(f[5]) == int(c["count"])
        assert int(f[7]) == int(c["stop_sum"]) and int(f[9]) == int(c["stop_xor"])
        assert int(f[11], 16) == int(c["hash"])
        assert int(f[13]) == int(c["max_start"][0]) | (int(c["max_start"][1]) << 64)
    assert f"hash {int(total['hash']):08X}" in p.stderr


@pytest.mark.gpu
def test_range_result_file(tmp_path, oracle):
    """--out writes the versioned result file: the library's records for the range (the
    oracle's, bit for bit), byte-identical to collatz_amd.chunkfile's own writer; with a low
    cutoff it lists the overflow starts, and with --lookahead it records the filter."""
    import collatz_amd as cz
    from collatz_amd import chunkfile as CF
    o = oracle.sieve(20)
    lo, hi = (1 << 52) + 11, (1 << 52) + (1 << 23)     # a few hundred values reach 2^64
    for extra, cutoff, la in (([], 187, 0), (["--cutoff", "64"], 64, 0),
                              (["--lookahead", "16"], 187, 16)):
        path = str(tmp_path / f"r{cutoff}_{la}.txt")
        p = run(["--range", str(lo), str(hi), "--k", "20", "--chunk", "20", "--out", path] + extra)
        assert p.returncode == 0, p.stderr
        r = CF.read_chunks(path)
        chunks, total = oracle.verify_range(o, lo, hi, 20, cutoff, lookahead_bits=la)
        assert r["chunks"].tobytes() == chunks.tobytes()
        assert r["total"].tobytes() == total.tobytes()
        assert r["lookahead_bits"] == la and r["cutoff_log2"] == cutoff
        sv = cz.Sieve(20, lookahead_bits=la)
        g = sv.verify_range(lo, hi, 20, cutoff_log2=cutoff)
        starts = g.overflow_starts
        if cutoff == 64:
            assert len(starts) > 0 and r["overflow_starts"] == sorted(starts)
        same = CF.format_chunks(k=20, chunk_log2=20, cutoff_log2=cutoff, lo=lo, hi=hi,
                                chunks=g.chunks, total=g.total, overflow_starts=starts,
                                lookahead_bits=la)
        assert open(path, "rb").read() == same
        sv.close()


@pytest.mark.gpu
def test_case_file_driver():
    """--cases runs every line of a case file: the per-line hashes equal the library's for
    each line, and their XOR (the lines of cases_tiny partition the tree) equals the
    whole-tree run's complete hash."""
    import collatz_amd as cz
    path = os.path.join(ROOT, "tests", "golden", "cases_tiny.txt")
    p = run(["--N", "40", "--A", "10", "--semantics", "cuda", "--cases", path])
    assert p.returncode == 0, p.stderr
    per_line = re.findall(r"^(.*): launches (\d+) hash ([0-9A-F]+)$", p.stderr, re.M)
    lines = [l for l in open(path).read().splitlines() if l.strip()]
    assert len(per_line) == len(lines)
    for (line, launches, h), want in zip(per_line[:6], lines[:6]):
        r = cz.reference_run(tuple(int(x) for x in want.split()), 40, 10, semantics=cz.SEM_CUDA)
        assert line.strip() == want.strip() and int(h, 16) == r.complete_hash
        assert int(launches) == r.launches
    m = re.search(r"XOR of complete hashes = ([0-9A-F]+)", p.stderr)
    full = cz.reference_run((3, 2, 2, 9, 1, 0), 40, 10, semantics=cz.SEM_CUDA)
    assert m and int(m.group(1), 16) == full.complete_hash


@pytest.mark.gpu
def test_rust_protocol(oracle):
    """main.rs's protocol: the first four numbers of stdin with doubleeven and numevens 0
    (main.rs:88-97), "Done with precompute" and "numcases = X" on stderr, and the start of
    every trajectory whose 16-step block reached 2^cutoff in decimal on stdout
    (main.rs:336-338); X and the starts equal the oracle's main.rs restatement."""
    N, cutoff = 32, 44
    p = run(["--N", str(N), "--semantics", "rust", "--cutoff", str(cutoff)], "3 2 2 9\n")
    assert p.returncode == 0, p.stderr
    assert p.stderr.startswith("Done with precompute\n")
    m = re.search(r"numcases = (\d+)\n$", p.stderr)
    assert m, p.stderr
    kn = oracle.knobs(N, 6, num_bv=8, semantics=oracle.SEM_RUST, cutoff=cutoff)
    want, starts = oracle.case_run_starts(kn, oracle.precompute_bv(kn), oracle.mod9_rows(N - 6, 6),
                                          oracle.Node.make(3, 2, 2, 9, 0, 0), cap=1 << 20,
                                          threads=16)
    assert int(m.group(1)) == want.numcases > 0
    got = sorted(int(x) for x in p.stdout.split())
    assert got == sorted(starts) and len(got) == want.overflow_count > 0
