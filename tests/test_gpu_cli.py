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
        m = re.search(r"Number of kernel launches = (\d+) \nSpot checks = 0\. Complete hash = ([0-9A-F]+)\.",
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
