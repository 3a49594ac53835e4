"""The oracle against the reference's own code: golden vectors written by
tests/golden/make_ref_fixtures.py from oracle/_ref (collatz_openclv2.cpp compiled as is),
and the live reference build where it exists."""
import hashlib
import random

import numpy as np
import pytest

from conftest import read_cases


def test_constants(golden, oracle):
    c = golden["constants"]
    assert (c["MAXBITS"], c["LASTBITS"], c["STEPSIZEINITIAL"], c["CUTOFFPOWER"]) == (45, 10, 24, 100)
    assert oracle.tpo(45 - 10) == c["THREEPOWEROFFSET"] == 23
    assert [oracle.magic(23 + i) for i in range(16)] == golden["magic"]


def test_magic_inverts_powers_of_three(oracle):
    for f in range(0, 80):
        assert (oracle.magic(f) * pow(3, f, 1 << 32)) % (1 << 32) == 1


def test_bit_vectors_opencl_semantics(golden, bv_opencl_n45):
    """precompute() of collatz_openclv2.cpp:294-366, all 16 BV_i, word for word, including
    the mirrored 32 words at the end (:343-345)."""
    _, bv, _ = bv_opencl_n45
    for i in range(16):
        w = bv[i].view(np.uint32)
        full = np.concatenate([w, w[:32]])
        assert hashlib.sha256(full.tobytes()).hexdigest() == golden["bv_sha256"][i], i
        assert int(np.unpackbits(w.view(np.uint8)).sum()) == golden["bv_popcount"][i]


def test_bit_vectors_cuda_differ_only_by_unsigned_compare(bv_cuda_n45, bv_opencl_n45, oracle):
    """collatz_cudav7.cu:461 compares signed; collatz_openclv2.cpp:331 unsigned.  The CUDA
    vectors are a subset of the OpenCL ones (the unsigned compare can only lower maxval),
    and they differ, so the two programs do not share a hash."""
    _, bc, _ = bv_cuda_n45
    _, bo, _ = bv_opencl_n45
    assert np.all((bc & ~bo) == 0)
    assert not np.array_equal(bc, bo)
    # The highest vectors keep everything in both.
    assert np.array_equal(bc[15], bo[15])


def test_bit_vectors_cuda_differ_only_where_the_unsigned_compare_flips(bv_cuda_n45,
                                                                      bv_opencl_n45, oracle):
    """The two programs' precompute() differ in one compare: the path-merge test
    curval + (DENOMINATOR - NUMERATOR) > maxval, signed in collatz_cudav7.cu:461 and unsigned
    in collatz_openclv2.cpp:331 (the odd-even-even test beside it never assigns in the C++
    programs, whose doubleeven() is always false).  Unsigned, a negative curval + 179
    compares above any maxval >= 0 and becomes maxval, which can only lower it; so every bit
    the OpenCL vectors have beyond the CUDA ones must belong to a residue n0 whose 24-step
    parity word reaches a negative path-merge term, and the CUDA vectors have no bit the
    OpenCL ones lack.  Bit j of BV_a holds residue n0 = j 3^(TPO + a) mod 2^24 (its index is
    n0 times the magic number 3^-(TPO + a)).  The parity word and running value here follow
    the definitions (curval += 306 per even step, -= 179 per odd one), not the oracle's
    code."""
    _, bc, _ = bv_cuda_n45
    _, bo, _ = bv_opencl_n45
    B, mask = 24, (1 << 24) - 1
    tpo = oracle.tpo(45 - 10)

    def negative_merge_term(n0):
        n, cur, evens = n0, 0, 0
        for _ in range(B):
            if n % 2 == 0:  # an "up" step of the shifted map: n -> 3n/2
                n += n // 2
                cur -= 179
                evens = 0
            else:
                n = n // 2 + 1
                cur += 306
                evens += 1
                if evens % 2 == 0 and cur + 179 < 0:
                    return True
        return False

    extra = 0
    for a in range(16):
        wc, wo = bc[a].view(np.uint32), bo[a].view(np.uint32)
        assert not np.any(wc & ~wo), a
        idx = np.nonzero(np.unpackbits((wo & ~wc).view(np.uint8), bitorder="little"))[0]
        extra += len(idx)
        p = pow(3, tpo + a, 1 << B)
        for j in idx.tolist():
            assert negative_merge_term((j * p) & mask), (a, j)
    assert extra > 0


def test_mod9_rows(golden, oracle):
    rows = oracle.mod9_rows(35, 10)
    assert rows.view(np.uint32).reshape(9, -1).tolist() == golden["mod9"]
    # Closed form: bit i of row a is set iff (a + i * 2^35) mod 9 is in {0,1,3,6,7}.
    bits = np.unpackbits(rows.view(np.uint8), bitorder="little").reshape(9, -1)
    for a in range(9):
        for i in range(1024):
            assert bits[a, i] == (((a + i * (1 << 35)) % 9) in (0, 1, 3, 6, 7))


def test_search3_schedule(golden, oracle):
    """The per-candidate hash of the schedule, including overflow resets at 2^100."""
    for base, num, h in golden["search3"]:
        got = oracle.traj(int(base), int(num), cutoff=100, second_gt=True)[0]
        assert got == h, (base, num)


def test_case_lines(golden, bv_opencl_n45, oracle):
    """Whole case lines: complete hash, launches, leaves per bin (search1 + search2)."""
    kn, bv, m9 = bv_opencl_n45
    for c in golden["cases"][:120]:
        root = oracle.Node.make(*c["case"])
        o = oracle.case_run(kn, bv, m9, root)
        assert o.complete_hash == c["hash"], c["case"]
        assert o.launches == c["launches"]
        assert list(o.bin_leaves)[:16] == c["bins"]
        leaves = oracle.search1(root, 35)
        rows = sorted((lf.base, lf.number, lf.threepower - 23) for lf in leaves
                      if lf.threepower - 23 < 16)
        h = hashlib.sha256()
        for row in rows:
            h.update(("%d %d %d\n" % row).encode())
        assert h.hexdigest() == c["leaves_sha256"]


@pytest.mark.skipif("not __import__('oracle.bind').bind.ref_available()")
def test_live_reference_search2(bv_opencl_n45, oracle):
    kn, bv, m9 = bv_opencl_n45
    rng = random.Random(5)
    lines = read_cases("cases_big_sample.txt")
    for t in rng.sample(lines, 10):
        for lf in oracle.search1(oracle.Node.make(*t), 35)[:50]:
            b = lf.threepower - 23
            h, _ = oracle.leaf_hash(kn, bv, m9, lf)
            assert h == oracle.ref_search2(lf.base, lf.threepower, lf.number, b)


@pytest.mark.skipif("not __import__('oracle.bind').bind.ref_available()")
def test_live_reference_search3_random(oracle):
    rng = random.Random(99)
    for _ in range(3000):
        base = rng.randrange(3, 1 << 64)
        num = rng.randrange(2, 1 << rng.choice([30, 64, 96, 99]))
        assert oracle.traj(base, num, cutoff=100, second_gt=True)[0] == oracle.ref_search3(base, num)
