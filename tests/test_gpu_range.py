"""collatz_verify_range (plain and persistent forms) against the oracle, bit for bit."""
import random

import numpy as np
import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def s20():
    import collatz_amd as cz
    return cz.Sieve(20)


@pytest.fixture(scope="module")
def o20(oracle):
    return oracle.sieve(20)


@pytest.fixture(scope="module")
def s32():
    import collatz_amd as cz
    return cz.Sieve(32)


@pytest.fixture(scope="module")
def o32(oracle):
    return oracle.sieve(32)


def check(sv, osv, oracle, lo, hi, C=24, limbs=2, cutoff=187, forms=(0, 1)):
    chunks, total = oracle.verify_range(osv, lo, hi, C, cutoff, threads=16)
    outs = []
    for form in forms:
        g = sv.verify_range(lo, hi, C, limbs, cutoff, form=form, overflow_cap=1 << 18)
        assert_chunks_equal(g.chunks, chunks, f"[{lo:#x},{hi:#x}) form {form} limbs {limbs}")
        assert_total_equal(g.total, total, "total")
        outs.append(g)
    return outs, total


def test_config0_full(s20, o20, oracle):
    """Every n in [1, 2^32), k = 20, 2^24-wide chunks (256 records)."""
    outs, total = check(s20, o20, oracle, 1, 1 << 32)
    assert len(outs[0].chunks) == 256
    assert int(total["n_checked"]) == (1 << 32) - 1
    assert int(total["overflow_count"]) == 0


def test_config1_slice(s32, o32, oracle):
    """A 2^30-wide slice of [2^40, 2^40 + 2^36) with the k = 32 sieve (64 chunk records)."""
    lo = (1 << 40) + (1 << 35)
    check(s32, o32, oracle, lo, lo + (1 << 30))


def test_config1_full_properties(s32):
    """The whole [2^40, 2^40 + 2^36): forms agree, records add up."""
    lo, hi = 1 << 40, (1 << 40) + (1 << 36)
    a = s32.verify_range(lo, hi, 24, form=0)
    b = s32.verify_range(lo, hi, 24, form=1)
    assert_chunks_equal(a.chunks, b.chunks, "plain vs persistent")
    assert len(a.chunks) == 4096
    assert int(a.chunks["n_checked"].sum()) == 1 << 36
    assert int(a.chunks["count"].sum()) == int(a.total["count"])
    h = 0
    for x in a.chunks["hash"]:
        h ^= int(x)
    assert h == int(a.total["hash"])
    assert a.escalated == 0 and int(a.total["overflow_count"]) == 0


@pytest.mark.parametrize("lo,hi,C", [
    (1, 2, 24), (3, 4, 24), (27, 28, 10), (0, 1 << 21, 20), (3, (1 << 26) + 12345, 20),
    ((1 << 40) + 777, (1 << 40) + (1 << 25), 22), ((1 << 64) - (1 << 21), (1 << 64) + (1 << 21), 20),
    ((1 << 100) + 3, (1 << 100) + (1 << 21), 19), ((1 << 127) - (1 << 20), (1 << 127), 16),
    ((1 << 50) + 1, (1 << 50) + (1 << 23) + 5, 32), ((1 << 33) + 9, (1 << 33) + (1 << 22), 63),
])
def test_edge_ranges(s20, o20, oracle, lo, hi, C):
    check(s20, o20, oracle, lo, hi, C)


@pytest.mark.parametrize("cutoff", [64, 70, 100, 123, 127, 128, 150, 187])
def test_cutoffs_and_overflow_lists(s20, o20, oracle, cutoff):
    lo = (1 << 62) + 5
    hi = lo + (1 << 21)
    outs, total = check(s20, o20, oracle, lo, hi, 20, cutoff=cutoff)
    a, b = outs
    assert sorted(a.overflow_starts) == sorted(b.overflow_starts)
    assert len(a.overflow_starts) == int(total["overflow_count"])
    if cutoff == 70:
        assert len(a.overflow_starts) > 0


def test_three_limbs_and_escalation(s20, o20, oracle):
    """Starts near 2^110 reach 2^123: two-limb requests that far up run on three limbs from
    the start, and give the same records as a three-limb request."""
    lo = (1 << 110) + 12345
    hi = lo + (1 << 20)
    outs2, total = check(s20, o20, oracle, lo, hi, 18, limbs=2)
    outs3, _ = check(s20, o20, oracle, lo, hi, 18, limbs=3)
    assert outs3[0].escalated == 0
    assert int(total["flags"]) & 2


def test_two_limb_escalation(s20, o20, oracle):
    """n = 0x7737a959485e163f092f585bf (< 2^100, so two u64 limbs) climbs to a 127-bit
    excursion: it ends a check above 2^118, is escalated and rerun on three limbs."""
    n = 0x7737A959485E163F092F585BF
    outs, total = check(s20, o20, oracle, n - (1 << 10), n + (1 << 10), C=10)
    assert outs[0].escalated >= 1 and outs[1].escalated >= 1
    assert int(total["flags"]) & 2


def test_narrow_width_escalation(s20, o20, oracle):
    """Two-limb ranges below 2^68 run on 96 bits and escalate above 2^86.  n =
    0x8b3aed82f55cb4c9b (< 2^68) climbs to a 92-bit excursion, so its candidate is escalated
    and rerun on three limbs; the records must equal the oracle's (which runs everything on
    wide integers)."""
    n = 0x8B3AED82F55CB4C9B
    outs, _ = check(s20, o20, oracle, n - (1 << 12), n + (1 << 12), C=12)
    assert outs[0].escalated >= 1 and outs[1].escalated >= 1
    assert int(outs[0].total["max_exc"][1]) >> 27 != 0


@pytest.mark.parametrize("limbs,cutoff", [(2, 187), (2, 100), (2, 64), (3, 187)])
def test_low_ranges(s20, o20, oracle, limbs, cutoff):
    """Ranges ending at or below 2^35 take the low-key test of the wave max (the plain
    kernel on three words, or four under a cutoff below 2^91; the persistent kernel on three
    words): records equal to the oracle's in both forms."""
    check(s20, o20, oracle, 3, (1 << 24) + 12345, 20, limbs=limbs, cutoff=cutoff)


def test_low_range_excursion_past_2p64(s20, o20, oracle):
    """n = 23035537407 (< 2^35) climbs to 34419078320774113520 > 2^64 in the shifted map:
    under the low-key test the wave's best then has a nonzero upper word, and every value
    below 2^64 must stop reaching it."""
    n = 23035537407
    outs, total = check(s20, o20, oracle, n - (1 << 20), n + (1 << 20), 20)
    assert int(total["max_exc"][1]) == 1 and int(total["max_start"][0]) == n


@pytest.mark.parametrize("lo,hi", [
    ((1 << 68) - (1 << 22), 1 << 68),             # last range on 96 bits
    ((1 << 68) - (1 << 21), (1 << 68) + (1 << 21)),  # first above: two u64 limbs
    ((1 << 86) - (1 << 20), (1 << 86) + 5),       # starts at the 96-bit check bound (four words)
    ((1 << 100) - (1 << 20), (1 << 100)),         # last range on four words
    ((1 << 100) - (1 << 20), (1 << 100) + 1),     # first one on three limbs
])
def test_width_boundaries(s20, o20, oracle, lo, hi):
    check(s20, o20, oracle, lo, hi, C=20)


def test_limbs3_low_cutoff(s20, o20, oracle):
    check(s20, o20, oracle, (1 << 62) + 5, (1 << 62) + (1 << 20), 20, limbs=3, cutoff=100)
    check(s20, o20, oracle, (1 << 62) + 5, (1 << 62) + (1 << 20), 20, limbs=3, cutoff=125)


def test_mod9_off(oracle):
    import collatz_amd as cz
    sv = cz.Sieve(20, mod9=False)
    o = oracle.sieve(20)
    lo, hi = (1 << 36) + 1, (1 << 36) + (1 << 24)
    chunks, total = oracle.verify_range(o, lo, hi, 22, mod9=False)
    g = sv.verify_range(lo, hi, 22)
    assert_chunks_equal(g.chunks, chunks, "mod9 off")


def test_random_ranges(s20, o20, oracle):
    rng = random.Random(11)
    for _ in range(25):
        lo = rng.randrange(1, 1 << rng.choice([30, 45, 61, 64, 80]))
        hi = lo + rng.randrange(1, 1 << rng.choice([8, 16, 22]))
        check(s20, o20, oracle, lo, hi, rng.choice([12, 16, 20, 24]))


def test_bad_arguments(s20):
    import collatz_amd as cz
    with pytest.raises(cz.CollatzError):
        s20.verify_range(5, 5)
    with pytest.raises(cz.CollatzError):
        s20.verify_range(1, 100, cutoff_log2=60)
    with pytest.raises(cz.CollatzError):
        s20.verify_range(1, 100, limbs=4)


def test_low_cutoff_start_above_2p123(oracle):
    """Two limbs with a cutoff of 2^64 and starts near 2^122 (k = 16): start values reach
    2^128 after the jump and odd runs of starts above 2^123 leave 128 bits before the
    overflow reset, so such launches run on six words; the max excursion (taken before
    the reset) must be exact."""
    import collatz_amd as cz
    lo, hi = 0x32b68ae4d0284133bd56836441af003, 0x32b68ae4d0284133bd568364427b3e6
    o16 = oracle.sieve(16)
    for la, sem in ((0, 0), (8, 2)):
        chunks, total = oracle.verify_range(o16, lo, hi, 20, 64, threads=16, lookahead_bits=la,
                                            lookahead_semantics=sem)
        sv = cz.Sieve(16, lookahead_bits=la, lookahead_semantics=sem)
        for form in (0, 1):
            g = sv.verify_range(lo, hi, 20, 2, 64, form=form, overflow_cap=1 << 20)
            assert_chunks_equal(g.chunks, chunks, f"la {la} form {form}")
            assert_total_equal(g.total, total, "total")
        assert int(total["max_exc"][2]) >> 4 != 0   # the excursion itself is above 2^132
        sv.close()
