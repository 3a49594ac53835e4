"""Seeded random combinations of the knobs (sieve depth, look-ahead depth and compares,
cutoff, word width, form) on small ranges: every record equal to the oracle's."""
import random

import pytest

from gpu_util import assert_chunks_equal, assert_total_equal

pytestmark = pytest.mark.gpu


def cases(seed, n):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        k = rng.choice([12, 16, 20, 24])
        la = rng.choice([0, 0, 8, 12, 16, 20])
        sem = rng.choice([0, 2]) if la else 0
        cutoff = rng.choice([64, 70, 96, 100, 123, 128, 150, 187])
        bits = rng.choice([30, 45, 60, 64, 66, 90, 110, 126])
        lo = rng.getrandbits(bits) | 1
        width = rng.randrange(1 << 18, 1 << rng.choice([20, 22, 24, 26]))
        C = rng.choice([16, 20, 24])
        limbs = rng.choice([2, 2, 3])
        out.append((k, la, sem, cutoff, lo, lo + width, C, limbs))
    return out


@pytest.mark.parametrize("seed", [2024, 7, 99])
def test_random_combinations(oracle, seed):
    import collatz_amd as cz
    sieves, osieves = {}, {}
    for k, la, sem, cutoff, lo, hi, C, limbs in cases(seed, 200):
        key = (k, la, sem)
        if key not in sieves:
            sieves[key] = cz.Sieve(k, lookahead_bits=la, lookahead_semantics=sem)
        if k not in osieves:
            osieves[k] = oracle.sieve(k)
        chunks, total = oracle.verify_range(osieves[k], lo, hi, C, cutoff, threads=16,
                                            lookahead_bits=la, lookahead_semantics=sem)
        tag = f"k={k} la={la}/{sem} cutoff={cutoff} [{lo:#x},{hi:#x}) C={C} limbs={limbs}"
        for form in (cz.FORM_PLAIN, cz.FORM_PERSISTENT):
            g = sieves[key].verify_range(lo, hi, C, limbs, cutoff, form=form,
                                         overflow_cap=1 << 20)
            assert_chunks_equal(g.chunks, chunks, f"{tag} form {form}")
            assert_total_equal(g.total, total, tag)
        b = sieves[key].verify_batch([(lo, hi)], chunk_log2=C, limbs=limbs, cutoff_log2=cutoff)
        assert_chunks_equal(b[0].chunks, chunks, f"{tag} batch")


@pytest.mark.parametrize("seed", [5, 6])
def test_random_batches_and_server(oracle, seed):
    """Many intervals of mixed magnitude (2^20 .. 2^127) in one batch and through the
    server, with a random cutoff, chunk size and look-ahead: each interval's records equal
    the oracle's."""
    import collatz_amd as cz
    rng = random.Random(seed)
    k = rng.choice([16, 20, 24])
    la = rng.choice([0, 12, 16])
    cutoff = rng.choice([64, 80, 100, 128, 187])
    C = rng.choice([18, 22])
    ivs = []
    for _ in range(40):
        lo = rng.getrandbits(rng.choice([20, 40, 62, 65, 90, 120, 127])) | 1
        ivs.append((lo, lo + rng.randrange(1 << 16, 1 << 23)))
    sv = cz.Sieve(k, lookahead_bits=la)
    o = oracle.sieve(k)
    want = [oracle.verify_range(o, lo, hi, C, cutoff, threads=16, lookahead_bits=la)
            for lo, hi in ivs]
    tag = f"k={k} la={la} cutoff={cutoff} C={C}"
    for graph in (False, True):
        got = sv.verify_batch(ivs, chunk_log2=C, cutoff_log2=cutoff, use_graph=graph)
        for (lo, hi), g, (chunks, total) in zip(ivs, got, want):
            assert_chunks_equal(g.chunks, chunks, f"{tag} batch [{lo:#x},{hi:#x})")
            assert_total_equal(g.total, total, tag)
    srv = cz.Server(sv, chunk_log2=C, cutoff_log2=cutoff)
    tickets = [srv.submit(lo, hi) for lo, hi in ivs]
    for t, (lo, hi), (chunks, total) in zip(tickets, ivs, want):
        tot, ch, _ = srv.wait(t)
        assert_chunks_equal(ch, chunks, f"{tag} server [{lo:#x},{hi:#x})")
    srv.close()
    sv.close()
