"""Known answers the reference's paper publishes, and structural facts of the case files."""
import numpy as np
import pytest

from conftest import read_cases

ROOT_NODE = (3, 2, 2, 9, 1, 0)  # collatz_cudav7.cu:6, :173


def test_fig41_bit_vector_densities(oracle):
    """paper p.9 Fig 4.1: BV_0..7 popcounts for N=72, A=6, B=24 under main.rs semantics."""
    kn = oracle.knobs(72, 6, num_bv=8, semantics=oracle.SEM_RUST)
    bv = oracle.precompute_bv(kn)
    pop = [int(np.unpackbits(bv[i].view(np.uint8)).sum()) for i in range(8)]
    assert pop == [333343, 1590275, 3837929, 6750160, 9809263, 12495505, 14479396, 15712298]


@pytest.mark.parametrize("N,expected", [(35, 31109345), (36, 56414851)])
def test_fig51_cases_to_check(oracle, N, expected):
    """paper p.10 Fig 5.1: integers to iterate for N = 35, 36 (A=6, B=24, 8 BVs, Rust)."""
    kn = oracle.knobs(N, 6, num_bv=8, semantics=oracle.SEM_RUST)
    bv = oracle.precompute_bv(kn)
    m9 = oracle.mod9_rows(N - 6, 6)
    assert oracle.rust_count(kn, bv, m9, oracle.Node.make(*ROOT_NODE)) == expected


@pytest.mark.parametrize("k,count", [(20, 16264), (24, 178304), (28, 2051228)])
def test_sieve_sizes(oracle, k, count):
    assert oracle.search1(oracle.Node.make(*ROOT_NODE), k, store=False) == count


def test_cases_tiny_partition_the_tree(oracle):
    """cases_tiny are frontier nodes of search1 from the root: their subtrees' leaves at any
    deeper depth add up to the root's (they partition the tree)."""
    lines = read_cases("cases_tiny.txt")
    depth = 26
    total = sum(oracle.search1(oracle.Node.make(*t), depth, store=False) for t in lines)
    assert total == oracle.search1(oracle.Node.make(*ROOT_NODE), depth, store=False)


def test_case_lines_are_search1_nodes(oracle):
    """Every cases_tiny line is a node reachable from the root with identical state."""
    lines = read_cases("cases_tiny.txt")
    by_depth = {}
    for t in lines:
        by_depth.setdefault(t[1], []).append(tuple(t))
    for d, ts in by_depth.items():
        nodes = {(n.base, n.sigfig, n.threepower, n.number, n.doubleeven, n.numevens)
                 for n in oracle.search1(oracle.Node.make(*ROOT_NODE), d)}
        for t in ts:
            assert t in nodes, t
