"""ctypes bindings for the CPU oracle (liboracle.so) and for the reference's own CPU code
compiled by oracle/Makefile (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

SEM_CXX = 0
SEM_RUST = 1
SEM_OPENCL = 2
WIDE_LOG2 = 123
MASK64 = (1 << 64) - 1


class Node(C.Structure):
    _fields_ = [
        ("base_lo", C.c_uint64), ("base_hi", C.c_uint64),
        ("number_lo", C.c_uint64), ("number_hi", C.c_uint64),
        ("sigfig", C.c_uint32), ("threepower", C.c_uint32),
        ("doubleeven", C.c_uint32), ("numevens", C.c_uint32),
    ]

    @classmethod
    def make(cls, base, sigfig, threepower, number, doubleeven, numevens):
        return cls(base & MASK64, base >> 64, number & MASK64, number >> 64,
                   sigfig, threepower, doubleeven, numevens)

    @property
    def base(self):
        return self.base_lo | (self.base_hi << 64)

    @property
    def number(self):
        return self.number_lo | (self.number_hi << 64)


class Knobs(C.Structure):
    _fields_ = [("N", C.c_uint32), ("A", C.c_uint32), ("B", C.c_uint32),
                ("num_bv", C.c_uint32), ("semantics", C.c_uint32), ("cutoff", C.c_uint32)]


class TrajOut(C.Structure):
    _fields_ = [("hash", C.c_uint32), ("overflow", C.c_uint32), ("steps", C.c_uint64),
                ("maxnum", C.c_uint64 * 3), ("final", C.c_uint64 * 3)]


class CaseOut(C.Structure):
    _fields_ = [("complete_hash", C.c_uint32), ("pad", C.c_uint32), ("launches", C.c_uint64),
                ("leaves", C.c_uint64), ("candidates", C.c_uint64),
                ("bin_leaves", C.c_uint64 * 17), ("overflow_count", C.c_uint64),
                ("numcases", C.c_uint64)]


class Chunk(C.Structure):
    _fields_ = [
        ("chunk_index", C.c_uint64), ("n_checked", C.c_uint64), ("count", C.c_uint64),
        ("stop_sum", C.c_uint64), ("stop_xor", C.c_uint64), ("max_exc", C.c_uint64 * 3),
        ("max_start", C.c_uint64 * 2), ("overflow_count", C.c_uint64), ("hash", C.c_uint32),
        ("flags", C.c_uint32),
    ]


CHUNK_DTYPE = np.dtype([
    ("chunk_index", "<u8"), ("n_checked", "<u8"), ("count", "<u8"), ("stop_sum", "<u8"),
    ("stop_xor", "<u8"), ("max_exc", "<u8", (3,)), ("max_start", "<u8", (2,)),
    ("overflow_count", "<u8"), ("hash", "<u4"), ("flags", "<u4"),
])
assert CHUNK_DTYPE.itemsize == C.sizeof(Chunk)

_P = C.POINTER
_u64p = _P(C.c_uint64)


def _ptr(a, ctype):
    return a.ctypes.data_as(_P(ctype))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.or_pow3.argtypes = [C.c_uint32, _u64p]
        L.or_magic.argtypes = [C.c_uint32]
        L.or_magic.restype = C.c_uint32
        L.or_tpo.argtypes = [C.c_uint32]
        L.or_tpo.restype = C.c_uint32
        L.or_search1.argtypes = [_P(Node), C.c_uint32, C.c_uint32, _P(Node), C.c_int64]
        L.or_search1.restype = C.c_int64
        L.or_sieve_sorted.argtypes = [C.c_uint32, _u64p, _u64p, _P(C.c_uint8), C.c_int64]
        L.or_sieve_sorted.restype = C.c_int64
        L.or_sieve_windows.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _u64p, C.c_int,
                                       _u64p, _u64p, _P(C.c_uint8), C.c_int64]
        L.or_sieve_windows.restype = C.c_int64
        L.or_precompute_bv.argtypes = [_P(Knobs), _u64p]
        L.or_mod9_rows.argtypes = [C.c_uint32, C.c_uint32, _u64p]
        L.or_doubleeven.argtypes = [C.c_uint32] * 4
        L.or_traj.argtypes = [_u64p, _u64p, C.c_uint32, C.c_int, _P(TrajOut)]
        L.or_traj.restype = None
        L.or_leaf_hash.argtypes = [_P(Knobs), _u64p, _u64p, _P(Node), _u64p]
        L.or_leaf_hash.restype = C.c_uint32
        L.or_case_run.argtypes = [_P(Knobs), _u64p, _u64p, _P(Node), _P(CaseOut)]
        L.or_case_run_starts.argtypes = [_P(Knobs), _u64p, _u64p, _P(Node), _P(CaseOut),
                                         _u64p, C.c_uint64]
        L.or_case_run_mt.argtypes = [_P(Knobs), _u64p, _u64p, _P(Node), _P(CaseOut),
                                     _u64p, C.c_uint64, C.c_int]
        L.or_rust_search3.argtypes = [_u64p, _u64p, C.c_uint32, _u64p]
        L.or_rust_search3.restype = C.c_int
        L.or_rust_count.argtypes = [_P(Knobs), _u64p, _u64p, _P(Node)]
        L.or_rust_count.restype = C.c_int64
        L.or_verify_range.argtypes = [C.c_uint32, C.c_int64, _u64p, _u64p, _P(C.c_uint8),
                                      C.c_int, _u64p, _u64p, C.c_uint32, C.c_uint32, C.c_int,
                                      _P(Chunk), _P(Chunk)]
        L.or_verify_range_la.argtypes = [C.c_uint32, C.c_int64, _u64p, _u64p,
                                         _P(C.c_uint8), C.c_int, C.c_uint32, C.c_uint32,
                                         _u64p, _u64p, C.c_uint32, C.c_uint32, C.c_int,
                                         _P(Chunk), _P(Chunk)]
        L.or_la_dip.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.or_la_dip.restype = C.c_int32
        _lib = L
    return _lib


def pow3(f):
    out = (C.c_uint64 * 2)()
    lib().or_pow3(f, out)
    return out[0] | (out[1] << 64)


def magic(e):
    return lib().or_magic(e)


def tpo(mbi):
    return lib().or_tpo(mbi)


def search1(root: Node, depth: int, semantics: int = SEM_CXX, store: bool = True):
    n = lib().or_search1(C.byref(root), depth, semantics, None, 0)
    if n < 0:
        raise ValueError("bad search1 input")
    if not store:
        return n
    arr = (Node * max(n, 1))()
    lib().or_search1(C.byref(root), depth, semantics, arr, n)
    return [arr[i] for i in range(n)]


@dataclass
class Sieve:
    k: int
    r: np.ndarray
    number: np.ndarray
    f: np.ndarray


def sieve(k: int) -> Sieve:
    n = lib().or_sieve_sorted(k, None, None, None, 0)
    r = np.empty(n, np.uint64)
    num = np.empty(n, np.uint64)
    f = np.empty(n, np.uint8)
    got = lib().or_sieve_sorted(k, _ptr(r, C.c_uint64), _ptr(num, C.c_uint64),
                                _ptr(f, C.c_uint8), n)
    assert got == n
    return Sieve(k, r, num, f)


def sieve_windows(k: int, w: int, tops, threads: int = 1) -> Sieve:
    """The survivors of sieve(k) with r >> w in tops: all that chunks whose residues mod 2^k
    lie in the windows [top 2^w, (top + 1) 2^w) use."""
    t = np.array(sorted(set(int(x) for x in tops)), np.uint64)
    cap = len(t) * ((1 << w) // 64) + 4096  # above the survivor density of any k <= 40
    while True:
        r = np.empty(cap, np.uint64)
        num = np.empty(cap, np.uint64)
        f = np.empty(cap, np.uint8)
        n = lib().or_sieve_windows(k, w, len(t), _ptr(t, C.c_uint64), threads,
                                   _ptr(r, C.c_uint64), _ptr(num, C.c_uint64),
                                   _ptr(f, C.c_uint8), cap)
        assert n >= 0
        if n <= cap:
            return Sieve(k, r[:n].copy(), num[:n].copy(), f[:n].copy())
        cap = n


def knobs(N, A, B=24, num_bv=16, semantics=SEM_CXX, cutoff=100):
    return Knobs(N, A, B, num_bv, semantics, cutoff)


def precompute_bv(kn: Knobs) -> np.ndarray:
    words = 1 << (kn.B - 6)
    bv = np.zeros((kn.num_bv, words), np.uint64)
    assert lib().or_precompute_bv(C.byref(kn), _ptr(bv, C.c_uint64)) == 0
    return bv


def mod9_rows(mbi, A) -> np.ndarray:
    words = max(1, 1 << (A - 6)) if A >= 6 else 1
    rows = np.zeros((9, words), np.uint64)
    assert lib().or_mod9_rows(mbi, A, _ptr(rows, C.c_uint64)) == 0
    return rows


def traj(base: int, num0: int, cutoff: int = 187, second_gt: bool = False, final=False):
    b = (C.c_uint64 * 2)(base & MASK64, base >> 64)
    v = (C.c_uint64 * 3)(num0 & MASK64, (num0 >> 64) & MASK64, num0 >> 128)
    o = TrajOut()
    lib().or_traj(b, v, cutoff, int(second_gt), C.byref(o))
    mx = o.maxnum[0] | (o.maxnum[1] << 64) | (o.maxnum[2] << 128)
    if final:
        fin = o.final[0] | (o.final[1] << 64) | (o.final[2] << 128)
        return o.hash, o.overflow, o.steps, mx, fin
    return o.hash, o.overflow, o.steps, mx


def leaf_hash(kn, bv, mod9, leaf: Node):
    nc = C.c_uint64()
    h = lib().or_leaf_hash(C.byref(kn), _ptr(bv, C.c_uint64), _ptr(mod9, C.c_uint64),
                           C.byref(leaf), C.byref(nc))
    return h, nc.value


def case_run(kn, bv, mod9, root: Node) -> CaseOut:
    out = CaseOut()
    assert lib().or_case_run(C.byref(kn), _ptr(bv, C.c_uint64), _ptr(mod9, C.c_uint64),
                             C.byref(root), C.byref(out)) == 0
    return out


def case_run_starts(kn, bv, mod9, root: Node, cap: int = 1 << 16, threads: int = 1):
    """case_run plus the start n of up to `cap` overflow reports (on `threads` threads)."""
    out = CaseOut()
    buf = np.zeros(2 * max(cap, 1), np.uint64)
    assert lib().or_case_run_mt(C.byref(kn), _ptr(bv, C.c_uint64), _ptr(mod9, C.c_uint64),
                                C.byref(root), C.byref(out), _ptr(buf, C.c_uint64), cap,
                                threads) == 0
    k = min(out.overflow_count, cap)
    return out, [int(buf[2 * i]) | (int(buf[2 * i + 1]) << 64) for i in range(k)]


def rust_search3(base: int, num0: int, cutoff: int = 120):
    """main.rs's search3 from start value num0: (overflowed, tally of 16-step blocks)."""
    b = (C.c_uint64 * 2)(base & MASK64, base >> 64)
    v = (C.c_uint64 * 3)(num0 & MASK64, (num0 >> 64) & MASK64, num0 >> 128)
    t = C.c_uint64()
    r = lib().or_rust_search3(b, v, cutoff, C.byref(t))
    return bool(r), t.value


def rust_count(kn, bv, mod9, root: Node) -> int:
    return lib().or_rust_count(C.byref(kn), _ptr(bv, C.c_uint64), _ptr(mod9, C.c_uint64),
                               C.byref(root))


def nchunks(lo, hi, C_log2):
    return ((hi - 1) >> C_log2) - (lo >> C_log2) + 1


def la_dip(n0: int, B: int, semantics: int = SEM_CXX) -> int:
    return lib().or_la_dip(n0, B, semantics)


def verify_range(sv: Sieve, lo: int, hi: int, chunk_log2: int, cutoff: int = 187,
                 mod9: bool = True, threads: int = 1, lookahead_bits: int = 0,
                 lookahead_semantics: int = SEM_CXX):
    n = nchunks(lo, hi, chunk_log2)
    chunks = np.zeros(n, CHUNK_DTYPE)
    total = Chunk()
    lo2 = (C.c_uint64 * 2)(lo & MASK64, lo >> 64)
    hi2 = (C.c_uint64 * 2)(hi & MASK64, hi >> 64)
    rc = lib().or_verify_range_la(sv.k, len(sv.r), _ptr(sv.r, C.c_uint64),
                                  _ptr(sv.number, C.c_uint64), _ptr(sv.f, C.c_uint8),
                                  int(mod9), lookahead_bits, lookahead_semantics, lo2, hi2,
                                  chunk_log2, cutoff, threads,
                                  chunks.ctypes.data_as(_P(Chunk)), C.byref(total))
    if rc != 0:
        raise ValueError("or_verify_range failed")
    tot = np.frombuffer(bytearray(total), CHUNK_DTYPE)[0]
    return chunks, tot


# ----------------------------------------------------------------------------------------
# The reference's own CPU functions (collatz_openclv2.cpp), built by oracle/Makefile.
# ----------------------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_constants.argtypes = [_P(C.c_uint32)]
        R.ref_init.argtypes = []
        R.ref_bv.argtypes = [C.c_uint32, _P(C.c_uint32)]
        R.ref_mod9.argtypes = [_P(C.c_uint32)]
        R.ref_magic.argtypes = [_P(C.c_uint32)]
        R.ref_search3.argtypes = [C.c_uint64] * 4
        R.ref_search3.restype = C.c_uint32
        R.ref_search2.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64,
                                  C.c_uint32]
        R.ref_search2.restype = C.c_uint32
        R.ref_case.argtypes = [_u64p, _P(C.c_uint32), _u64p, _u64p, _u64p, C.c_int64]
        R.ref_case.restype = C.c_int64
        R.ref_init()
        _ref = R
    return _ref


def ref_constants():
    out = (C.c_uint32 * 8)()
    ref().ref_constants(out)
    keys = ["MAXBITS", "LASTBITS", "STEPSIZEINITIAL", "CUTOFFPOWER", "THREEPOWEROFFSET",
            "INITIALBITARRAYSIZE", "LASTBITSTEPS", "BATCHSIZE"]
    return dict(zip(keys, list(out)))


def ref_bv(i):
    n = ref_constants()["INITIALBITARRAYSIZE"]
    out = np.zeros(n, np.uint32)
    assert ref().ref_bv(i, _ptr(out, C.c_uint32)) == 0
    return out


def ref_mod9():
    n = ref_constants()["LASTBITSTEPS"]
    out = np.zeros((9, n), np.uint32)
    assert ref().ref_mod9(_ptr(out, C.c_uint32)) == 0
    return out


def ref_magic():
    out = np.zeros(16, np.uint32)
    ref().ref_magic(_ptr(out, C.c_uint32))
    return out


def ref_search3(base, number):
    return ref().ref_search3(base & MASK64, base >> 64, number & MASK64, number >> 64)


def ref_search2(base, threepower, number, bigindex):
    return ref().ref_search2(base & MASK64, base >> 64, threepower, number & MASK64,
                             number >> 64, bigindex)


def ref_case(t, cap=1 << 20):
    tt = (C.c_uint64 * 6)(*t)
    h = C.c_uint32()
    launches = C.c_uint64()
    bins = (C.c_uint64 * 16)()
    leaves = np.zeros((cap, 3), np.uint64)
    n = ref().ref_case(tt, C.byref(h), C.byref(launches), bins, _ptr(leaves, C.c_uint64), cap)
    if n < 0:
        return None
    return dict(hash=h.value, launches=launches.value, bins=list(bins),
                leaves=leaves[:min(n, cap)].copy(), nleaves=n)
