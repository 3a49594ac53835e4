"""Python bindings for libcollatz_amd.so, the MI355X Collatz range verifier.

The library is hand-written HIP for gfx950 behind the C ABI declared in include/collatz.h;
this module only loads it with ctypes and converts results to numpy records.  There is no
fallback: if the library is missing, or has no HIP device to run on, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcollatz_amd.so")
MASK64 = (1 << 64) - 1

OK, EINVAL, ENOMEM, EDEVICE, ECAPACITY, ENOTREADY = 0, -1, -2, -3, -4, -5
SEM_CUDA, SEM_RUST, SEM_OPENCL = 0, 1, 2
FLAG_OVERFLOW, FLAG_WIDE = 1, 2
FORM_PLAIN, FORM_PERSISTENT = 0, 1

# Entry points declared in include/collatz.h.
EXPORTS = [
    "collatz_sieve_build", "collatz_sieve_query", "collatz_sieve_copy", "collatz_sieve_free",
    "collatz_verify_range", "collatz_verify_range_noalloc", "collatz_verify_batch",
    "collatz_reference_run",
    "collatz_reference_tables",
    "collatz_server_create", "collatz_server_submit", "collatz_server_wait",
    "collatz_server_destroy", "collatz_version", "collatz_last_error",
]


class SieveParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("mod9", C.c_uint32), ("device", C.c_int32),
                ("lookahead_bits", C.c_uint32), ("lookahead_semantics", C.c_uint32),
                ("reserved", C.c_uint32)]


class SieveInfo(C.Structure):
    _fields_ = [("k", C.c_uint32), ("mod9", C.c_uint32), ("survivors", C.c_uint64),
                ("bitarray_bits", C.c_uint64), ("device_bytes", C.c_uint64),
                ("build_ms", C.c_double), ("lookahead_bits", C.c_uint32),
                ("lookahead_semantics", C.c_uint32)]


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


class RangeResult(C.Structure):
    _fields_ = [("total", Chunk), ("chunks", C.POINTER(Chunk)), ("chunks_cap", C.c_uint64),
                ("chunks_written", C.c_uint64), ("overflow_starts", C.POINTER(C.c_uint64)),
                ("overflow_cap", C.c_uint64), ("overflow_written", C.c_uint64),
                ("escalated", C.c_uint64), ("checks", C.c_uint64), ("device_ms", C.c_double),
                ("spot_checks", C.c_uint64)]


RESULT_DTYPE = np.dtype([
    ("total", CHUNK_DTYPE), ("chunks", "<u8"), ("chunks_cap", "<u8"), ("chunks_written", "<u8"),
    ("overflow_starts", "<u8"), ("overflow_cap", "<u8"), ("overflow_written", "<u8"),
    ("escalated", "<u8"), ("checks", "<u8"), ("device_ms", "<f8"), ("spot_checks", "<u8"),
])
assert RESULT_DTYPE.itemsize == C.sizeof(RangeResult)


class VerifyOpts(C.Structure):
    _fields_ = [("chunk_log2", C.c_uint32), ("limbs", C.c_uint32), ("cutoff_log2", C.c_uint32),
                ("form", C.c_uint32)]


class Interval(C.Structure):
    _fields_ = [("lo", C.c_uint64 * 2), ("hi", C.c_uint64 * 2)]


class BatchOpts(C.Structure):
    _fields_ = [("chunk_log2", C.c_uint32), ("limbs", C.c_uint32), ("cutoff_log2", C.c_uint32),
                ("use_graph", C.c_uint32), ("tile_residues", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("checks_per_refill", C.c_uint32),
                ("spot_checks", C.c_uint32)]


class Node(C.Structure):
    _fields_ = [("base", C.c_uint64 * 2), ("number", C.c_uint64 * 2), ("sigfig", C.c_uint32),
                ("threepower", C.c_uint32), ("doubleeven", C.c_uint32),
                ("numevens", C.c_uint32)]


class RefParams(C.Structure):
    _fields_ = [("N", C.c_uint32), ("A", C.c_uint32), ("B", C.c_uint32), ("num_bv", C.c_uint32),
                ("semantics", C.c_uint32), ("cutoff_log2", C.c_uint32), ("device", C.c_int32),
                ("spot_check_freq", C.c_uint32)]


class RefResult(C.Structure):
    _fields_ = [("complete_hash", C.c_uint32), ("pad", C.c_uint32), ("launches", C.c_uint64),
                ("leaves", C.c_uint64), ("candidates", C.c_uint64), ("checks", C.c_uint64),
                ("bin_leaves", C.c_uint64 * 17), ("overflow_count", C.c_uint64),
                ("overflow_starts", C.POINTER(C.c_uint64)), ("overflow_cap", C.c_uint64),
                ("overflow_written", C.c_uint64), ("device_ms", C.c_double),
                ("numcases", C.c_uint64), ("spot_checks", C.c_uint64),
                ("spot_check_failures", C.c_uint64)]


class ServerOpts(C.Structure):
    _fields_ = [("chunk_log2", C.c_uint32), ("cutoff_log2", C.c_uint32),
                ("max_device_bytes", C.c_uint64), ("slice_candidates", C.c_uint64),
                ("max_inflight", C.c_uint32), ("reserved", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("total", Chunk), ("chunks", C.c_uint64), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("done_ms", C.c_double)]


class CollatzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"collatz error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    """Loads the HIP library (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    P = C.POINTER
    L.collatz_version.restype = C.c_int
    L.collatz_last_error.restype = C.c_char_p
    L.collatz_sieve_build.argtypes = [P(SieveParams), P(C.c_void_p)]
    L.collatz_sieve_query.argtypes = [C.c_void_p, P(SieveInfo)]
    L.collatz_sieve_copy.argtypes = [C.c_void_p, P(C.c_uint64), P(C.c_uint64), P(C.c_uint8),
                                     C.c_uint64]
    L.collatz_sieve_free.argtypes = [C.c_void_p]
    L.collatz_verify_range.argtypes = [C.c_void_p, P(C.c_uint64), P(C.c_uint64), P(VerifyOpts),
                                       C.c_void_p, P(RangeResult)]
    L.collatz_verify_range_noalloc.argtypes = L.collatz_verify_range.argtypes
    L.collatz_verify_batch.argtypes = [C.c_void_p, C.c_uint64, P(Interval), P(BatchOpts),
                                       C.c_void_p, P(RangeResult)]
    L.collatz_reference_run.argtypes = [P(RefParams), P(Node), C.c_void_p, P(RefResult)]
    L.collatz_reference_tables.argtypes = [P(RefParams), P(C.c_uint32), P(C.c_uint32)]
    L.collatz_server_create.argtypes = [C.c_void_p, P(ServerOpts), P(C.c_void_p)]
    L.collatz_server_submit.argtypes = [C.c_void_p, P(Interval)]
    L.collatz_server_submit.restype = C.c_int64
    L.collatz_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_double, P(RequestResult),
                                      P(Chunk), C.c_uint64]
    L.collatz_server_destroy.argtypes = [C.c_void_p]
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise CollatzError(rc, load().collatz_last_error().decode())


def u128_pair(x: int):
    return (C.c_uint64 * 2)(x & MASK64, (x >> 64) & MASK64)


def words_to_int(ws) -> int:
    v = 0
    for i, w in enumerate(ws):
        v |= int(w) << (64 * i)
    return v


@dataclass
class RangeOutput:
    total: np.void
    chunks: np.ndarray
    overflow_starts: list
    escalated: int
    device_ms: float
    checks: int = 0
    overflow_written: int = 0  # overflow reports made (may exceed the starts listed)
    spot_checks: int = 0       # chunk records re-run by the plain kernel and found equal


class Sieve:
    """Resident survivor list of the 2^k residue sieve (collatz_sieve_build).

    lookahead_bits = B adds the reference's Step 2 filter at depth k (a candidate is
    iterated only if the B-step dip of its start value mod 2^B is at most 485 f - 306 k,
    collatz_cudav7.cu:219-235, :423-480), with the BV_i compare semantics of
    lookahead_semantics (SEM_CUDA or SEM_OPENCL)."""

    def __init__(self, k: int, mod9: bool = True, device: int = -1, lookahead_bits: int = 0,
                 lookahead_semantics: int = 0):
        L = load()
        h = C.c_void_p()
        p = SieveParams(k, int(mod9), device, lookahead_bits, lookahead_semantics, 0)
        _check(L.collatz_sieve_build(C.byref(p), C.byref(h)))
        self._h = h
        self.k = k
        self.mod9 = mod9
        self.lookahead_bits = lookahead_bits
        self.lookahead_semantics = lookahead_semantics

    def close(self):
        if self._h:
            load().collatz_sieve_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self) -> SieveInfo:
        inf = SieveInfo()
        _check(load().collatz_sieve_query(self._h, C.byref(inf)))
        return inf

    def survivors(self):
        n = self.info().survivors
        r = np.zeros(n, np.uint64)
        num = np.zeros(n, np.uint64)
        f = np.zeros(n, np.uint8)
        P = C.POINTER
        _check(load().collatz_sieve_copy(self._h, r.ctypes.data_as(P(C.c_uint64)),
                                         num.ctypes.data_as(P(C.c_uint64)),
                                         f.ctypes.data_as(P(C.c_uint8)), n))
        return r, num, f

    def verify_range(self, lo: int, hi: int, chunk_log2: int = 24, limbs: int = 2,
                     cutoff_log2: int = 187, form: int = FORM_PLAIN, overflow_cap: int = 4096,
                     stream=None, noalloc: bool = False) -> RangeOutput:
        """collatz_verify_range over [lo, hi); with noalloc, collatz_verify_range_noalloc
        (the plain form on the library's fixed device array)."""
        n = ((hi - 1) >> chunk_log2) - (lo >> chunk_log2) + 1
        chunks = np.zeros(n, CHUNK_DTYPE)
        ovf = np.zeros(2 * max(overflow_cap, 1), np.uint64)
        res = RangeResult()
        res.chunks = chunks.ctypes.data_as(C.POINTER(Chunk))
        res.chunks_cap = n
        res.overflow_starts = ovf.ctypes.data_as(C.POINTER(C.c_uint64))
        res.overflow_cap = overflow_cap
        opts = VerifyOpts(chunk_log2, limbs, cutoff_log2, form)
        fn = load().collatz_verify_range_noalloc if noalloc else load().collatz_verify_range
        _check(fn(self._h, u128_pair(lo), u128_pair(hi), C.byref(opts), stream, C.byref(res)))
        total = np.frombuffer(bytearray(res.total), CHUNK_DTYPE)[0]
        nov = min(res.overflow_written, overflow_cap)
        starts = [int(ovf[2 * i]) | (int(ovf[2 * i + 1]) << 64) for i in range(nov)]
        return RangeOutput(total, chunks, starts, res.escalated, res.device_ms, res.checks,
                           res.overflow_written)

    def verify_batch(self, intervals, chunk_log2: int = 24, limbs: int = 2,
                     cutoff_log2: int = 187, use_graph: bool = False, tile_residues: int = 0,
                     waves_per_cu: int = 0, checks_per_refill: int = 0, overflow_cap: int = 0,
                     stream=None, spot_checks: int = 0):
        """collatz_verify_batch over [(lo, hi), ...]; one RangeOutput per interval.  The
        chunk records of all intervals share one array (each output holds a view).  With
        spot_checks, that many chunk records are re-run by the plain kernel and compared."""
        m = len(intervals)
        iv = np.array([(lo & MASK64, lo >> 64, hi & MASK64, hi >> 64) for lo, hi in intervals],
                      dtype=np.uint64).reshape(m, 4)
        nch = np.array([((hi - 1) >> chunk_log2) - (lo >> chunk_log2) + 1
                        for lo, hi in intervals], dtype=np.int64)
        offs = np.zeros(m + 1, np.int64)
        np.cumsum(nch, out=offs[1:])
        chunks = np.zeros(int(offs[-1]), CHUNK_DTYPE)
        res = np.zeros(m, RESULT_DTYPE)
        res["chunks"] = chunks.ctypes.data + offs[:-1].astype(np.uint64) * CHUNK_DTYPE.itemsize
        res["chunks_cap"] = nch
        ovf = None
        if overflow_cap:
            ovf = np.zeros((m, 2 * overflow_cap), np.uint64)
            res["overflow_starts"] = ovf.ctypes.data + np.arange(m, dtype=np.uint64) * (
                16 * overflow_cap)
            res["overflow_cap"] = overflow_cap
        opts = BatchOpts(chunk_log2, limbs, cutoff_log2, int(use_graph), tile_residues,
                         waves_per_cu, checks_per_refill, spot_checks)
        _check(load().collatz_verify_batch(self._h, m, iv.ctypes.data_as(C.POINTER(Interval)),
                                           C.byref(opts), stream,
                                           res.ctypes.data_as(C.POINTER(RangeResult))))
        totals = res["total"]
        esc = res["escalated"].tolist()
        dms = res["device_ms"].tolist()
        chk = res["checks"].tolist()
        ow = res["overflow_written"].tolist()
        sc = res["spot_checks"].tolist()
        o = offs.tolist()
        if ovf is None:
            starts = [[] for _ in range(m)]
        else:
            starts = [[int(ovf[i, 2 * j]) | (int(ovf[i, 2 * j + 1]) << 64)
                       for j in range(min(ow[i], overflow_cap))] for i in range(m)]
        return [RangeOutput(t, chunks[a:b], st, e, d, c, w, k)
                for t, a, b, st, e, d, c, w, k in zip(totals, o[:-1], o[1:], starts, esc, dms,
                                                      chk, ow, sc)]


class Server:
    """Served form: submit ranges at any time; a worker thread batches the pending work of all
    admitted requests into back-to-back persistent launches (collatz_server_*)."""

    def __init__(self, sieve: Sieve, chunk_log2=24, cutoff_log2=187, max_device_bytes=0,
                 slice_candidates=0, max_inflight=0):
        h = C.c_void_p()
        opts = ServerOpts(chunk_log2, cutoff_log2, max_device_bytes, slice_candidates,
                          max_inflight, 0)
        _check(load().collatz_server_create(sieve.handle, C.byref(opts), C.byref(h)))
        self._h = h
        self.sieve = sieve
        self.chunk_log2 = chunk_log2
        self._meta = {}

    def submit(self, lo: int, hi: int) -> int:
        iv = Interval((C.c_uint64 * 2)(lo & MASK64, lo >> 64), (C.c_uint64 * 2)(hi & MASK64, hi >> 64))
        t = load().collatz_server_submit(self._h, C.byref(iv))
        if t < 0:
            _check(int(t))
        self._meta[t] = (lo, hi)
        return t

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        lo, hi = self._meta.pop(ticket)
        n = ((hi - 1) >> self.chunk_log2) - (lo >> self.chunk_log2) + 1
        chunks = np.zeros(n, CHUNK_DTYPE)
        out = RequestResult()
        _check(load().collatz_server_wait(self._h, ticket, timeout_ms, C.byref(out),
                                          chunks.ctypes.data_as(C.POINTER(Chunk)), n))
        total = np.frombuffer(bytearray(out.total), CHUNK_DTYPE)[0]
        return total, chunks, (out.submit_ms, out.admit_ms, out.done_ms)

    def close(self):
        if getattr(self, "_h", None):
            load().collatz_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reference_tables(N=45, A=10, B=24, num_bv=16, semantics=SEM_CUDA, device=-1):
    """BV_i vectors (num_bv x 2^(B-5) u32 words) and mod 9 rows (9 x ceil(2^A/32))."""
    bv = np.zeros((num_bv, 1 << (B - 5)), np.uint32)
    m9 = np.zeros((9, max(1, (1 << A) // 32)), np.uint32)
    params = RefParams(N, A, B, num_bv, semantics, 100, device, 0)
    _check(load().collatz_reference_tables(C.byref(params), bv.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           m9.ctypes.data_as(C.POINTER(C.c_uint32))))
    return bv, m9


def reference_run(t, N=45, A=10, B=24, num_bv=16, semantics=SEM_CUDA, cutoff_log2=100,
                  overflow_cap=1024, device=-1, stream=None, spot_check_freq=0) -> RefResult:
    """One case line (base, sigfig, threepower, number, doubleeven, numevens).  With
    spot_check_freq, one tile of 256 leaves in every spot_check_freq is re-run by a second
    kernel and compared (RefResult.spot_checks)."""
    base, sigfig, tp, number, de, ne = t
    node = Node((C.c_uint64 * 2)(base & MASK64, base >> 64),
                (C.c_uint64 * 2)(number & MASK64, number >> 64), sigfig, tp, de, ne)
    ovf = np.zeros(2 * max(overflow_cap, 1), np.uint64)
    out = RefResult()
    out.overflow_starts = ovf.ctypes.data_as(C.POINTER(C.c_uint64))
    out.overflow_cap = overflow_cap
    params = RefParams(N, A, B, num_bv, semantics, cutoff_log2, device, spot_check_freq)
    _check(load().collatz_reference_run(C.byref(params), C.byref(node), stream, C.byref(out)))
    out._ovf = ovf
    return out
