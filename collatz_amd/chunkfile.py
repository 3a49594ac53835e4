"""Versioned per-chunk result file, format "collatz-chunks 1".

Text, one item per line (the same bytes as `collatz_run --range ... --out FILE` writes):

    collatz-chunks 1
    k <k>
    mod9 <0|1>
    lookahead <bits> <cuda|opencl>
    chunk_log2 <C>
    cutoff_log2 <c>
    range <lo> <hi>                  0x-prefixed lowercase hex
    chunks <n>
    fields chunk_index n_checked count stop_sum stop_xor max_exc max_start overflow_count hash flags
    chunk <fields>                   n lines, one per aligned chunk record
    total <fields>                   the merge of all records
    overflow <m>
    start <hex>                      m lines: the listed starts that reached 2^cutoff, in
                                     ascending order (one line per event)
    end <fnv>                        FNV-1a 32-bit hash of every byte before this line

Fields: counts in decimal, max_exc (192 bits) and max_start (128 bits) as 0x-prefixed
lowercase hex, hash as eight uppercase hex digits, flags in decimal.  A reader refuses
another format name or version and a file whose hash does not match.
"""
import numpy as np

from . import CHUNK_DTYPE

FORMAT = "collatz-chunks"
VERSION = 1
FIELDS = ("chunk_index", "n_checked", "count", "stop_sum", "stop_xor", "max_exc", "max_start",
          "overflow_count", "hash", "flags")
_M64 = (1 << 64) - 1


def _fnv1a(data: bytes) -> int:
    h = 2166136261
    for b in data:
        h = ((h ^ b) * 16777619) & 0xFFFFFFFF
    return h


def _words_int(ws) -> int:
    v = 0
    for i, w in enumerate(ws):
        v |= int(w) << (64 * i)
    return v


def _fields(c) -> str:
    return "%d %d %d %d %d 0x%x 0x%x %d %08X %d" % (
        int(c["chunk_index"]), int(c["n_checked"]), int(c["count"]), int(c["stop_sum"]),
        int(c["stop_xor"]), _words_int(c["max_exc"]), _words_int(c["max_start"]),
        int(c["overflow_count"]), int(c["hash"]), int(c["flags"]))


def format_chunks(*, k, chunk_log2, cutoff_log2, lo, hi, chunks, total, overflow_starts=(),
                  mod9=True, lookahead_bits=0, lookahead_semantics=0) -> bytes:
    """The file's bytes for one verified range (records as collatz_verify_range gives them)."""
    sem = "opencl" if lookahead_semantics == 2 else "cuda"
    lines = [f"{FORMAT} {VERSION}", f"k {k}", f"mod9 {int(bool(mod9))}",
             f"lookahead {lookahead_bits} {sem}", f"chunk_log2 {chunk_log2}",
             f"cutoff_log2 {cutoff_log2}", "range 0x%x 0x%x" % (lo, hi),
             f"chunks {len(chunks)}", "fields " + " ".join(FIELDS)]
    lines += ["chunk " + _fields(c) for c in chunks]
    lines.append("total " + _fields(total))
    lines.append(f"overflow {len(overflow_starts)}")
    lines += ["start 0x%x" % s for s in sorted(overflow_starts)]
    body = ("\n".join(lines) + "\n").encode()
    return body + ("end %08x\n" % _fnv1a(body)).encode()


def write_chunks(path, **kw) -> None:
    with open(path, "wb") as f:
        f.write(format_chunks(**kw))


def _parse_fields(parts):
    rec = np.zeros((), CHUNK_DTYPE)
    (rec["chunk_index"], rec["n_checked"], rec["count"], rec["stop_sum"],
     rec["stop_xor"]) = (int(x) for x in parts[:5])
    mx = int(parts[5], 16)
    rec["max_exc"] = [(mx >> (64 * i)) & _M64 for i in range(3)]
    ms = int(parts[6], 16)
    rec["max_start"] = [(ms >> (64 * i)) & _M64 for i in range(2)]
    rec["overflow_count"] = int(parts[7])
    rec["hash"] = int(parts[8], 16)
    rec["flags"] = int(parts[9])
    return rec


def parse_chunks(data: bytes) -> dict:
    """Parses a file's bytes; returns the header values, `chunks` (a CHUNK_DTYPE array),
    `total` and `overflow_starts`."""
    if not data.endswith(b"\n"):
        raise ValueError("truncated chunk file")
    cut = data.rstrip(b"\n").rfind(b"\n") + 1
    last = data[cut:].decode().split()
    if len(last) != 2 or last[0] != "end" or int(last[1], 16) != _fnv1a(data[:cut]):
        raise ValueError("chunk file hash mismatch or missing end line")
    lines = data[:cut].decode().splitlines()
    head = lines[0].split()
    if head != [FORMAT, str(VERSION)]:
        raise ValueError(f"not a {FORMAT} {VERSION} file: {lines[0]!r}")
    out = {"overflow_starts": []}
    recs = []
    for line in lines[1:]:
        key, *parts = line.split()
        if key == "chunk":
            recs.append(_parse_fields(parts))
        elif key == "total":
            out["total"] = _parse_fields(parts)
        elif key == "start":
            out["overflow_starts"].append(int(parts[0], 16))
        elif key == "range":
            out["lo"], out["hi"] = int(parts[0], 16), int(parts[1], 16)
        elif key == "lookahead":
            out["lookahead_bits"] = int(parts[0])
            out["lookahead_semantics"] = 2 if parts[1] == "opencl" else 0
        elif key == "fields":
            if tuple(parts) != FIELDS:
                raise ValueError("unexpected field list")
        elif key == "chunks":
            out["nchunks"] = int(parts[0])
        elif key in ("k", "mod9", "chunk_log2", "cutoff_log2", "overflow"):
            out[key] = int(parts[0])
        else:
            raise ValueError(f"unknown line {line!r}")
    out["chunks"] = np.array(recs, CHUNK_DTYPE) if recs else np.zeros(0, CHUNK_DTYPE)
    if len(recs) != out.get("nchunks") or "total" not in out:
        raise ValueError("chunk count mismatch or missing total")
    if len(out["overflow_starts"]) != out.get("overflow"):
        raise ValueError("overflow count mismatch")
    return out


def read_chunks(path) -> dict:
    with open(path, "rb") as f:
        return parse_chunks(f.read())
