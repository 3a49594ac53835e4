"""Shared helpers for the GPU tests."""
import numpy as np

FIELDS = ["chunk_index", "n_checked", "count", "stop_sum", "stop_xor", "max_exc", "max_start",
          "overflow_count", "hash", "flags"]


def assert_chunks_equal(got, want, tag=""):
    assert len(got) == len(want), (tag, len(got), len(want))
    for f in FIELDS:
        a, b = got[f], want[f]
        if not np.array_equal(a, b):
            bad = np.nonzero(np.any((a != b).reshape(len(a), -1), axis=1))[0][:5]
            raise AssertionError(f"{tag}: field {f} differs at chunks {bad.tolist()}: "
                                 f"{a[bad].tolist()} vs {b[bad].tolist()}")


def assert_total_equal(got, want, tag=""):
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), (tag, f, got[f], want[f])
