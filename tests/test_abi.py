"""The C-ABI library loads and exports every entry point include/collatz.h declares.
No compute calls (this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "collatz.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(collatz_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("collatz_sieve_build", "collatz_verify_range", "collatz_verify_batch"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import collatz_amd as cz
    lib = cz.load()
    names = declared()
    assert names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(cz.EXPORTS)


def test_version_and_error_string():
    import collatz_amd as cz
    lib = cz.load()
    assert lib.collatz_version() == 2
    assert isinstance(lib.collatz_last_error(), bytes)


def test_null_arguments_are_rejected():
    import collatz_amd as cz
    lib = cz.load()
    assert lib.collatz_sieve_build(None, None) == cz.EINVAL
    assert lib.collatz_verify_range(None, None, None, None, None, None) == cz.EINVAL
    assert lib.collatz_verify_batch(None, 0, None, None, None, None) == cz.EINVAL
    assert lib.collatz_sieve_query(None, None) == cz.EINVAL
    assert lib.collatz_sieve_free(None) == cz.OK


def test_struct_layouts_match_header():
    import collatz_amd as cz
    assert ctypes.sizeof(cz.Chunk) == 11 * 8 + 8
    assert ctypes.sizeof(cz.SieveParams) == 24
    assert ctypes.sizeof(cz.Interval) == 32


def test_sieve_params_are_checked_before_the_device():
    """Bad depths and look-ahead settings are refused with EINVAL before any HIP call."""
    import collatz_amd as cz
    lib = cz.load()
    h = ctypes.c_void_p()
    for k, bits, sem in ((7, 0, 0), (39, 0, 0), (32, 7, 0), (32, 25, 0), (32, 16, cz.SEM_RUST)):
        p = cz.SieveParams(k, 1, -1, bits, sem, 0)
        assert lib.collatz_sieve_build(ctypes.byref(p), ctypes.byref(h)) == cz.EINVAL
        assert b"k must" in lib.collatz_last_error() or b"lookahead" in lib.collatz_last_error()
    assert ctypes.sizeof(cz.SieveInfo) == 48


def test_reference_params_are_checked_before_the_device():
    """Reference mode refuses knobs outside the reference's domain (a candidate window
    longer than a bit vector, A > B) and outside a run's word widths (N > 64, N - A > 62,
    a main.rs cutoff above 2^180) with EINVAL before any HIP call."""
    import collatz_amd as cz
    lib = cz.load()
    node = cz.Node((ctypes.c_uint64 * 2)(3, 0), (ctypes.c_uint64 * 2)(9, 0), 2, 2, 1, 0)
    res = cz.RefResult()
    for N, A, B, sem, cut in ((30, 10, 8, 0, 100), (40, 13, 24, 0, 100), (30, 0, 24, 0, 100),
                              (70, 10, 24, 0, 100), (65, 2, 24, 0, 100),
                              (40, 6, 24, cz.SEM_RUST, 181)):
        p = cz.RefParams(N, A, B, 16, sem, cut, -1, 0)
        assert lib.collatz_reference_run(ctypes.byref(p), ctypes.byref(node), None,
                                         ctypes.byref(res)) == cz.EINVAL, (N, A, B)
    p = cz.RefParams(30, 10, 8, 16, 0, 100, -1, 0)
    lib.collatz_reference_run(ctypes.byref(p), ctypes.byref(node), None, ctypes.byref(res))
    assert b"must not exceed B" in lib.collatz_last_error()


def test_header_compiles_as_c_and_layouts_match(tmp_path):
    """include/collatz.h is plain C99: a C program built against it links the library,
    calls entry points that need no GPU, and sees the struct sizes the Python binding
    declares."""
    import shutil
    import subprocess
    import collatz_amd as cz
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        import pytest
        pytest.skip("no C compiler")
    src = tmp_path / "abi.c"
    src.write_text(r'''
#include <stdio.h>
#include "collatz.h"
int main(void) {
  collatz_sieve* sv = 0;
  collatz_sieve_params bad = {7, 1, -1, 0, COLLATZ_SEM_CUDA, 0};
  int rc = collatz_sieve_build(&bad, &sv);
  printf("%d %d %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", collatz_version(), rc,
         sizeof(collatz_sieve_params), sizeof(collatz_sieve_info), sizeof(collatz_chunk),
         sizeof(collatz_range_result), sizeof(collatz_verify_opts), sizeof(collatz_batch_opts),
         sizeof(collatz_ref_params), sizeof(collatz_ref_result), sizeof(collatz_server_opts));
  return 0;
}
''')
    exe = tmp_path / "abi"
    lib = os.path.join(ROOT, "collatz_amd")
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", lib, "-lcollatz_amd",
                    "-Wl,-rpath," + lib], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    v, rc, *sizes = (int(x) for x in out)
    assert v == cz.load().collatz_version() and rc == cz.EINVAL
    assert sizes == [ctypes.sizeof(t) for t in (cz.SieveParams, cz.SieveInfo, cz.Chunk,
                                                cz.RangeResult, cz.VerifyOpts, cz.BatchOpts,
                                                cz.RefParams, cz.RefResult, cz.ServerOpts)]
