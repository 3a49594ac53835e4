"""The failure paths of the kernels' own bounds, forced from the host: with
COLLATZ_CHECK_LIMIT set to a few checks, candidates that need more are abandoned, and with
COLLATZ_LOOP_LIMIT set to one iteration the kernels abandon their work loops; every
kernel still exits normally and each form reports the failure instead of a result.  Also the
served form's isolation of such a failure to the request it belongs to, and the lifetime
rules of its tickets."""
import os

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def tiny_check_limit():
    os.environ["COLLATZ_CHECK_LIMIT"] = "2"
    yield
    del os.environ["COLLATZ_CHECK_LIMIT"]


def test_check_bound_failure_is_reported(tiny_check_limit):
    import collatz_amd as cz
    sv = cz.Sieve(20)
    lo, hi = (1 << 40) + 1, (1 << 40) + (1 << 22)
    for form in (cz.FORM_PLAIN, cz.FORM_PERSISTENT):
        with pytest.raises(cz.CollatzError) as e:
            sv.verify_range(lo, hi, 20, form=form)
        assert e.value.code == cz.EDEVICE and "did not descend" in str(e.value)
    with pytest.raises(cz.CollatzError):
        sv.verify_batch([(lo, hi)], chunk_log2=20, use_graph=True)
    with pytest.raises(cz.CollatzError):
        cz.reference_run((3, 2, 2, 9, 1, 0), 36, 10)
    srv = cz.Server(sv, chunk_log2=20)
    t = srv.submit(lo, hi)
    with pytest.raises(cz.CollatzError):
        srv.wait(t, 60000)
    srv.close()


def test_loop_bound_failure_is_reported():
    """The work-loop bound of the kernels (iterations per tile of the persistent kernels and
    of their take loop per refill, iterations per wave share of the plain kernel), forced to
    one iteration with COLLATZ_LOOP_LIMIT: waves abandon their work and exit normally, and
    the plain, no-allocation, persistent, batched, served and reference-mode forms fail
    with the bound named."""
    import collatz_amd as cz
    os.environ["COLLATZ_LOOP_LIMIT"] = "1"
    try:
        sv = cz.Sieve(20)
        sla = cz.Sieve(20, lookahead_bits=16)
        lo, hi = (1 << 40) + 1, (1 << 40) + (1 << 24)
        for s in (sv, sla):
            for form in (cz.FORM_PERSISTENT, cz.FORM_PLAIN):
                with pytest.raises(cz.CollatzError) as e:
                    s.verify_range(lo, hi, 20, form=form)
                assert e.value.code == cz.EDEVICE and "iteration bound" in str(e.value)
            with pytest.raises(cz.CollatzError) as e:
                s.verify_range(lo, hi, 20, form=cz.FORM_PLAIN, noalloc=True)
            assert e.value.code == cz.EDEVICE and "iteration bound" in str(e.value)
            with pytest.raises(cz.CollatzError) as e:
                s.verify_batch([(lo, hi), (lo << 8, (lo << 8) + (1 << 22))], chunk_log2=20,
                               use_graph=True)
            assert "iteration bound" in str(e.value)
        with pytest.raises(cz.CollatzError) as e:
            cz.reference_run((3, 2, 2, 9, 1, 0), 36, 10)
        assert "iteration bound" in str(e.value)
        srv = cz.Server(sv, chunk_log2=20)
        t = srv.submit(lo, hi)
        with pytest.raises(cz.CollatzError) as e:
            srv.wait(t, 60000)
        assert "iteration bound" in str(e.value)
        srv.close()
    finally:
        del os.environ["COLLATZ_LOOP_LIMIT"]
    # The bound is read per call: without the variable the same calls succeed.
    assert int(sv.verify_range(lo, hi, 20, form=cz.FORM_PLAIN).total["count"]) > 0
    assert int(sv.verify_range(lo, hi, 20, form=cz.FORM_PLAIN, noalloc=True)
               .total["count"]) > 0


def test_default_bound_after_the_hook():
    """Without the variables the same calls succeed (the launch sequences captured under the
    lowered bounds are not replayed)."""
    import collatz_amd as cz
    sv = cz.Sieve(20)
    lo, hi = (1 << 40) + 1, (1 << 40) + (1 << 22)
    a = sv.verify_range(lo, hi, 20, form=cz.FORM_PERSISTENT)
    b = sv.verify_range(lo, hi, 20, form=cz.FORM_PLAIN)
    assert int(a.total["count"]) == int(b.total["count"]) > 0
    assert cz.reference_run((3, 2, 2, 9, 1, 0), 36, 10).candidates > 0


# A start near 2^40 that needs 182 T-steps (20 checks) to descend, and a 4096-wide window
# whose 28 candidates all descend within 34 steps (both found with the oracle, k = 20).
LONG_START = 1099511984379
SHORT_LO, SHORT_HI = 1099511697408, 1099511697408 + 4096


def test_served_stall_fails_only_its_request(oracle):
    """With the check bound at 20 checks (40 T-steps in the persistent kernel), a request
    holding the 182-step trajectory fails and names its chunk, while a request admitted with
    it into the same launch (a device budget smaller than the first, large request holds both
    back until that one has finished) returns the oracle's records; the server keeps serving
    afterwards."""
    import collatz_amd as cz

    from gpu_util import assert_chunks_equal, assert_total_equal
    o20 = oracle.sieve(20)
    want = oracle.verify_range(o20, SHORT_LO, SHORT_HI, 20)
    assert int(want[1]["count"]) == 28
    os.environ["COLLATZ_CHECK_LIMIT"] = "20"
    try:
        sv = cz.Sieve(20)
        srv = cz.Server(sv, chunk_log2=20, max_device_bytes=1 << 20)
        big = srv.submit(1 << 40, (1 << 40) + (1 << 34))
        bad = srv.submit(LONG_START, LONG_START + 1)
        good = srv.submit(SHORT_LO, SHORT_HI)
        for t in (big, bad):
            with pytest.raises(cz.CollatzError) as e:
                srv.wait(t, 120000)
            assert e.value.code == cz.EDEVICE and "did not descend" in str(e.value)
        assert f"chunk {LONG_START >> 20}" in str(e.value)
        total, chunks, _ = srv.wait(good, 120000)
        assert_chunks_equal(chunks, want[0], "short request")
        assert_total_equal(total, want[1], "short request")
        again = srv.submit(SHORT_LO, SHORT_HI)
        assert_chunks_equal(srv.wait(again, 120000)[1], want[0], "later request")
        srv.close()
    finally:
        del os.environ["COLLATZ_CHECK_LIMIT"]


def test_server_wait_twice_and_close_with_waiters():
    """A ticket's result is handed out once (a second wait on it fails cleanly, as does a
    wait on a ticket never issued), and a server closed while a thread waits on an
    unfinished request finishes it and lets that waiter return."""
    import ctypes as C
    import threading
    import time

    import collatz_amd as cz
    L = cz.load()
    sv = cz.Sieve(24)
    srv = cz.Server(sv, chunk_log2=30)
    t = srv.submit(1 << 40, (1 << 40) + (1 << 24))
    out = cz.RequestResult()
    assert L.collatz_server_wait(srv._h, t + 1000, 1000.0, C.byref(out), None, 0) == cz.EINVAL
    assert L.collatz_server_wait(srv._h, t, 60000.0, C.byref(out), None, 0) == cz.OK
    assert out.total.count > 0
    assert L.collatz_server_wait(srv._h, t, 1000.0, C.byref(out), None, 0) == cz.EINVAL
    # About 6.5e9 candidates (327,680 rows): still running when close() is called.
    t2 = srv.submit(1 << 44, (1 << 44) + (1 << 40))
    got = []

    def waiter():
        r = cz.RequestResult()
        got.append((L.collatz_server_wait(srv._h, t2, -1.0, C.byref(r), None, 0),
                    r.total.count))

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.02)
    srv.close()
    th.join(60)
    assert not th.is_alive() and got and got[0][0] == cz.OK and got[0][1] > 0
