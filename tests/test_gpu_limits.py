"""The failure path of the kernels' own check bound, forced from the host: with
COLLATZ_CHECK_LIMIT set to a few checks, candidates that need more are abandoned (every
kernel still exits normally) and each form reports the failure instead of a result."""
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


def test_default_bound_after_the_hook():
    """Without the variable the same calls succeed (the launch sequence captured under the
    lowered bound is not replayed)."""
    import collatz_amd as cz
    sv = cz.Sieve(20)
    lo, hi = (1 << 40) + 1, (1 << 40) + (1 << 22)
    a = sv.verify_range(lo, hi, 20, form=cz.FORM_PERSISTENT)
    b = sv.verify_range(lo, hi, 20, form=cz.FORM_PLAIN)
    assert int(a.total["count"]) == int(b.total["count"]) > 0
    assert cz.reference_run((3, 2, 2, 9, 1, 0), 36, 10).candidates > 0
