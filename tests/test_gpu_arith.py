"""Device arithmetic the traversal relies on for bit-exact hit records.

ray_triangle (RayTracer.cu:257-280) takes a correctly rounded reciprocal of
the Moller-Trumbore determinant; the kernels compute it as one Newton step on
the hardware estimate for exponents 2..252 and by IEEE division elsewhere.
The library's self-check compares the two on every float of that range.
"""
import pytest

pytestmark = pytest.mark.gpu


def test_newton_reciprocal_is_correctly_rounded(rsh):
    tested, bad = rsh.selfcheck_rcp(0)
    assert tested == 2 * 251 * (1 << 23)      # both signs, exponent fields 2..252, every mantissa
    assert bad == 0


def test_host_wait_gives_up_at_its_deadline(rsh):
    """A pass size that never arrives ends the host wait at its deadline with
    RSH_INTERNAL (4) instead of spinning for ever; the stream is held busy
    from the host (a host function that sleeps), so no kernel is involved."""
    st, ms = rsh.selfcheck_wait_deadline(100, 1500)
    assert st == 4, st
    assert 100 <= ms < 1400, ms
    # a stream that drains without publishing the word: an error, at once
    st, ms = rsh.selfcheck_wait_deadline(5000, 50)
    assert st == 3, st
    assert ms < 2000, ms


def test_wait_timeout_argument_checked(rsh):
    e = rsh.Engine(0)
    try:
        assert rsh.lib().rsh_engine_set_wait_timeout(e.h, 0) == 1
        assert rsh.lib().rsh_engine_set_wait_timeout(e.h, 250) == 0
    finally:
        e.close()
