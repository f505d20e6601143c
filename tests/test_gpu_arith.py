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
