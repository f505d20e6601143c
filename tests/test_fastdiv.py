"""The multiply-high division used by the traversal's item counting.

ray-tracer_amd/csrc/rsh_device.hpp (div_hi / div_hi_magic / div_hi_ok) divides
by a fixed c as the high word of x * ceil(2^32 / c) when the largest numerator
satisfies x <= (2^32 - 1) / (c - 1) (checked once on the host; the kernels
divide plainly otherwise).  The bound is checked here against exact integer
division at its edge and on random numerators, for every chunk size the
engine uses and beyond, and the multiply-high alone is shown to fail for some
numerators past it.
"""
import numpy as np


def div_by(c):
    m = 0 if c <= 1 else (2 ** 32 + c - 1) // c
    lim = 0 if c <= 1 else (2 ** 32 - 1) // (c - 1)
    return c, m, lim


def div_q(x, d):
    c, m, lim = d
    if c == 1:
        return x
    return (x * m) >> 32 if x <= lim else x // c


def test_exact_at_the_bound_and_on_samples():
    rng = np.random.default_rng(7)
    for c in list(range(1, 257)) + [1000, 4096, 65535, 1 << 20]:
        d = div_by(c)
        lim = min(d[2], 2 ** 32 - 1) if c > 1 else 2 ** 32 - 1
        xs = {0, 1, c - 1, c, c + 1, lim, lim - 1, 2 ** 32 - 1} | set(int(v) for v in rng.integers(0, 2 ** 32, 200))
        xs |= set(int(v) for v in rng.integers(max(0, lim - 5000), lim + 1, 200))
        for x in xs:
            if 0 <= x < 2 ** 32:
                assert div_q(x, d) == x // c, (c, x)


def test_multiply_high_alone_fails_beyond_some_bound():
    # without the fallback the high word is wrong for some large x: the
    # bound is needed (c = 7: e = 4, the first failure is near 2^32 / 4)
    c, m, _ = div_by(7)
    bad = [x for x in range(2 ** 32 - 7 * 2000, 2 ** 32) if (x * m) >> 32 != x // c]
    assert bad
