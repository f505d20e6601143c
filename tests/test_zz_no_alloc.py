"""rsh_sort_pairs_noalloc (the sort stage with its scratch in a file-scope
device array, for callers that cannot allocate) against rsh_sort_pairs, the
allocating entry point it mirrors: bit-exact keys and values on two calls in
a row, the second smaller, and sizes beyond the maximum rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 21


def noalloc(rsh):
    L = rsh.lib()
    f = L.rsh_sort_pairs_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def reference_sort(rsh, keys, vals, bits):
    L = rsh.lib()
    L.rsh_compat_set_exit_on_error(0)
    L.rsh_compat_reset_status()
    n = len(keys)
    dk, dv = rsh.DeviceArray.from_host(keys), rsh.DeviceArray.from_host(vals)
    ok_, ov = rsh.DeviceArray((n,), np.uint32), rsh.DeviceArray((n,), np.uint32)
    L.rsh_sort_pairs(dk.ptr, dv.ptr, n, bits, ok_.ptr, ov.ptr)
    assert L.rsh_compat_status() == 0, L.rsh_compat_error()
    return ok_.download(), ov.download()


@pytest.mark.parametrize("bits", [27, 32])
def test_noalloc_sort_matches_allocating_sort(rsh, bits):
    f = noalloc(rsh)
    L = rsh.lib()
    rng = np.random.default_rng(bits)
    for n in (MAX_PAIRS, 300_001):          # two calls in a row, the second smaller
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.permutation(n).astype(np.uint32)
        dk, dv = rsh.DeviceArray.from_host(keys), rsh.DeviceArray.from_host(vals)
        ok_, ov = rsh.DeviceArray((n,), np.uint32), rsh.DeviceArray((n,), np.uint32)
        assert f(dk.ptr, dv.ptr, n, bits, ok_.ptr, ov.ptr, None) == 0
        assert L.rsh_device_synchronize() == 0
        want_k, want_v = reference_sort(rsh, keys, vals, bits)
        np.testing.assert_array_equal(ok_.download(), want_k)
        np.testing.assert_array_equal(ov.download(), want_v)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(want_v, vals[order])


def test_noalloc_sort_in_place_and_limits(rsh):
    f = noalloc(rsh)
    L = rsh.lib()
    rng = np.random.default_rng(5)
    n = 77_777
    keys = rng.integers(0, 1 << 27, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = rsh.DeviceArray.from_host(keys), rsh.DeviceArray.from_host(vals)
    assert f(dk.ptr, dv.ptr, n, 27, dk.ptr, dv.ptr, None) == 0       # outputs alias inputs, 3 digit passes
    assert L.rsh_device_synchronize() == 0
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(dk.download(), keys[order])
    np.testing.assert_array_equal(dv.download(), vals[order])
    # beyond the maximum: RSH_CAPACITY before anything runs; bad bits: RSH_INVALID_ARG
    assert f(dk.ptr, dv.ptr, MAX_PAIRS + 1, 27, dk.ptr, dv.ptr, None) == 2
    assert f(dk.ptr, dv.ptr, n, 0, dk.ptr, dv.ptr, None) == 1
    assert f(dk.ptr, dv.ptr, n, 33, dk.ptr, dv.ptr, None) == 1
