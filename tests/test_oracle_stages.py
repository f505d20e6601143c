"""Trim / compress / sort / decompress restatements against textbook algorithms.

Reference stages: RayTracer.cu:2276 RayTrimmingWrapper, :2326
RayCompressionWrapper, :2385 RaySortingWrapper (CUB stable LSD radix sort),
:2411 RayDecompressionWrapper.
"""
import ctypes as C

import numpy as np
import pytest


def run_trim(oracle, keys, values, flags, W, H, L):
    S = W * H * L
    scan = np.zeros(S, np.uint32)
    tk = np.zeros(S, np.uint32)
    tv = np.zeros(S, np.uint32)
    total = C.c_uint32(0)
    oracle.call("oracle_RayTrimmingWrapper", keys, values, W, H, L, flags, scan, tk, tv, C.byref(total))
    return tk[:total.value], tv[:total.value], scan


@pytest.mark.parametrize("dead", [0.0, 0.3, 0.97, 1.0])
def test_trim_is_stable_copy_if(oracle, dead):
    rng = np.random.default_rng(11)
    W, H, L = 37, 29, 2
    S = W * H * L
    keys = rng.integers(0, 1 << 29, S).astype(np.uint32)
    values = np.arange(S, dtype=np.uint32)
    flags = (rng.random(S) < dead).astype(np.uint32)
    tk, tv, scan = run_trim(oracle, keys, values, flags, W, H, L)
    live = flags == 0
    np.testing.assert_array_equal(tk, keys[live])
    np.testing.assert_array_equal(tv, values[live])
    np.testing.assert_array_equal(scan, np.cumsum(flags))


def compress(oracle, tk):
    N = tk.size
    flags = np.zeros(N, np.uint32)
    scan = np.zeros(N, np.uint32)
    bases, sizes, ck, cv = (np.zeros(N, np.uint32) for _ in range(4))
    total = C.c_uint32(0)
    oracle.call("oracle_RayCompressionWrapper", tk, np.zeros(N, np.uint32), N, flags, scan, bases, sizes, ck, cv,
                C.byref(total))
    c = total.value
    return bases[:c], sizes[:c], ck[:c], cv[:c]


def test_compress_is_run_length_encoding(oracle):
    rng = np.random.default_rng(12)
    tk = np.repeat(rng.integers(0, 50, 3000), rng.integers(1, 6, 3000)).astype(np.uint32)
    bases, sizes, ck, cv = compress(oracle, tk)
    starts = np.concatenate([[0], np.nonzero(tk[1:] != tk[:-1])[0] + 1])
    np.testing.assert_array_equal(bases, starts)
    np.testing.assert_array_equal(sizes, np.diff(np.concatenate([starts, [tk.size]])))
    np.testing.assert_array_equal(ck, tk[starts])
    np.testing.assert_array_equal(cv, np.arange(starts.size))
    assert sizes.sum() == tk.size


def test_sort_is_stable(oracle):
    rng = np.random.default_rng(13)
    n = 5000
    ck = rng.integers(0, 64, n).astype(np.uint32) << np.uint32(20)
    cv = np.arange(n, dtype=np.uint32)
    sk, sv = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    oracle.call("oracle_RaySortingWrapper", ck, cv, n, sk, sv)
    order = np.argsort(ck, kind="stable")
    np.testing.assert_array_equal(sk, ck[order])
    np.testing.assert_array_equal(sv, cv[order])


def test_compress_sort_decompress_is_stable_sort(oracle):
    """Chunks sorted stably then expanded == stable sort of the trimmed rays."""
    rng = np.random.default_rng(14)
    tk = rng.integers(0, 200, 20000).astype(np.uint32)
    tv = rng.permutation(20000).astype(np.uint32)
    N = tk.size
    bases, sizes, ck, cv = compress(oracle, tk)
    c = ck.size
    sk, sv = np.zeros(c, np.uint32), np.zeros(c, np.uint32)
    oracle.call("oracle_RaySortingWrapper", np.ascontiguousarray(ck), np.ascontiguousarray(cv), c, sk, sv)
    skel, scan = np.zeros(c, np.uint32), np.zeros(c, np.uint32)
    rk, rv = np.zeros(N, np.uint32), np.zeros(N, np.uint32)
    oracle.call("oracle_RayDecompressionWrapper", np.ascontiguousarray(bases), np.ascontiguousarray(sizes), sk, sv,
                tk, tv, skel, scan, c, rk, rv)
    order = np.argsort(tk, kind="stable")
    np.testing.assert_array_equal(rk, tk[order])
    np.testing.assert_array_equal(rv, tv[order])
