"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py).

The oracle must reproduce every committed vector bit for bit (a change to the
oracle that moves any key, sphere, hit or pixel fails here), and a few facts
of the vectors are checked against closed forms independent of the oracle.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def assert_same(a, b):
    assert set(a) == set(b)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)   # NaN compares equal to NaN here


def test_keys_reproduce(oracle):
    assert_same(make_golden.make_keys(), load("keys.npz"))


def test_geometry_reproduces(oracle):
    assert_same(make_golden.make_geometry(), load("geometry.npz"))


@pytest.mark.parametrize("name", sorted(make_golden.FRAMES))
def test_frames_reproduce(oracle, name):
    assert_same(make_golden.make_frame(name), load("frame_%s.npz" % name))


def test_key_vectors_closed_forms():
    k = load("keys.npz")
    d = k["d"]
    # RayTracer.cu:147-168: bits 30.. = light + 1; azimuth bin of atan2(x, x)
    # is 11 for x > 0, 7 for x = +0 and 0 for x < 0 or x = -0.
    az = np.where(d[:, 0] > 0, 11, np.where((d[:, 0] == 0) & ~np.signbit(d[:, 0]), 7, 0))
    for l in range(3):
        s = k["shadow_ref"][l].astype(np.uint64)
        assert np.all((s >> 30) == l + 1)
        assert np.array_equal((s >> 15) & 0x7FFF, az)
    pol = (k["shadow_ref"][0] & 0x7FFF) - 1
    assert pol[0] == 0 and pol[1] == 15          # z = 1 -> acos 0; z = -1 -> floor(5 pi) = 15
    assert np.all(np.diff(pol[np.argsort(-d[:, 2], kind="stable")].astype(np.int64)) >= 0)  # monotone in -z
    # octahedral keys: type in the top 2 of 29 bits
    assert np.all((k["oct"][:, :, :] >> 27) == np.arange(4)[None, :, None])


def test_geometry_vectors_closed_forms():
    g = load("geometry.npz")
    s = g["tri_spheres"]
    np.testing.assert_allclose(s[0, :3], [5, 5 / np.sqrt(3), 0], rtol=1e-6)       # equilateral circumcentre
    np.testing.assert_allclose(s[0, 3], 10 / np.sqrt(3), rtol=1e-6)
    np.testing.assert_allclose(s[1], [2, 1.5, 0, 2.5], rtol=1e-6)                   # right: hypotenuse midpoint
    np.testing.assert_allclose(s[2], [1, 1, 1, np.sqrt(3)], rtol=1e-6)              # collinear: longest edge
    assert s[4, 3] == 0
    t = g["t_hit"]
    # ray from (1,1,0) to the centroid of the z = 10 triangle: t = distance
    np.testing.assert_allclose(t[0], np.linalg.norm(np.array([10 / 3, 10 / 3, 10]) - [1, 1, 0]), rtol=1e-6)
    assert t[9] == -1.0                        # parallel ray
    assert t[10] < 0                           # triangle behind the origin
    assert np.all(g["sphere_node"][:5] == 1)   # HALF_PI spread always passes


def test_frame_vectors_are_consistent():
    f = load("frame_c0_64.npz")
    assert f["pixels"].shape == (64, 64)
    assert list(f["pass_kind"]) == [0, 1, 2]
    assert f["pass_rays"][0] == 64 * 64           # every primary ray is live
    hits = f["pass0_hits"]
    miss = hits == np.uint64(0xFFFFFFFFFFFFFFFF)
    assert not miss.all()
    tri = (hits[~miss] & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert tri.max() < 2060                       # 12 room + 2 x 1024 sphere triangles
