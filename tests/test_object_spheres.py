"""Object bounding spheres (rsh_object_spheres; reference BoundingSphere.cpp:10-62).

The library's host function computes the smallest ball around each object's
vertices.  It is checked against:
  * brute force on small point sets: the smallest ball through every pair,
    triple and quadruple of points that holds all points (the minimal ball is
    determined by at most four of them);
  * the reference's own Miniball (Gaertner's header, built from the
    reference sources into oracle/_ref when they are present);
  * containment of every vertex of the test scenes, and being no larger than
    the generators' own (bounding-box centred) spheres;
  * the reference's radius rule (quirk Q12) in RSH_SPHERE_REFERENCE mode.
No device is used: the function is host code.
"""
import itertools

import numpy as np
import pytest

from rsh_amd import scenes
import rsh_amd
from oracle import pyoracle


def spheres_of_points(pts, mode=rsh_amd.SPHERE_MINIMAL):
    """One object made of the given points (as triangles, padded by repeats)."""
    pts = np.asarray(pts, np.float64)
    n = len(pts)
    m = -(-n // 3) * 3
    p = np.concatenate([pts, np.repeat(pts[-1:], m - n, 0)])
    pos = np.concatenate([p, np.ones((m, 1))], 1).astype(np.float32)
    return rsh_amd.object_spheres(pos, np.array([[0, m // 3 - 1]], np.uint32), mode)[0]


def circumsphere(s):
    """Smallest sphere through the points s (2-4 of them), float64."""
    s = np.asarray(s, np.float64)
    p0 = s[0]
    A = s[1:] - p0
    if len(s) == 2:
        return (s[0] + s[1]) / 2, np.linalg.norm(s[1] - s[0]) / 2
    # centre p0 + A^T y with 2 A A^T y = |A_i|^2 (centre in the affine hull)
    G = 2 * A @ A.T
    if abs(np.linalg.det(G)) < 1e-12 * max(1.0, np.abs(G).max()) ** len(A):
        return None
    y = np.linalg.solve(G, (A * A).sum(1))
    c = p0 + A.T @ y
    return c, np.linalg.norm(c - p0)


def brute_force(pts):
    pts = np.asarray(pts, np.float64)
    best = None
    for k in (2, 3, 4):
        for idx in itertools.combinations(range(len(pts)), k):
            cs = circumsphere(pts[list(idx)])
            if cs is None:
                continue
            c, r = cs
            if np.all(np.linalg.norm(pts - c, axis=1) <= r * (1 + 1e-9) + 1e-9):
                if best is None or r < best[1]:
                    best = (c, r)
    return best


@pytest.mark.parametrize("seed", range(12))
def test_small_sets_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 9))
    pts = rng.normal(0, 1, (n, 3)) * rng.uniform(0.1, 10) + rng.uniform(-50, 50, 3)
    out = spheres_of_points(pts)
    c, r = brute_force(pts)
    scale = np.abs(pts).max()
    assert np.allclose(out[:3], c, atol=1e-5 * scale), (out[:3], c)
    assert abs(out[3] - r) <= 1e-5 * scale, (out[3], r)


def test_special_point_sets():
    # one point, two points, collinear, duplicates, a regular tetrahedron
    assert np.allclose(spheres_of_points([[1, 2, 3]])[:4], [1, 2, 3, 0])
    assert np.allclose(spheres_of_points([[0, 0, 0], [2, 0, 0]])[:4], [1, 0, 0, 1])
    assert np.allclose(spheres_of_points([[0, 0, 0], [1, 0, 0], [4, 0, 0]])[:4], [2, 0, 0, 2], atol=1e-6)
    assert np.allclose(spheres_of_points([[0, 0, 0]] * 5 + [[0, 2, 0]] * 4)[:4], [0, 1, 0, 1], atol=1e-6)
    t = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64)
    assert np.allclose(spheres_of_points(t)[:4], [0, 0, 0, np.sqrt(3)], atol=1e-6)
    # a point inside the tetrahedron changes nothing
    assert np.allclose(spheres_of_points(np.concatenate([t, [[0.1, 0.2, 0.0]]]))[:4], [0, 0, 0, np.sqrt(3)],
                       atol=1e-6)


@pytest.mark.parametrize("make", [scenes.config0_scene, scenes.config1_scene])
def test_scene_objects_contained_and_no_larger(make):
    sc = make()
    rg = sc.object_spheres[:, 4:6].astype(np.uint32)
    out = rsh_amd.object_spheres(sc.tri_pos, rg)
    assert np.array_equal(out[:, 4:6], sc.object_spheres[:, 4:6])
    v = sc.tri_pos.reshape(-1, 3, 4)[:, :, :3]
    for k, (a, b) in enumerate(rg):
        pts = v[a:b + 1].reshape(-1, 3)
        d = np.sqrt(((pts - out[k, :3]) ** 2).sum(1, dtype=np.float32), dtype=np.float32)
        assert (d <= out[k, 3]).all()                                     # in float
        assert out[k, 3] <= sc.object_spheres[k, 3] * (1 + 1e-6)          # never larger


def test_reference_radius_rule():
    sc = scenes.config0_scene()
    rg = sc.object_spheres[:, 4:6].astype(np.uint32)
    mn = rsh_amd.object_spheres(sc.tri_pos, rg, rsh_amd.SPHERE_MINIMAL)
    rf = rsh_amd.object_spheres(sc.tri_pos, rg, rsh_amd.SPHERE_REFERENCE)
    assert np.array_equal(mn[:, :3], rf[:, :3])
    for k in range(len(rg)):
        r2 = np.float32(mn[k, 3]) ** 2
        expect = min(np.float32(1e5), max(np.float32(r2 * r2), mn[k, 3]))
        assert rf[k, 3] == pytest.approx(expect, rel=1e-5)
    # a unit-scale object keeps the true radius (r^4 < r), world-scale ones cap at 1e5
    small = spheres_of_points([[0, 0, 0], [0.5, 0, 0]], rsh_amd.SPHERE_REFERENCE)
    assert small[3] == pytest.approx(0.25, rel=1e-6)
    assert rf[0, 3] == 1e5


def test_against_reference_miniball():
    if pyoracle.ref_miniball(np.zeros((1, 3))) is None:
        pytest.skip("the reference's Miniball is built only where the reference sources are present")
    rng = np.random.default_rng(3)
    sc = scenes.config1_scene()
    rg = sc.object_spheres[:, 4:6].astype(np.uint32)
    v = sc.tri_pos.reshape(-1, 3, 4)[:, :, :3]
    sel = rng.choice(len(rg), 200, replace=False)
    out = rsh_amd.object_spheres(sc.tri_pos, rg[sel])
    for k, o in zip(sel, out):
        a, b = rg[k]
        pts = v[a:b + 1].reshape(-1, 3)
        c, r2 = pyoracle.ref_miniball(pts)
        scale = np.abs(pts).max()
        assert np.allclose(o[:3], c, atol=2e-5 * scale), (k, o[:3], c)
        assert o[3] == pytest.approx(np.sqrt(r2), abs=2e-5 * scale)
    for seed in range(20):
        pts = rng.normal(0, 1, (int(rng.integers(3, 200)), 3)).astype(np.float32) * 7 + 3
        c, r2 = pyoracle.ref_miniball(pts)
        o = spheres_of_points(pts)
        assert np.allclose(o[:3], c, atol=1e-4), (o[:3], c)
        assert o[3] == pytest.approx(np.sqrt(r2), abs=1e-4)
