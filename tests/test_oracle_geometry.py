"""Known answers and invariants of the oracle's geometric primitives.

RayTracer.cu:257-280 RayTriangleIntersection, :283-298 SphereNodeIntersection,
:301-376 CreateTriangleBoundingSphere, :379-446 cone/sphere unions.
"""
import ctypes as C

import numpy as np
import pytest

PF = C.POINTER(C.c_float)


def p(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(PF)


def mt(oracle, o, d, v0, v1, v2):
    v0 = np.asarray(v0, np.float32)
    e1 = np.asarray(v1, np.float32) - v0
    e2 = np.asarray(v2, np.float32) - v0
    arrs = [p(x) for x in (o, d, v0, e1, e2)]
    return oracle.lib().oracle_ray_triangle(*[a[1] for a in arrs])


TRI = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


def test_moller_trumbore_known_answers(oracle):
    down = [0, 0, -1]
    assert mt(oracle, [0.25, 0.25, 1], down, *TRI) == pytest.approx(1.0, abs=1e-7)
    assert mt(oracle, [1 / 3, 1 / 3, 2], down, *TRI) == pytest.approx(2.0, abs=1e-6)      # centroid
    assert mt(oracle, [0.5, 0.0, 1], down, *TRI) == pytest.approx(1.0, abs=1e-7)          # on edge v = 0
    assert mt(oracle, [0.0, 0.0, 1], down, *TRI) == pytest.approx(1.0, abs=1e-7)          # vertex
    assert mt(oracle, [0.6, 0.6, 1], down, *TRI) == -1.0                                  # u + v > 1
    assert mt(oracle, [-0.1, 0.5, 1], down, *TRI) == -1.0                                 # u < 0
    assert mt(oracle, [0.25, 0.25, -1], down, *TRI) == pytest.approx(-1.0, abs=1e-7)      # behind the origin
    t = mt(oracle, [0.25, 0.25, 1], [1, 0, 0], *TRI)                                      # parallel: det = 0
    assert not (t > 0.01)
    # t scales with the direction length
    assert mt(oracle, [0.25, 0.25, 1], [0, 0, -2], *TRI) == pytest.approx(0.5, abs=1e-7)


def tri_sphere(oracle, a, b, c):
    out = np.zeros(4, np.float32)
    arrs = [p(x) for x in (a, b, c)]
    oracle.lib().oracle_triangle_sphere(*[x[1] for x in arrs], out.ctypes.data_as(PF))
    return out


def test_triangle_sphere_closed_forms(oracle):
    s = 10.0   # large enough that 2*|AB x AC|^2 > 0.01 (non-degenerate branch)
    # right triangle: centre = hypotenuse midpoint
    r = tri_sphere(oracle, [0, 0, 0], [s, 0, 0], [0, s, 0])
    np.testing.assert_allclose(r[:3], [s / 2, s / 2, 0], atol=1e-5)
    assert r[3] == pytest.approx(s * np.sqrt(2) / 2, rel=1e-6)
    # equilateral: circumcentre = centroid, radius = side / sqrt(3)
    a, b, c = np.array([0, 0, 0.0]), np.array([s, 0, 0.0]), np.array([s / 2, s * np.sqrt(3) / 2, 0.0])
    r = tri_sphere(oracle, a, b, c)
    np.testing.assert_allclose(r[:3], (a + b + c) / 3, atol=1e-4)
    assert r[3] == pytest.approx(s / np.sqrt(3), rel=1e-5)
    # obtuse: the longest edge's midpoint
    r = tri_sphere(oracle, [0, 0, 0], [s, 0, 0], [s / 2, 0.1 * s, 0])
    np.testing.assert_allclose(r[:3], [s / 2, 0, 0], atol=1e-5)
    assert r[3] == pytest.approx(s / 2, rel=1e-6)
    # collinear (degenerate branch): midpoint of the longest side
    r = tri_sphere(oracle, [0, 0, 0], [1, 0, 0], [3, 0, 0])
    np.testing.assert_allclose(r, [1.5, 0, 0, 1.5], atol=1e-6)


def test_triangle_sphere_contains_vertices(oracle):
    rng = np.random.default_rng(21)
    bad = 0
    for _ in range(2000):
        v = rng.uniform(0, 100, size=(3, 3)).astype(np.float32)
        r = tri_sphere(oracle, *v)
        dist = np.linalg.norm(v.astype(np.float64) - r[:3], axis=1)
        bad += int(np.any(dist > r[3] * (1 + 1e-5) + 1e-4))
    assert bad == 0


def sphere_node(oracle, sphere, cone, tri):
    arrs = [p(x) for x in (sphere, cone, tri)]
    return oracle.lib().oracle_sphere_node(*[a[1] for a in arrs])


def test_sphere_node_cases(oracle):
    # cone along +z from a point; triangle sphere on the axis ahead -> hit
    assert sphere_node(oracle, [0, 0, 0, 0], [0, 0, 1, 0.1], [0, 0, 10, 0.5]) == 1
    # far off-axis, small cone -> miss
    assert sphere_node(oracle, [0, 0, 0, 0], [0, 0, 1, 0.1], [10, 0, 10, 0.5]) == 0
    # inside the cone angle: perpendicular distance 0.5 at depth 10 with tan(0.1)*10 ~ 1.0 -> hit
    assert sphere_node(oracle, [0, 0, 0, 0], [0, 0, 1, 0.1], [0.5, 0, 10, 0.0]) == 1
    # the node sphere radius widens the test
    assert sphere_node(oracle, [0, 0, 0, 2.0], [0, 0, 1, 0.0], [1.5, 0, 10, 0.0]) == 1
    assert sphere_node(oracle, [0, 0, 0, 1.0], [0, 0, 1, 0.0], [1.5, 0, 10, 0.0]) == 0
    # spread of exactly HALF_PI short-circuits to true (RayTracer.cu:286-287)
    half_pi = float(np.float32(np.float32(3.1415927) / np.float32(2)))
    assert sphere_node(oracle, [0, 0, 0, 0], [0, 0, 1, half_pi], [0, 0, -100, 0.1]) == 1


def union(oracle, name, a, b, *extra):
    out = np.zeros(4, np.float32)
    arrs = [p(x) for x in (a, b)]
    getattr(oracle.lib(), name)(*[x[1] for x in arrs], *extra, out.ctypes.data_as(PF))
    return out


def test_sphere_union_contains_both(oracle):
    rng = np.random.default_rng(22)
    for _ in range(2000):
        s1 = np.concatenate([rng.uniform(0, 100, 3), rng.uniform(0, 5, 1)]).astype(np.float32)
        s2 = np.concatenate([rng.uniform(0, 100, 3), rng.uniform(0, 5, 1)]).astype(np.float32)
        u = union(oracle, "oracle_sphere_union", s1, s2)
        for s in (s1, s2):
            assert np.linalg.norm(s[:3].astype(np.float64) - u[:3]) + s[3] <= u[3] * (1 + 1e-5) + 1e-4


def test_sphere_union_closed_forms(oracle):
    """CreateHierarchySphere (RT:427-446): a sphere holding the other is
    returned unchanged; two points at distance D give their midpoint and
    radius D/2, and two spheres of equal radius r the minimal D/2 + r."""
    rng = np.random.default_rng(23)
    for _ in range(500):
        a, b = rng.uniform(0, 100, 3), rng.uniform(0, 100, 3)
        D = np.linalg.norm(a - b)
        u = union(oracle, "oracle_sphere_union", np.r_[a, 0.0], np.r_[b, 0.0])
        np.testing.assert_allclose(u[:3], (a.astype(np.float32) + b.astype(np.float32)) / 2, rtol=1e-6, atol=1e-5)
        assert u[3] == pytest.approx(D / 2, rel=1e-5, abs=1e-5)
        r = rng.uniform(0.1, 10)
        u = union(oracle, "oracle_sphere_union", np.r_[a, r], np.r_[b, r])
        assert u[3] == pytest.approx(D / 2 + r, rel=1e-5, abs=1e-5)
        big = np.r_[a, D + 5.0].astype(np.float32)
        small = np.r_[b, 2.0].astype(np.float32)
        assert np.array_equal(union(oracle, "oracle_sphere_union", big, small), big)
        assert np.array_equal(union(oracle, "oracle_sphere_union", small, big), big)


def angle(a, b):
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    return np.arccos(np.clip(np.dot(a, b), -1, 1))


@pytest.mark.parametrize("mode", [0, 1])
def test_cone0_contains_rays(oracle, mode):
    rng = np.random.default_rng(23)
    for _ in range(300):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        dirs = axis + rng.normal(scale=0.2, size=(8, 3))
        dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
        cone = np.concatenate([dirs[0], [0.0]]).astype(np.float32)
        for d in dirs[1:]:
            cone = union(oracle, "oracle_cone0", cone, np.concatenate([d, [0.0]]), C.c_uint32(mode))
        if np.any(np.isnan(cone)):
            assert mode == 0
            continue
        for d in dirs:
            assert angle(cone[:3].astype(np.float64), d.astype(np.float64)) <= cone[3] + 1e-4


def test_cone0_duplicate_direction(oracle):
    """Reference union of two identical directions: normalize(0) -> NaN (Q5);
    the guarded union keeps the cone."""
    d = np.array([0.0, 0.6, 0.8, 0.0], np.float32)
    ref = union(oracle, "oracle_cone0", d, d, C.c_uint32(0))
    assert np.all(np.isnan(ref))
    g = union(oracle, "oracle_cone0", d, d, C.c_uint32(1))
    np.testing.assert_array_equal(g, d)


def test_coneN_contains_children(oracle):
    rng = np.random.default_rng(24)
    half_pi = np.float32(np.float32(3.1415927) / np.float32(2))
    for _ in range(2000):
        a = rng.normal(size=3); a /= np.linalg.norm(a)
        b = a + rng.normal(scale=0.3, size=3); b /= np.linalg.norm(b)
        c1 = np.concatenate([a, [rng.uniform(0, 0.3)]]).astype(np.float32)
        c2 = np.concatenate([b, [rng.uniform(0, 0.3)]]).astype(np.float32)
        u = union(oracle, "oracle_coneN", c1, c2)
        if u[3] == half_pi:
            continue
        for c in (c1, c2):
            assert angle(u[:3].astype(np.float64), c[:3].astype(np.float64)) + c[3] <= u[3] + 1e-4
