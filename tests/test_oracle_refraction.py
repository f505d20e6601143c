"""Refraction rays in the oracle: helper_math.h's refract (helper_math.h:1422-1431)
and CreateRefractionRayIndex (RayTracer.cu:200-227), whose kernel the
reference never launches.

refract is pinned by Snell's law (unit length, coplanar with the incident
ray and the normal, sin t = ior sin i, across the surface), normal incidence
and total internal reflection.  The key is pinned by its relation to the
separately restated reflection key (RayTracer.cu:170-198): the same four
fields below a distance field shifted by 26 in a 32-bit word.
"""
import numpy as np
import pytest


def refract(oracle, i, n, ior):
    out = np.zeros(3, np.float32)
    oracle.lib().oracle_refract(oracle.fp(np.asarray(i, np.float32)), oracle.fp(np.asarray(n, np.float32)),
                                float(ior), oracle.fp(out))
    return out.astype(np.float64)


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def test_refract_snell(oracle):
    rng = np.random.default_rng(11)
    ior = 1.0 / 1.52
    for _ in range(200):
        n = unit(rng.normal(size=3))
        i = unit(rng.normal(size=3))
        if np.dot(i, n) > -0.05:
            i = unit(i - 2 * np.dot(i, n) * n - 0.1 * n)
        d = refract(oracle, i, n, ior)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        assert abs(np.linalg.det(np.stack([i, n, d]).astype(np.float64))) < 1e-5           # coplanar
        sin_i = np.linalg.norm(np.cross(i, n))
        sin_t = np.linalg.norm(np.cross(d, n))
        assert sin_t == pytest.approx(ior * sin_i, abs=2e-6)
        assert np.dot(d, n) < 0                                                          # through the surface
    n = np.array([0, 1, 0], np.float32)
    assert np.allclose(refract(oracle, [0, -1, 0], n, ior), [0, -1, 0], atol=1e-7)       # normal incidence
    graze = unit([1, -0.05, 0])
    assert np.array_equal(refract(oracle, graze, n, 1.52), [0, 0, 0])                    # total internal reflection


def test_refraction_key_against_reflection_key(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(12)
    for _ in range(300):
        o = (rng.normal(size=3) * rng.uniform(1, 150)).astype(np.float32)
        d = unit(rng.normal(size=3))
        dist = float(np.linalg.norm(o.astype(np.float64)))
        if abs(dist - round(dist)) < 1e-3:
            continue                                  # the distance field's float rounding at an integer
        field = min(int(dist), 128)
        refl = L.oracle_reflection_key_ref(oracle.fp(o), oracle.fp(d))
        refr = L.oracle_refraction_key_ref(oracle.fp(o), oracle.fp(d))
        assert refr == ((((field << 26) & 0xFFFFFFFF) | (refl - 1)) + 1) & 0xFFFFFFFF, (o, d)
    # the distance field keeps only its low 6 bits (7 + 7 + 7 + 7 + 5 = 33 bits in 32)
    o = np.array([100.5, 0.0, 0.0], np.float32)
    d = unit([0, 0, 1])
    refl = L.oracle_reflection_key_ref(oracle.fp(o), oracle.fp(d))
    assert L.oracle_refraction_key_ref(oracle.fp(o), oracle.fp(d)) == (((100 & 63) << 26) | (refl - 1)) + 1
