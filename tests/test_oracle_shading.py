"""Independent checks of the oracle's shading, hit-surface and ray-creation
stages and of its per-frame object update.

Each check evaluates the reference's formula in float64 numpy, or a closed
form, and compares the oracle with it.  To show that the inputs discriminate,
every check also runs mutated numpy models (a coefficient swapped, a clamp or
factor dropped, an operand transposed, a sign flipped) and requires each of
them to disagree with the oracle: a restatement carrying that mistake would
fail the check.

Reference (projects/CUDA-Engine/source/RayTracer.cu): :100-107 RgbToInt;
:1616-1709 ColorPrimaryShadowRay; :1776-1882 ColorReflectionRay (packed
integer add of the colour x Ns); :1807-1843 hit point and area-ratio normal;
:682-740 CreateShadowRays; :742-791 CreateReflectionRays; :448-570
UpdateVertex / UpdateBoundingSphere.
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

PF = C.POINTER(C.c_float)
EPS = 0.01
MISS32 = 0xFFFFFFFF


def fp(a):
    return a.ctypes.data_as(PF)


# ----------------------------------------------------------------- models ---
def norm(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def rgb_to_int64(rgb, mut=None):
    """RgbToInt of float64 channels [..., 3]: clamp to [0, 255], truncate,
    R | G << 8 | B << 16."""
    c = np.clip(np.nan_to_num(rgb, nan=255.0), 0.0, 255.0)
    if mut == "round":
        c = np.rint(c)
    if mut == "clamp256":
        c = np.clip(np.nan_to_num(rgb, nan=255.0), 0.0, 256.0)
    i = np.floor(c).astype(np.int64)
    if mut == "bgr":
        i = i[..., ::-1]
    return (i[..., 0] | (i[..., 1] << 8) | (i[..., 2] << 16)).astype(np.uint32)


def blinn_phong(lp, lc, li, P, N, eye, Kd, Ks, scale=1.0, scaled=False, mut=None):
    """float64 colour of one point light (RT:1655-1702; scaled: RT:1863-1866).
    Arrays broadcast over a leading axis; N is used as given (not normalised)."""
    lp, lc = lp[..., :3], lc[..., :3]
    Ld = lp - P
    dist = np.linalg.norm(Ld, axis=-1)
    Ld = Ld / dist[..., None]
    if mut == "att_swap":
        att = 1.0 / (0.75 + dist * 0.00005 + dist * dist * 0.0005)
    elif mut == "att_linear":
        att = 1.0 / (0.75 + dist * 0.0005 + dist * 0.00005)
    else:
        att = 1.0 / (0.75 + dist * 0.0005 + dist * dist * 0.00005)
    ndl = np.sum(Ld * N, axis=-1)
    diff = np.clip(np.maximum(ndl, 0.0), 0.0, 1.0) if mut != "no_clamp" else np.maximum(ndl, 0.0)
    V = norm(P - eye)
    Hv = norm(Ld + V) if mut == "half_sign" else norm(Ld - V)
    ndh = np.maximum(np.sum(Hv * N, axis=-1), 0.0)
    spec = np.power(ndh, Ks[..., 3])
    if mut != "no_clamp":
        spec = np.clip(spec, 0.0, 1.0)
    Ix, Iy = (li[..., 1], li[..., 0]) if mut == "int_swap" else (li[..., 0], li[..., 1])
    kd, ks = (Ks[..., :3], Kd[..., :3]) if mut == "kd_ks_swap" else (Kd[..., :3], Ks[..., :3])
    s = scale if scaled else 1.0
    cd = kd * lc * (diff * Ix * att * s)[..., None]
    cs = ks * lc * (spec * Iy * att * (1.0 if mut == "scale_diffuse_only" else s))[..., None]
    gate = (diff > 0.0) if mut != "spec_ungated" else np.ones_like(diff, bool)
    out = np.where((diff > 0.0)[..., None], cd, 0.0) + np.where(gate[..., None], cs, 0.0)
    return out


def mt64(o, d, v0, e1, e2, bounded=True):
    """Moller-Trumbore (1997) in float64; NaN where the ray misses (bounded),
    else the distance to the triangle's plane."""
    pv = np.cross(d, e2)
    det = np.sum(e1 * pv, axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        tv = o - v0
        u = np.sum(tv * pv, axis=-1) * inv
        qv = np.cross(tv, e1)
        v = np.sum(d * qv, axis=-1) * inv
        t = np.sum(e2 * qv, axis=-1) * inv
        ok = (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1)      # det = 0 gives NaN / inf: a miss
    return (np.where(ok, t, np.nan) if bounded else t), u, v


def closest_hits(scene, orig, dirs, chunk=256):
    """Brute-force closest hit (float64) of rays o + d*eps over every
    triangle: (triangle or -1, t from o, barycentric u, v)."""
    v = scene.tri_pos.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    v0, e1, e2 = v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    n = len(orig)
    tri = np.full(n, -1, np.int64)
    tt = np.full(n, np.inf)
    uu = np.zeros(n)
    vv = np.zeros(n)
    for s in range(0, n, chunk):
        o = (orig[s:s + chunk] + dirs[s:s + chunk] * EPS)[:, None, :]
        d = dirs[s:s + chunk][:, None, :]
        t, u, w = mt64(o, d, v0[None], e1[None], e2[None])
        t = np.where(t > EPS, t, np.inf)
        k = np.argmin(t, axis=1)
        rows = np.arange(len(k))
        best = t[rows, k]
        hit = np.isfinite(best)
        tri[s:s + chunk] = np.where(hit, k, -1)
        tt[s:s + chunk] = np.where(hit, best + EPS, np.inf)
        uu[s:s + chunk] = u[rows, k]
        vv[s:s + chunk] = w[rows, k]
    return tri, tt, uu, vv


def any_hit(scene, orig, dirs, tmax, chunk=256):
    """Brute-force any-hit (float64): some triangle at eps < t < tmax, t
    measured from o + d*eps as RT:1555-1611 does."""
    v = scene.tri_pos.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    v0, e1, e2 = v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    out = np.zeros(len(orig), bool)
    for s in range(0, len(orig), chunk):
        o = (orig[s:s + chunk] + dirs[s:s + chunk] * EPS)[:, None, :]
        t, _, _ = mt64(o, dirs[s:s + chunk][:, None, :], v0[None], e1[None], e2[None])
        tm = tmax[s:s + chunk][:, None]
        out[s:s + chunk] = np.any((t > EPS) & (t < tm), axis=1)
    return out


def exact_where_unambiguous(want_rgb, got, tol=2e-5):
    """Compares packed pixels with a float64 expectation: exact wherever no
    channel lies within tol (relative, at least 1e-4 absolute) of an integer
    step below 255, where FP32 rounding may decide the truncation (the
    oracle's FP32 light term is within 4.5e-6 relative of float64 over the
    random cases of test_light_term_against_float64).  Returns
    (mismatches among unambiguous pixels, share of ambiguous pixels)."""
    v = np.asarray(want_rgb, np.float64)
    frac = v - np.floor(v)
    near = (np.minimum(frac, 1.0 - frac) < np.maximum(tol * np.abs(v), 1e-4)) & (v > 0.5) & (v < 255.5)
    amb = near.any(axis=-1)
    return int(np.sum((rgb_to_int64(v) != got) & ~amb)), float(amb.mean())


# ------------------------------------------------------------- RgbToInt ----
def test_rgb_to_int_known_answers(oracle):
    f = oracle.lib().oracle_rgb_to_int
    assert f(0.0, 0.0, 0.0) == 0
    assert f(254.99, 0.0, 0.0) == 254                 # truncation, not rounding
    assert f(255.0, 0.0, 0.0) == 255
    assert f(300.0, 0.0, 0.0) == 255                  # clamp at 255
    assert f(-5.0, 0.0, 0.0) == 0                     # clamp at 0
    assert f(0.0, 1.0, 0.0) == 1 << 8                 # channel order R | G << 8 | B << 16
    assert f(0.0, 0.0, 1.0) == 1 << 16
    assert f(12.7, 300.0, 128.5) == 12 | (255 << 8) | (128 << 16)
    # helper_math's clamp is fmaxf(lo, fminf(x, hi)): NaN becomes hi
    assert f(float("nan"), 0.0, 0.0) == 255


def test_rgb_to_int_against_model(oracle):
    f = oracle.lib().oracle_rgb_to_int
    rng = np.random.default_rng(1)
    v = rng.uniform(-20, 300, size=(3000, 3)).astype(np.float32)
    got = np.array([f(*map(float, x)) for x in v], np.uint32)
    assert np.array_equal(got, rgb_to_int64(v.astype(np.float64)))
    for mut in ("round", "clamp256", "bgr"):
        assert np.mean(rgb_to_int64(v.astype(np.float64), mut) != got) > 0.05, mut


# ------------------------------------------------------- Blinn-Phong term ---
def light_term(oracle, lp, lc, li, P, N, eye, Kd, Ks, scale=1.0, scaled=False):
    L = oracle.lib()
    out = np.zeros(3, np.float32)
    args = [np.ascontiguousarray(a, np.float32) for a in (lp, lc, li, P, N, eye, Kd, Ks)]
    L.oracle_light_term(*[fp(a) for a in args], float(scale), int(scaled), fp(out))
    return out


def random_shading_cases(rng, n):
    lp = np.c_[rng.uniform(0, 100, (n, 3)), np.ones(n)]
    lc = np.c_[rng.uniform(0.2, 1, (n, 3)), np.ones(n)]
    li = rng.uniform(0.2, 1, (n, 2))
    P = rng.uniform(0, 100, (n, 3))
    N = norm(rng.normal(size=(n, 3))) * rng.uniform(0.5, 2.0, (n, 1))     # lengths != 1 exercise the clamps
    eye = rng.uniform(0, 100, (n, 3))
    Kd = np.c_[rng.uniform(0, 1, (n, 3)), rng.integers(0, 2, n)]
    Ks = np.c_[rng.uniform(0, 1, (n, 3)), rng.choice([1.0, 2.5, 10.0, 63.75, 102.0, 153.0], n)]
    return [a.astype(np.float32) for a in (lp, lc, li, P, N, eye, Kd, Ks)]


@pytest.mark.parametrize("scaled", [False, True])
def test_light_term_against_float64(oracle, scaled):
    rng = np.random.default_rng(2 + scaled)
    cases = random_shading_cases(rng, 1500)
    got = np.array([light_term(oracle, *[a[i] for a in cases], 0.75, scaled) for i in range(1500)])
    c64 = [a.astype(np.float64) for a in cases]
    want = blinn_phong(*c64, scale=0.75, scaled=scaled)
    lit = np.any(want > 0, axis=1)
    assert 0.2 < lit.mean() < 0.8                     # both branches of the diffuse gate are exercised
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=1e-6)
    for mut in ("att_swap", "att_linear", "no_clamp", "half_sign", "int_swap", "kd_ks_swap", "spec_ungated") + \
            (("scale_diffuse_only",) if scaled else ()):
        bad = np.any(np.abs(blinn_phong(*c64, scale=0.75, scaled=scaled, mut=mut) - got) > 1e-3 * (1 + np.abs(got)),
                     axis=1)
        assert bad.mean() > 0.02, mut


def test_light_term_closed_forms(oracle):
    """N.L = 1 at distance d with the eye at the light: att = 1/((0.75 +
    0.0005 d) + 0.00005 d^2), H = N so the specular factor is 1 for any Ns."""
    for d in (1.0, 10.0, 37.5, 120.0):
        P = np.array([50.0, 50.0, 50.0])
        lp = np.r_[P + [0.0, d, 0.0], 1.0]
        N = np.array([0.0, 1.0, 0.0])
        Kd = np.array([0.5, 0.25, 1.0, 0.0])
        Ks = np.array([0.125, 0.75, 0.0, 153.0])
        lc = np.array([1.0, 1.0, 1.0, 1.0])
        li = np.array([0.75, 0.5])
        att = 1.0 / ((0.75 + 0.0005 * d) + 0.00005 * d * d)
        got = light_term(oracle, lp, lc, li, P, N, lp[:3], Kd, Ks)
        np.testing.assert_allclose(got, (Kd[:3] * 0.75 + Ks[:3] * 0.5) * att, rtol=1e-6)
        # the reflection variant multiplies both components by 0.75
        got_s = light_term(oracle, lp, lc, li, P, N, lp[:3], Kd, Ks, 0.75, True)
        np.testing.assert_allclose(got_s, (Kd[:3] * 0.75 + Ks[:3] * 0.5) * att * 0.75, rtol=1e-6)
    # grazing and back-facing light: no diffuse, and the specular term is gated by it
    P = np.array([50.0, 50.0, 50.0])
    for lp in ([60.0, 50.0, 50.0, 1.0], [60.0, 40.0, 50.0, 1.0]):
        got = light_term(oracle, lp, [1, 1, 1, 1], [1, 1], P, [0, 1, 0], [50, 60, 50], [1, 1, 1, 0],
                         [1, 1, 1, 1.0])
        assert not got.any()
    # Ns = 0: the specular factor is pow(x, 0) = 1 wherever the light is in front
    lp = np.array([50.0, 60.0, 50.0, 1.0])
    got = light_term(oracle, lp, [1, 1, 1, 1], [0, 1], P, [0, 1, 0], [80, 51, 50], [0, 0, 0, 0], [1, 1, 1, 0.0])
    np.testing.assert_allclose(got, np.full(3, 1.0 / (0.75 + 0.005 + 0.005)), rtol=1e-6)


# --------------------------------------------------- coloring wrappers ----
def _bind_scene(oracle, scene):
    os_ = oracle.make_scene(scene)
    oracle.lib().oracle_bind_scene(C.byref(os_))
    return os_


def _light_scene(rsh, lights):
    base = rsh.scenes.config0_scene(sphere_slices=8, sphere_stacks=5)
    L = len(lights)
    return dataclasses.replace(
        base, light_pos=np.c_[np.asarray(lights, np.float32), np.ones(L, np.float32)].astype(np.float32),
        light_col=np.tile(np.float32([0.9, 0.8, 1.0, 1.0]), (L, 1)),
        light_int=np.tile(np.float32([0.75, 0.6]), (L, 1)))


def test_shadow_ray_coloring_against_float64(oracle, rsh):
    """ColorPrimaryShadowRay over a random G-buffer with two lights and random
    shadow flags: pixels equal the float64 Blinn-Phong sum over the lit lights,
    x 255, clamped and truncated."""
    rng = np.random.default_rng(4)
    W, H, L = 40, 24, 2
    scene = _light_scene(rsh, [[80.0, 140.0, 90.0], [20.0, 130.0, 40.0]])
    keep = _bind_scene(oracle, scene)
    n = W * H
    pos = np.c_[rng.uniform(50, 150, (n, 3)), np.ones(n)].astype(np.float32)
    pos[rng.random(n) < 0.1] = 0.0                                     # background (RT:708)
    nrm = np.c_[norm(rng.normal(size=(n, 3)) + [0, 1.5, 0]), np.zeros(n)].astype(np.float32)
    dif = np.c_[rng.uniform(0, 1, (n, 3)), np.zeros(n)].astype(np.float32)
    spec = np.c_[rng.uniform(0, 1, (n, 3)), rng.choice([2.0, 63.75, 102.0, 153.0], n)].astype(np.float32)
    oracle.call("oracle_set_constants", L, 2, 8, 0)
    oracle.lib().oracle_bind_gbuffer(fp(dif), fp(spec), fp(pos), fp(nrm))
    sflags = np.where(rng.random(n * L) < 0.3, MISS32, 0).astype(np.uint32)
    eye = np.float32([100.0, 100.0, 10.0])
    pixels = np.zeros(n, np.uint32)
    oracle.call("oracle_ShadowRayColoringWrapper", W, H, L, eye, sflags, pixels)
    del keep

    def model(mut=None, planar=False):
        x, y = np.arange(n) % W, np.arange(n) // W
        col = np.zeros((n, 3))
        for l in range(L):
            idx = (x + y * W + W * H * l) if planar else (x + y * W * L) + W * l
            lit = (sflags[idx] != MISS32) & (np.linalg.norm(pos[:, :3], axis=1) != 0)
            c = blinn_phong(scene.light_pos[l].astype(np.float64)[:3], scene.light_col[l, :3].astype(np.float64),
                            scene.light_int[l].astype(np.float64), pos[:, :3].astype(np.float64),
                            nrm[:, :3].astype(np.float64), eye.astype(np.float64), dif.astype(np.float64),
                            spec.astype(np.float64), mut=mut)
            col += np.where(lit[:, None], c, 0.0)
        return col * 255.0

    bad, amb = exact_where_unambiguous(model(), pixels)
    assert bad == 0 and amb < 0.05, (bad, amb)
    assert (pixels != 0).mean() > 0.5
    for mut, planar in (("att_swap", False), ("half_sign", False), ("int_swap", False), (None, True)):
        assert (rgb_to_int64(model(mut, planar)) != pixels).mean() > 0.05, (mut, planar)


def surface64(scene, tri, o, d):
    """Hit point o + d (eps + t) and the area-ratio barycentric normal
    (RT:1807-1843) in float64: weights (1 - u - v, u, v) of the three vertex
    normals, unnormalised."""
    v = scene.tri_pos.reshape(-1, 3, 4)[tri, :, :3].astype(np.float64)
    nv = scene.tri_nrm.reshape(-1, 3, 4)[tri, :, :3].astype(np.float64)
    # unbounded: a ray through a shared edge may fall a rounding outside the
    # triangle the oracle chose at the tie
    t, u, w = mt64(o + d * EPS, d, v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], bounded=False)
    P = o + d * (EPS + t)[:, None]
    N = (1 - u - w)[:, None] * nv[:, 0] + u[:, None] * nv[:, 1] + w[:, None] * nv[:, 2]
    return P, N, (u, w)


def test_reflection_ray_coloring_against_float64(oracle, rsh):
    """ColorReflectionRay / ColorAndCreateReflectionRay on rays that hit the
    512 scene's spheres and walls: the colour at the hit (area-ratio normal,
    Blinn-Phong x 0.75) times Ns is added to the pixel as a packed integer,
    so a channel above 255 carries into the next; the next bounce is
    reflect(d, N) from the hit point."""
    rng = np.random.default_rng(7)
    scene = rsh.scenes.config0_scene(sphere_slices=8, sphere_stacks=5)
    # Ns spread so that some channels stay below 255 after x Ns
    spec = scene.mat_specular.copy()
    spec[:, 3] = np.float32([1000.0, 1.5, 102.0])
    scene = dataclasses.replace(scene, mat_specular=spec)
    keep = _bind_scene(oracle, scene)
    W, H, L = 32, 16, 1
    n = W * H
    o = rng.uniform(60, 140, (n, 3))
    d = norm(rng.normal(size=(n, 3)))
    tri, t, _, _ = closest_hits(scene, o, d)
    rays = np.zeros((n, 2, 3), np.float32)
    rays[:, 0], rays[:, 1] = o, d
    rays = np.ascontiguousarray(rays.reshape(-1))
    times = np.where(tri >= 0, ((np.minimum(t * 10 + 1, 4095)).astype(np.uint32) << 20) + np.maximum(tri, 0),
                     MISS32).astype(np.uint32)
    assert 0.5 < (tri >= 0).mean()
    pixels0 = rng.integers(0, 1 << 24, n).astype(np.uint32)
    pixels = pixels0.copy()
    flags = np.zeros(n, np.uint32)
    keys = np.zeros(n, np.uint32)
    values = np.zeros(n, np.uint32)
    eye = np.float32([100.0, 100.0, 55.0])
    oracle.call("oracle_set_constants", L, 2, 8, 0)
    oracle.call("oracle_ReflectionRayColoringWrapper", scene.tri_pos, scene.tri_nrm, W, H, L, eye, True, times,
                rays, flags, keys, values, pixels)
    del keep
    hit = tri >= 0
    o64 = o.astype(np.float32).astype(np.float64)
    d64 = d.astype(np.float32).astype(np.float64)
    P, N, _ = surface64(scene, np.maximum(tri, 0), o64, d64)
    m = scene.tri_mat[3 * np.maximum(tri, 0)]
    Kd = scene.mat_diffuse[m].astype(np.float64)
    Ks = scene.mat_specular[m].astype(np.float64)
    lp = scene.light_pos[0].astype(np.float64)[:3]

    def model(mut=None):
        c = blinn_phong(lp, scene.light_col[0, :3].astype(np.float64), scene.light_int[0].astype(np.float64), P, N,
                        eye.astype(np.float64), Kd, Ks, scale=0.75, scaled=True, mut=mut)
        return c * Ks[:, 3:4]

    want = model()
    add = rgb_to_int64(want)
    got_add = pixels - pixels0                                          # uint32 wrap-around = the packed add
    bad, amb = exact_where_unambiguous(want[hit], got_add[hit])
    assert bad == 0 and amb < 0.05, (bad, amb)
    assert np.array_equal(pixels[~hit], pixels0[~hit])
    carries = ((pixels0 & 0xFF) + (add & 0xFF) > 255) & hit
    assert carries.mean() > 0.05                                        # carries between channels occur
    assert (want[hit] < 255).any() and (want[hit] > 255).any()
    for mut in ("att_swap", "half_sign", "kd_ks_swap"):
        assert (rgb_to_int64(model(mut))[hit] != got_add[hit]).sum() >= 5, mut
    # a packed OR instead of the add, or a per-channel saturating add, differ
    assert ((pixels0 | add)[hit] != pixels[hit]).mean() > 0.05
    # the next bounce: origin P, direction reflect(d, N) with the unnormalised normal
    r = rays.reshape(n, 2, 3)
    nd = d64 - 2.0 * N * np.sum(N * d64, axis=1, keepdims=True)
    np.testing.assert_allclose(r[hit, 0], P[hit], atol=2e-3)
    np.testing.assert_allclose(r[hit, 1], nd[hit], atol=2e-4)
    nd_normalised_n = d64 - 2.0 * norm(N) * np.sum(norm(N) * d64, axis=1, keepdims=True)
    assert np.abs(r[hit, 1] - nd_normalised_n[hit]).max() > 1e-3
    assert np.array_equal(flags, np.where(hit, 0, 1).astype(np.uint32))
    assert np.array_equal(values[hit], np.arange(n, dtype=np.uint32)[hit])


# ------------------------------------------------------- ray creation -----
def _random_gbuffer(rng, n):
    pos = np.c_[rng.uniform(50, 150, (n, 3)), np.ones(n)].astype(np.float32)
    pos[rng.random(n) < 0.15] = 0.0
    nrm = np.c_[rng.normal(size=(n, 3)) * rng.uniform(0.5, 2, (n, 1)), np.zeros(n)].astype(np.float32)
    dif = np.zeros((n, 4), np.float32)
    spec = np.zeros((n, 4), np.float32)
    return pos, nrm, dif, spec


def test_shadow_ray_creation_against_float64(oracle, rsh):
    """CreateShadowRays: slot x + W (l + L y) of a live fragment facing light l
    holds origin = the light and direction normalize(P - light)."""
    rng = np.random.default_rng(8)
    W, H, L = 24, 20, 2
    lights = [[80.0, 140.0, 90.0], [20.0, 60.0, 40.0]]
    scene = _light_scene(rsh, lights)
    keep = _bind_scene(oracle, scene)
    n = W * H
    pos, nrm, dif, spec = _random_gbuffer(rng, n)
    oracle.call("oracle_set_constants", L, 2, 8, 0)
    oracle.lib().oracle_bind_gbuffer(fp(dif), fp(spec), fp(pos), fp(nrm))
    S = n * L
    rays = np.zeros(6 * S, np.float32)
    flags = np.full(S, 7, np.uint32)
    keys = np.zeros(S, np.uint32)
    values = np.zeros(S, np.uint32)
    oracle.call("oracle_ShadowRayCreationWrapper", W, H, L, rays, flags, keys, values)
    del keep
    i = np.arange(S)
    x, y, l = i % W, i // (W * L), (i // W) % L
    P = pos[x + y * W, :3].astype(np.float64)
    N = norm(nrm[x + y * W, :3].astype(np.float64))
    lp = np.asarray(lights, np.float64)[l]
    d = norm(P - lp)
    facing = np.sum(d * -N, axis=1)
    live = (np.linalg.norm(P, axis=1) != 0) & (facing > 0)
    clear = np.abs(facing) > 1e-5
    assert 0.2 < live.mean() < 0.8
    assert np.array_equal(flags[clear], np.where(live, 0, 1)[clear].astype(np.uint32))
    r = rays.reshape(S, 2, 3)
    lv = live & clear
    assert np.array_equal(r[lv, 0], np.asarray(lights, np.float32)[l[lv]])
    np.testing.assert_allclose(r[lv, 1], d[lv], atol=3e-7)
    assert np.array_equal(values[lv], i[lv].astype(np.uint32))
    # the key is the separately pinned CreateShadowRayIndex of the written ray
    L_ = oracle.lib()
    for s in np.flatnonzero(lv)[:200]:
        o_, d_ = r[s, 0].copy(), r[s, 1].copy()
        assert keys[s] == L_.oracle_shadow_key_ref(fp(o_), fp(d_), int(l[s]))
    # plausible mistakes: slot y * W * L + x * L + l, direction light - P, facing N instead of -N
    alt_l = i % L
    assert (np.asarray(lights, np.float32)[alt_l][lv] != r[lv, 0]).any(axis=1).mean() > 0.2
    assert np.abs(r[lv, 1] - (-d[lv])).max() > 1.0
    live_wrong = (np.linalg.norm(P, axis=1) != 0) & (np.sum(d * N, axis=1) > 0)
    assert (np.where(live_wrong, 0, 1)[clear] != flags[clear]).mean() > 0.2


def test_reflection_ray_creation_against_float64(oracle, rsh):
    """CreateReflectionRays: slot x + y W L of a covered pixel holds origin P
    and direction reflect(normalize(P - eye), normalize(N)) = I - 2 (N.I) N."""
    rng = np.random.default_rng(9)
    W, H, L = 24, 20, 2
    n = W * H
    pos, nrm, dif, spec = _random_gbuffer(rng, n)
    oracle.call("oracle_set_constants", L, 2, 8, 0)
    oracle.lib().oracle_bind_gbuffer(fp(dif), fp(spec), fp(pos), fp(nrm))
    S = n * L
    rays = np.zeros(6 * S, np.float32)
    flags = np.full(S, 7, np.uint32)
    keys = np.zeros(S, np.uint32)
    values = np.zeros(S, np.uint32)
    eye = np.float32([90.0, 95.0, 10.0])
    oracle.call("oracle_ReflectionRayCreationWrapper", W, H, eye, rays, flags, keys, values)
    p = np.arange(n)
    x, y = p % W, p // W
    slot = x + y * W * L
    P = pos[:, :3].astype(np.float64)
    live = np.linalg.norm(P, axis=1) != 0
    I = norm(P - eye.astype(np.float64))
    Nn = norm(nrm[:, :3].astype(np.float64))
    want = I - 2.0 * Nn * np.sum(Nn * I, axis=1, keepdims=True)
    r = rays.reshape(S, 2, 3)
    assert np.array_equal(flags[slot], np.where(live, 0, 1).astype(np.uint32))
    assert np.array_equal(r[slot[live], 0], pos[live, :3])
    np.testing.assert_allclose(r[slot[live], 1], want[live], atol=5e-7)
    assert np.array_equal(values[slot[live]], slot[live].astype(np.uint32))
    # gap slots (x + y W L + W l, l > 0) are left as they were
    assert (flags[slot + W] == 7).all()
    # mistakes: the unnormalised normal, I + 2 (N.I) N, slot x + y W
    unn = nrm[:, :3].astype(np.float64)
    bad = I - 2.0 * unn * np.sum(unn * I, axis=1, keepdims=True)
    assert np.abs(r[slot[live], 1] - bad[live]).max() > 0.1
    assert np.abs(r[slot[live], 1] - (I + 2.0 * Nn * np.sum(Nn * I, axis=1, keepdims=True))[live]).max() > 0.1


# ------------------------------------------------------- object updates ---
def test_update_wrappers_against_numpy(oracle):
    """UpdateVertex: position = M[obj] @ (x, y, z, w) with row-major M, w
    replaced by the object id; normal = normalize((Nm[obj] @ n).xyz).
    UpdateBoundingSphere: centre = T @ (c, 1), radius x max |S_ii| (i < 3)."""
    rng = np.random.default_rng(10)
    K, T = 5, 40
    obj = np.repeat(np.arange(K), 3 * T // K).astype(np.int32)
    pos = np.c_[rng.uniform(-10, 10, (3 * T, 3)), np.ones(3 * T)].astype(np.float32)
    nrm = np.c_[norm(rng.normal(size=(3 * T, 3))), np.zeros(3 * T)].astype(np.float32)
    spheres4 = np.c_[rng.uniform(-10, 10, (K, 3)), np.ones(K), rng.uniform(0.5, 3, K), np.arange(K) * 8,
                     np.arange(K) * 8 + 7, np.zeros(K)].astype(np.float32)
    model = rng.normal(size=(K, 4, 4)).astype(np.float32)
    model[:, 3] = [0, 0, 0, 1]
    normal = rng.normal(size=(K, 4, 4)).astype(np.float32)
    keep = [obj, pos, nrm, spheres4]
    oracle.lib().oracle_bind_original.argtypes = [PF, PF, C.POINTER(C.c_int32), PF]
    oracle.lib().oracle_bind_original(fp(pos), fp(nrm), obj.ctypes.data_as(C.POINTER(C.c_int32)), fp(spheres4))
    pos_out = np.zeros((3 * T, 4), np.float32)
    nrm_out = np.zeros((3 * T, 4), np.float32)
    oracle.call("oracle_TriangleUpdateWrapper", np.ascontiguousarray(model), np.ascontiguousarray(normal), T,
                pos_out, nrm_out)
    M64 = model.astype(np.float64)[obj]
    want = np.einsum("nij,nj->ni", M64, pos.astype(np.float64))
    np.testing.assert_allclose(pos_out[:, :3], want[:, :3], rtol=1e-5, atol=1e-4)
    assert np.array_equal(pos_out[:, 3], obj.astype(np.float32))
    wn = np.einsum("nij,nj->ni", normal.astype(np.float64)[obj], nrm.astype(np.float64))[:, :3]
    np.testing.assert_allclose(nrm_out[:, :3], norm(wn), atol=1e-5)
    assert (nrm_out[:, 3] == 0).all()
    transposed = np.einsum("nji,nj->ni", M64, pos.astype(np.float64))
    assert np.abs(pos_out[:, :3] - transposed[:, :3]).max() > 1.0
    assert np.abs(nrm_out[:, :3] - wn).max() > 1e-2                   # the normal is normalised

    trans = np.tile(np.eye(4, dtype=np.float32), (K, 1, 1))
    trans[:, :3, 3] = rng.uniform(-5, 5, (K, 3))
    trans[:, :3, :3] += rng.normal(scale=0.1, size=(K, 3, 3)).astype(np.float32)
    scale = np.tile(np.eye(4, dtype=np.float32), (K, 1, 1))
    for k in range(K):
        scale[k, [0, 1, 2], [0, 1, 2]] = rng.uniform(-0.9, 0.9, 3)       # |S_ii| < 1 = S_33
    out3 = np.zeros((K, 6), np.float32)
    oracle.call("oracle_BoundingSphereUpdateWrapper", np.ascontiguousarray(trans), np.ascontiguousarray(scale), K,
                out3)
    c = np.c_[spheres4[:, :3].astype(np.float64), np.ones(K)]
    wc = np.einsum("kij,kj->ki", trans.astype(np.float64), c)[:, :3]
    np.testing.assert_allclose(out3[:, :3], wc, rtol=1e-5, atol=1e-4)
    ms = np.abs(np.diagonal(scale, axis1=1, axis2=2)[:, :3]).max(axis=1)
    np.testing.assert_allclose(out3[:, 3], spheres4[:, 4] * ms, rtol=1e-6)
    assert np.array_equal(out3[:, 4:], spheres4[:, 5:7])
    assert np.abs(out3[:, 3] - spheres4[:, 4]).min() > 1e-3             # max over all 4 diagonal entries would be 1
    del keep


# ----------------------------------------------------- whole frame (C0) ---
@pytest.fixture(scope="module")
def c0_frames(oracle, rsh):
    scene = rsh.scenes.config0_scene()
    dif = scene.mat_diffuse.copy()
    dif[2, 3] = 0.0                                   # one sphere does not reflect: the bounce gate is exercised
    scene = dataclasses.replace(scene, mat_diffuse=dif)
    W = H = 32
    cam = rsh.scenes.config0_camera(W, H)
    base = rsh.config.config0(W, H).with_scene_bounds(scene)
    f0 = oracle.render(scene, cam, dataclasses.replace(base, bounces=0))
    f1 = oracle.render(scene, cam, base)
    f2 = oracle.render(scene, cam, dataclasses.replace(base, bounces=2, bounce_shadows=1))
    return scene, cam, W, H, f0, f1, f2


def primary_rays64(cam, W, H):
    p = np.arange(W * H)
    x, y = p % W, p // W
    a = (x / W - 0.5)[:, None]
    c = (y / H - 0.5)[:, None]
    D, R, U = (np.asarray(v, np.float64) for v in (cam.direction, cam.right, cam.up))
    d = norm(D + R * a + U * c)
    o = np.tile(np.asarray(cam.eye, np.float64), (W * H, 1))
    return o, d


def test_frame_surface_stage_brute_force(c0_frames):
    """The primary pass of a 32 x 32 frame of the 512 scene: per pixel the
    closest triangle over all 2,060 (float64 brute force from the pinhole
    rays of Camera.cpp:108-133), and the G-buffer: hit point, normalised
    area-ratio normal, the first vertex's material."""
    scene, cam, W, H, f0, _, _ = c0_frames
    o, d = primary_rays64(cam, W, H)
    tri, t, _, _ = closest_hits(scene, o, d)
    rec = f0.passes[0]["hits"]
    got_tri = np.where(rec == np.uint64(0xFFFFFFFFFFFFFFFF), -1, (rec & np.uint64(0xFFFFFFFF)).astype(np.int64))
    got_t = (rec >> np.uint64(32)).astype(np.uint32).view(np.float32).astype(np.float64) + EPS
    assert (tri >= 0).all()                                              # the room is closed
    same = got_tri == tri
    assert same.mean() >= 0.999, (~same).sum()
    assert np.all(np.abs(got_t[~same] - t[~same]) <= 1e-5 * t[~same])  # differing only at ties
    g = f0.gbuffer
    P, N, _ = surface64(scene, got_tri, o, d)
    err = np.linalg.norm(g["pos"][:, :3] - P, axis=1)
    assert err.max() < 2e-3, err.max()
    assert np.linalg.norm(P - (o + d * (t - EPS)[:, None]), axis=1).min() > 5e-3   # dropping eps is caught
    assert (g["pos"][:, 3] == 1).all()
    n_err = np.linalg.norm(g["nrm"][:, :3] - norm(N), axis=1)
    assert n_err.max() < 2e-4, n_err.max()
    v = scene.tri_pos.reshape(-1, 3, 4)[got_tri, :, :3].astype(np.float64)
    nv = scene.tri_nrm.reshape(-1, 3, 4)[got_tri, :, :3].astype(np.float64)
    _, u, w = mt64(o + d * EPS, d, v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    rotated = norm(u[:, None] * nv[:, 0] + w[:, None] * nv[:, 1] + (1 - u - w)[:, None] * nv[:, 2])
    assert np.linalg.norm(g["nrm"][:, :3] - rotated, axis=1).max() > 1e-2  # weights on the wrong vertices
    m = scene.tri_mat[3 * got_tri]
    assert np.array_equal(g["dif"], scene.mat_diffuse[m])
    assert np.array_equal(g["spec"], scene.mat_specular[m])


def test_frame_shadow_pass_and_primary_shading(c0_frames):
    """The shadow pass (any-hit from the light towards each lit fragment,
    occluded iff eps < t < |o - P| - 2 eps) against brute force, and the
    primary shading of the frame with no bounce against float64."""
    scene, cam, W, H, f0, _, _ = c0_frames
    g = f0.gbuffer
    P = g["pos"][:, :3].astype(np.float64)
    N = g["nrm"][:, :3].astype(np.float64)
    lp = scene.light_pos[0, :3].astype(np.float64)
    sh = f0.passes[1]
    assert sh["kind"] == 1
    d = norm(P - lp)
    facing = np.sum(d * -norm(N), axis=1)
    live = facing > 0
    assert sh["rays"] == int(live.sum())
    lp32 = scene.light_pos[0, :3].astype(np.float32).astype(np.float64)
    tmax = np.linalg.norm(lp32 - P, axis=1) - 2 * EPS
    occ = any_hit(scene, np.tile(lp32, (len(P), 1)), norm(P - lp32), tmax)
    got_occ = sh["hits"] == np.uint64(MISS32)
    assert np.array_equal(got_occ[~live], np.zeros((~live).sum(), bool))
    assert (got_occ[live] == occ[live]).mean() >= 0.995
    assert 0.01 < occ[live].mean() < 0.95
    lit = live & ~got_occ
    eye = np.asarray(cam.eye, np.float32).astype(np.float64)

    def model(mut=None):
        c = blinn_phong(lp32, scene.light_col[0, :3].astype(np.float64), scene.light_int[0].astype(np.float64), P, N,
                        eye, g["dif"].astype(np.float64), g["spec"].astype(np.float64), mut=mut)
        return np.where(lit[:, None], c, 0.0) * 255.0

    got = f0.pixels.reshape(-1)
    bad, amb = exact_where_unambiguous(model(), got)
    assert bad == 0 and amb < 0.05, (bad, amb)
    for mut in ("att_swap", "half_sign", "kd_ks_swap"):        # the light's two intensities are equal here
        assert (rgb_to_int64(model(mut)) != got).mean() > 0.02, mut
    # shading every live fragment regardless of its shadow flag is caught
    unshadowed = np.where(live[:, None], blinn_phong(
        lp32, scene.light_col[0, :3].astype(np.float64), scene.light_int[0].astype(np.float64), P, N, eye,
        g["dif"].astype(np.float64), g["spec"].astype(np.float64)), 0.0) * 255.0
    assert (rgb_to_int64(unshadowed) != got).mean() > 0.02


def rec_tri(rec):
    return np.where(rec == np.uint64(0xFFFFFFFFFFFFFFFF), -1, (rec & np.uint64(0xFFFFFFFF)).astype(np.int64))


def bounce_walk(scene, cam, f, shadows):
    """Walks the mirror bounces of an oracle frame in float64: per bounce,
    the reflection rays of the gated pixels (first: reflect(normalize(P -
    eye), normalize(N)) from the primary G-buffer; later: reflect(d, N) with
    the previous hit's unnormalised area-ratio normal), their brute-force
    closest hits, the bounce's shadow pass and the packed colour added to the
    pixel.  Yields per bounce (pass record, brute-force triangles, shadow
    check, colour x Ns, hit mask) over the frame's pixels."""
    g = f.gbuffer
    n = g["pos"].shape[0]
    eye = np.asarray(cam.eye, np.float32).astype(np.float64)
    lp = scene.light_pos[0, :3].astype(np.float64)
    P = g["pos"][:, :3].astype(np.float64)
    Nn = norm(g["nrm"][:, :3].astype(np.float64))
    active = g["dif"][:, 3] > 0
    I = norm(P - eye)
    d = I - 2.0 * Nn * np.sum(Nn * I, axis=1, keepdims=True)
    k = 2
    for bounce in range(f.cfg_bounces):
        rp = f.passes[k]
        assert rp["kind"] == 2 and rp["rays"] == int(active.sum())
        idx = np.flatnonzero(active)
        tri_bf = np.full(n, -1, np.int64)
        t_bf = np.full(n, np.inf)
        tri_bf[idx], t_bf[idx], _, _ = closest_hits(scene, P[idx], d[idx])
        got_tri = np.where(active, rec_tri(rp["hits"][:n]), -1)
        hit = got_tri >= 0
        P2, N2, _ = surface64(scene, np.maximum(got_tri, 0), P, d)
        m = scene.tri_mat[3 * np.maximum(got_tri, 0)]
        Kd = scene.mat_diffuse[m].astype(np.float64)
        Ks = scene.mat_specular[m].astype(np.float64)
        lit = hit.copy()
        shadow = None
        if shadows:
            sp = f.passes[k + 1]
            assert sp["kind"] == 1
            sd = norm(P2 - lp)
            facing = np.sum(sd * -norm(N2), axis=1)
            live = hit & (facing > 0)
            grazing = hit & (np.abs(facing) < 1e-4)        # FP32 may decide these either way
            assert abs(sp["rays"] - int(live.sum())) <= int(grazing.sum())
            lp32 = scene.light_pos[0, :3].astype(np.float32).astype(np.float64)
            occ = np.zeros(n, bool)
            li = np.flatnonzero(live)
            occ[li] = any_hit(scene, np.tile(lp32, (len(li), 1)), norm(P2[li] - lp32),
                              np.linalg.norm(lp32 - P2[li], axis=1) - 2 * EPS)
            got_occ = sp["hits"][:n] == np.uint64(MISS32)
            shadow = (got_occ, occ, live & ~grazing)
            lit &= ~got_occ
        c = blinn_phong(lp, scene.light_col[0, :3].astype(np.float64), scene.light_int[0].astype(np.float64), P2,
                        N2, eye, Kd, Ks, scale=0.75, scaled=True)
        colour = np.where(lit[:, None], c * Ks[:, 3:4], 0.0)
        t_got = np.where(hit, np.linalg.norm(P2 - P, axis=1), np.inf)
        yield rp, (tri_bf, t_bf, t_got), got_tri, shadow, colour, hit, (P2, N2, Kd, Ks, lit)
        k += 2 if shadows else 1
        d = d - 2.0 * N2 * np.sum(N2 * d, axis=1, keepdims=True)
        P = P2
        active = hit & (Kd[:, 3] > 0)


@pytest.mark.parametrize("which", ["one_bounce", "two_bounces_with_shadows"])
def test_frame_reflection_bounces_against_float64(c0_frames, which):
    """Gated mirror bounces of the 32 x 32 frame: reflection rays and their
    closest hits by brute force, the bounce shadow passes by brute-force
    any-hit, and each pixel = primary pixel + sum over bounces of
    RgbToInt(0.75-scaled Blinn-Phong at the hit x Ns) as packed integer adds."""
    scene, cam, W, H, f0, f1, f2 = c0_frames
    f, shadows = (f1, False) if which == "one_bounce" else (f2, True)
    f.cfg_bounces = 1 if which == "one_bounce" else 2
    gate = f.gbuffer["dif"][:, 3] > 0
    assert 0.02 < gate.mean() < 0.98
    expect = f0.pixels.reshape(-1).copy()
    amb_all = np.zeros(W * H, bool)
    mut_ns = f0.pixels.reshape(-1).copy()
    for rp, (tri_bf, t_bf, t_got), got_tri, shadow, colour, hit, extra in bounce_walk(scene, cam, f, shadows):
        act = (tri_bf >= 0) | (got_tri >= 0)
        diff = act & (got_tri != tri_bf)
        assert (~diff[act]).mean() >= 0.995
        # the rest are ties: rays through a shared edge, at the same distance
        assert np.all(np.abs(t_got[diff] - t_bf[diff]) <= 1e-5 * t_bf[diff]), (t_got[diff], t_bf[diff])
        assert hit.sum() > 20
        if shadow is not None:
            got_occ, occ, live = shadow
            assert not got_occ[~live & ~hit].any()
            assert (got_occ[live] == occ[live]).mean() >= 0.99
        v = colour
        frac = v - np.floor(v)
        amb_all |= ((np.minimum(frac, 1.0 - frac) < np.maximum(2e-5 * np.abs(v), 1e-4)) & (v > 0.5) &
                    (v < 255.5)).any(axis=1)
        expect = expect + rgb_to_int64(colour)
        P2, N2, Kd, Ks, lit = extra
        mut_ns = mut_ns + rgb_to_int64(colour / np.maximum(Ks[:, 3:4], 1e-30) * 255.0)
    got = f.pixels.reshape(-1)
    assert amb_all.mean() < 0.05
    assert np.array_equal(got[~amb_all], expect[~amb_all].astype(np.uint32))
    assert (mut_ns[~amb_all] != got[~amb_all]).mean() > 0.05            # x 255 instead of x Ns is caught
    assert not np.array_equal(got, f0.pixels.reshape(-1))




# ------------------------------------------------------ camera ray keys ---
def test_primary_key_layout(oracle):
    """Camera rays are keyed by screen position (the reference rasterises
    primary visibility, so this key is the library's own): 8 x 8 tiles in
    row-major order over the view, Morton order inside a tile with x taking
    the even bits, the view id at segment_shift.  Sorted, 8 consecutive keys
    are a 4 x 2 pixel block (one leaf) and 64 one tile (one top node)."""
    f = oracle.lib().oracle_primary_key
    for w, h in ((40, 24), (37, 19), (8, 8)):
        xs, ys = np.meshgrid(np.arange(w), np.arange(h))
        xs, ys = xs.reshape(-1), ys.reshape(-1)
        keys = np.array([f(int(x), int(y), w, 0, 0) for x, y in zip(xs, ys)], np.uint64)
        assert len(np.unique(keys)) == len(keys)
        tiles_x = (w + 7) // 8

        def interleave(x, y):
            m = 0
            for b in range(3):
                m |= ((x >> b) & 1) << (2 * b) | ((y >> b) & 1) << (2 * b + 1)
            return m

        want = np.array([((y // 8) * tiles_x + x // 8) << 6 | interleave(x % 8, y % 8) for x, y in zip(xs, ys)],
                        np.uint64)
        assert np.array_equal(keys, want)
        order = np.argsort(keys, kind="stable")
        if w % 8 == 0 and h % 8 == 0:
            for s in range(0, len(order), 8):
                bx, by = xs[order[s:s + 8]], ys[order[s:s + 8]]
                assert bx.max() - bx.min() == 3 and by.max() - by.min() == 1      # a 4 x 2 block, not 2 x 4
            for s in range(0, len(order), 64):
                tx, ty = xs[order[s:s + 64]] // 8, ys[order[s:s + 64]] // 8
                assert len(set(zip(tx, ty))) == 1
    assert f(3, 5, 40, 7, 26) == (7 << 26) | f(3, 5, 40, 0, 0)
    assert f(3, 5, 40, 7, 0) == f(3, 5, 40, 0, 0)


# ------------------------------------------------------ hierarchy build ---
@pytest.mark.parametrize("n_rays", [1000, 1001, 64])
def test_hierarchy_wrapper_invariants(oracle, n_rays):
    """CreateHierarchyLevel0 / LevelN (RT:1061-1150 via RT:2469-2535): leaf j
    covers the rays of sorted positions 8j .. 8j+7 (gathered through the
    sorted values), node j of level 1 covers leaves 8j .. 8j+7, level 1 starts
    after the n0 leaves, sphere at 2g and cone at 2g+1.  Every origin lies in
    its leaf sphere and every direction in its leaf cone; children lie in
    their parents; a leaf of one ray has radius 0 and spread 0."""
    rng = np.random.default_rng(11 + n_rays)
    o = rng.uniform(20, 180, (n_rays, 3))
    o = o[np.argsort(o[:, 0] // 10 + 100 * (o[:, 2] // 40), kind="stable")]       # loosely coherent runs
    d = norm(rng.normal(size=(n_rays, 3)) + [0.0, 4.0, 0.0])          # within the reference cone's reach
    perm = rng.permutation(n_rays).astype(np.uint32)                              # slot of sorted ray i
    rays = np.zeros((n_rays, 2, 3), np.float32)
    rays[perm, 0], rays[perm, 1] = o, d
    rays = np.ascontiguousarray(rays.reshape(-1))
    oracle.call("oracle_set_constants", 1, 2, 8, 0)
    n0 = (n_rays + 7) // 8
    n1 = (n0 + 7) // 8
    h = np.zeros((2 * (n0 + n1), 4), np.float32)
    oracle.call("oracle_HierarchyCreationWrapper", rays, perm, perm, n_rays, 0.0, 0.0, h)
    sph, cone = h[0::2].astype(np.float64), h[1::2].astype(np.float64)
    o32 = o.astype(np.float32).astype(np.float64)
    d32 = d.astype(np.float32).astype(np.float64)
    for j in range(n0):
        idx = np.arange(8 * j, min(8 * j + 8, n_rays))
        c, r = sph[j, :3], sph[j, 3]
        dist = np.linalg.norm(o32[idx] - c, axis=1)
        assert np.all(dist <= r * (1 + 1e-5) + 1e-4), j
        ang = np.arccos(np.clip(d32[idx] @ cone[j, :3] / np.linalg.norm(cone[j, :3]), -1, 1))
        assert np.all(ang <= cone[j, 3] + 1e-4), j
        if len(idx) == 1:
            assert r == 0 and cone[j, 3] == 0
            assert np.array_equal(c, o32[idx[0]])
    for j in range(n1):
        P = n0 + j
        for ch in range(8 * j, min(8 * j + 8, n0)):
            gap = np.linalg.norm(sph[ch, :3] - sph[P, :3]) + sph[ch, 3]
            assert gap <= sph[P, 3] * (1 + 1e-5) + 1e-4
            a = np.arccos(np.clip(cone[ch, :3] @ cone[P, :3] / np.linalg.norm(cone[ch, :3]) /
                                  np.linalg.norm(cone[P, :3]), -1, 1))
            assert a + cone[ch, 3] <= min(cone[P, 3], np.pi / 2) + 1e-4 or cone[P, 3] >= np.pi / 2 - 1e-6
    # a leaf over the rays in slot order (the gather dropped) fails containment
    wrong = np.linalg.norm(rays.reshape(n_rays, 2, 3)[:8, 0].astype(np.float64) - sph[0, :3], axis=1)
    assert np.any(wrong > sph[0, 3] * (1 + 1e-5) + 1e-4)
