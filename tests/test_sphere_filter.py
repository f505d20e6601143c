"""The traversal's sphere_node_fast (ray-tracer_amd/csrc/rsh_device.hpp) must
decide every node test exactly as the reference's expression
(RayTracer.cu:283-298, sphere_node) evaluated in correctly rounded FP32.  It
takes the hardware reciprocal and square root (v_rcp_f32 / v_sqrt_f32, within
1 ulp) and falls back to the exact expression when the two sides are within a
margin M.  Restated here in numpy float32 with the kernel's operation order;
each hardware estimate is replaced by the correctly rounded value moved by a
random -1 / 0 / +1 ulp, on samples placed on the test boundary (where the
estimate error matters) and on generic samples.  Planted mistake: a margin 64x
too small must produce decisions that contradict the exact test.
"""
import numpy as np

f32 = np.float32
HALF_PI = f32(f32(3.1415927) / f32(2.0))


def perturb(x, rng):
    step = rng.integers(-1, 2, x.shape)
    up = np.nextafter(x, np.float32(np.inf))
    dn = np.nextafter(x, np.float32(-np.inf))
    return np.where(step > 0, up, np.where(step < 0, dn, x)).astype(f32)


def node_terms(S, Cn, T, tanv):
    cd = Cn[:, :3]
    s2t = (T[:, :3] - S[:, :3]).astype(f32)
    pr = ((s2t[:, 0] * cd[:, 0] + s2t[:, 1] * cd[:, 1]).astype(f32) + s2t[:, 2] * cd[:, 2]).astype(f32)
    v = (s2t - (pr[:, None] * cd).astype(f32)).astype(f32)
    dd = ((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]).astype(f32) + v[:, 2] * v[:, 2]).astype(f32)
    a = (pr * tanv).astype(f32)
    R = (S[:, 3] + T[:, 3]).astype(f32)
    return a, R, dd


def exact_test(S, Cn, T, cosv, tanv):
    a, R, dd = node_terms(S, Cn, T, tanv)
    return (Cn[:, 3] == HALF_PI) | ((a + (R / cosv).astype(f32)).astype(f32) >= np.sqrt(dd).astype(f32))


def fast_test(S, Cn, T, cosv, tanv, rng, scale=f32(2.0 ** -18)):
    """Returns (decided, decision) of the estimate path; undecided lanes take
    the exact expression in the kernel."""
    a, R, dd = node_terms(S, Cn, T, tanv)
    rcos = perturb((f32(1) / cosv).astype(f32), rng)
    q = (R * rcos).astype(f32)
    rs = perturb(np.sqrt(dd).astype(f32), rng)
    diff = ((a + q).astype(f32) - rs).astype(f32)
    M = (scale * ((np.abs(a) + np.abs(q)).astype(f32) + rs).astype(f32) + f32(2.0 ** -60)).astype(f32)
    decided = (diff > M) | (diff < -M)
    return decided & (Cn[:, 3] != HALF_PI), diff > M


def samples(rng, n, tight):
    S = np.concatenate([rng.uniform(-150, 150, (n, 3)), rng.uniform(0, 10, (n, 1)) * (rng.random((n, 1)) < 0.7)], 1)
    cd = rng.normal(size=(n, 3))
    cd /= np.linalg.norm(cd, axis=1, keepdims=True)
    spread = np.where(rng.random(n) < 0.5, rng.uniform(0, 0.05, n), rng.uniform(0, 1.5707, n))
    Cn = np.concatenate([cd, spread[:, None]], 1).astype(f32)
    perp = rng.normal(size=(n, 3))
    perp -= (perp * cd).sum(1, keepdims=True) * cd
    perp /= np.linalg.norm(perp, axis=1, keepdims=True)
    along = rng.uniform(-50, 250, n)
    rt = rng.uniform(0, 5, n)
    rad = along * np.tan(spread) + (S[:, 3] + rt) / np.cos(spread)
    if tight:
        rad = rad * (1 + rng.normal(0, 1e-7, n))
    else:
        rad = rad * rng.uniform(0, 2, n)
    c = S[:, :3] + along[:, None] * cd + rad[:, None] * perp
    T = np.concatenate([c, rt[:, None]], 1).astype(f32)
    return S.astype(f32), Cn, np.cos(Cn[:, 3]).astype(f32), np.tan(Cn[:, 3]).astype(f32), T


def contradictions(seed, rounds, n, tight, **kw):
    rng = np.random.default_rng(seed)
    bad = decided_total = 0
    for _ in range(rounds):
        S, Cn, cosv, tanv, T = samples(rng, n, tight)
        decided, fast = fast_test(S, Cn, T, cosv, tanv, rng, **kw)
        exact = exact_test(S, Cn, T, cosv, tanv)
        bad += int((decided & (fast != exact)).sum())
        decided_total += int(decided.sum())
    return bad, decided_total


def test_fast_path_agrees_with_exact_test():
    bad, decided = contradictions(11, 4, 500_000, tight=True)
    assert bad == 0 and decided > 0
    bad, decided = contradictions(12, 2, 500_000, tight=False)
    assert bad == 0 and decided > 0.99 * 1_000_000     # the estimates decide nearly every generic test


def test_planted_small_margin_is_caught():
    bad, _ = contradictions(11, 1, 500_000, tight=True, scale=f32(2.0 ** -24))
    assert bad > 0
