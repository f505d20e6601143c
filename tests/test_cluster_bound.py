"""The traversal's cluster pre-test (cluster_may_pass in
ray-tracer_amd/csrc/rsh_device.hpp) must never reject a cluster holding a
triangle whose own FP32 node test (sphere_node, RayTracer.cu:283-298) passes;
otherwise hits would differ from the reference's.  Both predicates are
restated here in numpy float32 with the kernel's operation order and checked
on adversarial samples.  The kernel evaluates the cluster test with the
hardware square root and reciprocal (v_sqrt_f32, v_rcp_f32, within 1 ulp);
the restatement perturbs each of those results by 2 ulp in the direction that
rejects more clusters.  Samples: cluster centres placed on the cone surface, member
triangle spheres pushed to the cluster boundary towards the axis, and
triangle spheres equal to the cluster sphere on the exact test boundary.
A planted mistake (Lipschitz constant 1 instead of max(L, 1/cos), or no
rounding margin on tight cases) must be caught, so the check has teeth.
"""
import numpy as np

f32 = np.float32
HALF_PI = f32(f32(3.1415927) / f32(2.0))


def node_test(S, Cn, T, cosv, tanv):
    s2t = (T[:, :3] - S[:, :3]).astype(f32)
    cd = Cn[:, :3]
    pr = ((s2t[:, 0] * cd[:, 0] + s2t[:, 1] * cd[:, 1]).astype(f32) + s2t[:, 2] * cd[:, 2]).astype(f32)
    v = (s2t - (pr[:, None] * cd).astype(f32)).astype(f32)
    l = np.sqrt(((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]).astype(f32) + v[:, 2] * v[:, 2]).astype(f32)).astype(f32)
    lhs = ((pr * tanv).astype(f32) + ((S[:, 3] + T[:, 3]).astype(f32) / cosv).astype(f32)).astype(f32)
    return (Cn[:, 3] == HALF_PI) | (lhs >= l)


def cluster_test(S, Cn, Q, cosv, tanv, lipschitz_one=False, margin=1e-4, hw_err=2.0 ** -22):
    lo, hi = f32(1 - hw_err), f32(1 + hw_err)
    cd = Cn[:, :3].astype(f32)
    k = (np.sqrt((cd * cd).sum(1).astype(f32)) * lo).astype(f32)
    tk = (tanv * k).astype(f32)
    L = (np.sqrt(f32(1) + tk * tk + f32(2) * tk * np.abs(f32(1) - k * k)) * lo).astype(f32)
    rc = ((f32(1) / cosv) * lo).astype(f32)
    lmax = (np.maximum(L, rc) * f32(1.00001)).astype(f32)
    if lipschitz_one:
        lmax = np.ones_like(lmax)
    ok = (Cn[:, 3] != HALF_PI) & (cosv > 0) & (tanv >= 0) & (k > 0.5) & (k < 1.4) & np.isfinite(lmax)
    s2q = (Q[:, :3] - S[:, :3]).astype(f32)
    pr = (s2q * cd).sum(1).astype(f32)
    v = (s2q - pr[:, None] * cd).astype(f32)
    f = (np.sqrt((v * v).sum(1)) * hi - pr * tanv).astype(f32)
    rho = Q[:, 3]
    allow = (S[:, 3] * rc + rho * lmax).astype(f32)
    E = f32(margin) * (np.sqrt((s2q * s2q).sum(1)) * lo + rho + (np.abs(pr) + rho) * tanv + (S[:, 3] + rho) * lmax) \
        + f32(margin)
    return ~ok | ~(f > allow + E)


def nodes(rng, n):
    S = np.concatenate([rng.uniform(-150, 150, (n, 3)), rng.uniform(0, 10, (n, 1)) * (rng.random((n, 1)) < 0.7)], 1)
    cd = rng.normal(size=(n, 3))
    cd /= np.linalg.norm(cd, axis=1, keepdims=True)
    spread = np.where(rng.random(n) < 0.5, rng.uniform(0, 0.05, n), rng.uniform(0, 1.5707, n))
    Cn = np.concatenate([cd, spread[:, None]], 1).astype(f32)
    perp = rng.normal(size=(n, 3))
    perp -= (perp * cd).sum(1, keepdims=True) * cd
    perp /= np.linalg.norm(perp, axis=1, keepdims=True)
    return S.astype(f32), Cn, np.cos(Cn[:, 3]).astype(f32), np.tan(Cn[:, 3]).astype(f32), cd, spread, perp


def contained_samples(rng, n):
    S, Cn, cosv, tanv, cd, spread, perp = nodes(rng, n)
    along = rng.uniform(-50, 250, n)
    rad = along * np.tan(spread) + S[:, 3] / np.cos(spread)
    qc = S[:, :3] + along[:, None] * cd + (rad + rng.normal(0, 3, n))[:, None] * perp
    rho = rng.uniform(0.01, 20, n)
    Q = np.concatenate([qc, rho[:, None]], 1).astype(f32)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = np.where((rng.random(n) < 0.5)[:, None], -perp, u)        # towards the axis: the worst case
    rt = rho * rng.uniform(0, 1, n) * (rng.random(n) < 0.8)
    dist = (rho - rt) * np.where(rng.random(n) < 0.5, 1.0, rng.uniform(0, 1, n))
    T = np.concatenate([qc + dist[:, None] * u, rt[:, None]], 1).astype(f32)
    over = np.linalg.norm(T[:, :3].astype(np.float64) - Q[:, :3], axis=1) + T[:, 3] - Q[:, 3]
    T[:, 3] = np.where(over > 0, np.maximum(T[:, 3] - over * 1.0001, 0), T[:, 3]).astype(f32)
    return S, Cn, cosv, tanv, Q, T


def tight_samples(rng, n):
    S, Cn, cosv, tanv, cd, spread, perp = nodes(rng, n)
    along = rng.uniform(-50, 250, n)
    rt = rng.uniform(0, 5, n)
    rad = along * np.tan(spread) + (S[:, 3] + rt) / np.cos(spread)
    c = S[:, :3] + along[:, None] * cd + (rad * (1 + rng.normal(0, 1e-7, n)))[:, None] * perp
    T = np.concatenate([c, rt[:, None]], 1).astype(f32)
    return S, Cn, cosv, tanv, T.copy(), T


def violations(gen, seed, rounds, n, **kw):
    rng = np.random.default_rng(seed)
    bad = 0
    for _ in range(rounds):
        S, Cn, cosv, tanv, Q, T = gen(rng, n)
        bad += int((node_test(S, Cn, T, cosv, tanv) & ~cluster_test(S, Cn, Q, cosv, tanv, **kw)).sum())
    return bad


def test_cluster_pretest_is_conservative():
    assert violations(contained_samples, 1, 4, 500_000) == 0
    assert violations(tight_samples, 7, 4, 500_000) == 0


def test_planted_mistakes_are_caught():
    assert violations(contained_samples, 1, 1, 500_000, lipschitz_one=True) > 0
    assert violations(tight_samples, 7, 1, 500_000, margin=0.0) > 0


def test_pretest_culls():
    rng = np.random.default_rng(3)
    S, Cn, cosv, tanv, Q, T = contained_samples(rng, 200_000)
    assert (~cluster_test(S, Cn, Q, cosv, tanv)).mean() > 0.05
