"""Traversal candidate sets: closed form == literal transcription.

The literal path restates RayTracer.cu:2537-2852 (object culling through
atomicAdd run lengths + inclusive scan + trimming, top-level triangle-sphere
test, then one expand/test/trim step per level) over triangle batches.  The
closed form (every (leaf, triangle) whose ancestors all pass, with object
culling and the batch-boundary case of RayTracer.cu:1192-1244) is what the
oracle frame and the fused GPU kernel enumerate.
"""
import ctypes as C

import numpy as np
import pytest


def make_scene(oracle, rng, sizes, spread=100.0):
    """Triangle soup: object k owns `sizes[k]` consecutive triangles clustered
    around a random centre; spheres bound the object's vertices."""
    tris, spheres = [], []
    start = 0
    for n in sizes:
        c = rng.uniform(20, 180, 3)
        v = c + rng.normal(scale=rng.uniform(1, 12), size=(n, 3, 3))
        tris.append(v)
        flat = v.reshape(-1, 3)
        ctr = 0.5 * (flat.min(0) + flat.max(0))
        rad = np.linalg.norm(flat - ctr, axis=1).max() * 1.0001
        spheres.append([*ctr, rad, start, start + n - 1])
        start += n
    pos = np.concatenate(tris).reshape(-1, 3).astype(np.float32)
    T = pos.shape[0] // 3
    s = type("S", (), {})()
    s.tri_pos = np.hstack([pos, np.ones((3 * T, 1), np.float32)])
    s.tri_nrm = np.zeros_like(s.tri_pos)
    s.tri_mat = np.zeros(3 * T, np.int32)
    s.mat_diffuse = np.ones((1, 4), np.float32)
    s.mat_specular = np.ones((1, 4), np.float32)
    s.object_spheres = np.array(spheres, np.float32)
    s.light_pos = np.array([[100, 190, 100, 1]], np.float32)
    s.light_col = np.ones((1, 4), np.float32)
    s.light_int = np.ones((1, 2), np.float32)
    return s


def build_hierarchy(oracle, rng, n_rays, depth, sub, coherent):
    if coherent:
        o = np.tile(np.array([100.0, 190.0, 100.0], np.float32), (n_rays, 1))
        tgt = rng.uniform(20, 180, (n_rays, 3))
        tgt = tgt[np.lexsort((tgt[:, 0], (tgt[:, 2] // 20)))]
        d = tgt - o
    else:
        o = rng.uniform(20, 180, (n_rays, 3))
        d = rng.normal(size=(n_rays, 3))
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.stack([o, d], 1).reshape(-1, 3), np.float32)
    oracle.call("oracle_set_constants", 1, depth, sub, 0)
    sv = np.arange(n_rays, dtype=np.uint32)
    counts = []
    n = n_rays
    for _ in range(depth):
        n = (n + sub - 1) // sub
        counts.append(n)
    h = np.zeros((2 * sum(counts), 4), np.float32)
    oracle.call("oracle_HierarchyCreationWrapper", rays, sv, sv, n_rays, 0.0, 0.0, h)
    return h, np.array(counts, np.uint32)


def pairs(oracle, which, h, counts, depth, sub, scene, algorithm, tb, cap=4_000_000):
    L = oracle.lib()
    out = np.zeros(cap, np.uint64)
    osc = oracle.make_scene(scene)
    n = getattr(L, which)(h.ctypes.data_as(C.POINTER(C.c_float)), counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                          depth, sub, C.byref(osc), algorithm, tb, out.ctypes.data_as(C.POINTER(C.c_uint64)), cap)
    assert n <= cap
    return out[:n]


@pytest.mark.parametrize("algorithm", [0, 1])
@pytest.mark.parametrize("depth,sub", [(2, 8), (3, 4), (1, 8)])
@pytest.mark.parametrize("tb", [16, 37, 5000])
@pytest.mark.parametrize("coherent", [True, False])
def test_closed_form_equals_literal(oracle, algorithm, depth, sub, tb, coherent):
    rng = np.random.default_rng(31 + depth * 7 + tb + algorithm)
    # object ends at multiples of 16 exercise the batch-boundary case
    sizes = [20, 28, 16, 3, 33, 12, 40, 8, 1, 31, 24, 9]
    scene = make_scene(oracle, rng, sizes)
    h, counts = build_hierarchy(oracle, rng, 700, depth, sub, coherent)
    lit = pairs(oracle, "oracle_leaf_pairs_literal", h, counts, depth, sub, scene, algorithm, tb)
    clo = pairs(oracle, "oracle_leaf_pairs_closed", h, counts, depth, sub, scene, algorithm, tb)
    np.testing.assert_array_equal(lit, clo)
    assert lit.size > 0


def test_batch_boundary_case_drops_pairs(oracle):
    """An object that starts before a batch and ends exactly at the batch's
    upper bound contributes no candidates in that batch (no branch for it in
    RayTracer.cu:1192-1244); with one batch covering everything it does."""
    rng = np.random.default_rng(41)
    scene = make_scene(oracle, rng, [5, 43, 4])   # object 1 = [5, 47]
    scene.object_spheres[:, 3] = 1e4        # every node hits every object sphere
    h, counts = build_hierarchy(oracle, rng, 64, 1, 8, False)
    h[1::2, 3] = np.float32(np.float32(3.1415927) / np.float32(2))   # cones cover everything
    all_b = pairs(oracle, "oracle_leaf_pairs_literal", h, counts, 1, 8, scene, 0, 5000)
    b16 = pairs(oracle, "oracle_leaf_pairs_literal", h, counts, 1, 8, scene, 0, 16)
    T = scene.tri_pos.shape[0] // 3
    assert all_b.size == counts[0] * T
    # object 1 = [5, 47]: z = 47 is not a multiple of 16, nothing dropped
    assert b16.size == all_b.size
    scene2 = make_scene(oracle, rng, [5, 43, 6])   # object 1 = [5, 47], object 2 = [48, 53]
    scene2.object_spheres[:, 3] = 1e4
    scene2.object_spheres[1, 5] = 48.0             # make object 1 end at 48 = 3 * 16 ...
    scene2.object_spheres[2, 4] = 49.0             # ... and object 2 start at 49
    b16 = pairs(oracle, "oracle_leaf_pairs_literal", h, counts, 1, 8, scene2, 0, 16)
    tris = (b16 & 0xFFFFFFFF).astype(np.int64)
    dropped = set(range(32, 48))                   # batch [32, 48) of object 1
    assert not (set(tris.tolist()) & dropped)
    clo = pairs(oracle, "oracle_leaf_pairs_closed", h, counts, 1, 8, scene2, 0, 16)
    np.testing.assert_array_equal(b16, clo)
