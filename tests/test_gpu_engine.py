"""Native engine (include/rsh.h) against the oracle frame, pass by pass.

Each case renders the same seeded scene and camera on the GPU and with the
CPU oracle and compares live-ray counts, packet counts, every pass's per-slot
hit records, the primary G-buffer and the image.  The bar is the parity
target (hit IDs identical on >= 99.99% of rays, images within 1/255); the
implementation is expected to be bit-exact and the strict cases assert that.
"""
import dataclasses
import sys

import numpy as np
import pytest

from gpu_common import compare_frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(rsh):
    e = rsh.Engine(0)
    yield e
    e.close()


def run_case(engine, oracle, rsh, scene, cam, cfg, strict=True):
    cfg = cfg.with_scene_bounds(scene) if cfg.key_mode == 1 else cfg
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.KEEP_PASSES)
    ofr = oracle.render(scene, cam, cfg, threads=0)
    rep = compare_frames(engine, ofr, cfg, strict=strict)
    print(cfg.width, cfg.height, rep, file=sys.stderr)
    return rep


def c0(rsh, w):
    return rsh.scenes.config0_scene(), rsh.scenes.config0_camera(w, w)


@pytest.mark.parametrize("w", [64, 128, 512])
def test_config0_default(engine, oracle, rsh, w):
    scene, cam = c0(rsh, w)
    run_case(engine, oracle, rsh, scene, cam, rsh.config.config0(w, w))


@pytest.mark.parametrize("variant", [
    dict(algorithm=1),                       # RAH: unsorted hierarchy, no object culling
    dict(key_mode=0),                        # reference spherical keys
    dict(hit_mode=1),                        # reference 32-bit hit packing
    dict(cone_mode=0),                       # reference cone union (NaN nodes)
    dict(depth=3, branch=4),
    dict(depth=1),
    dict(leaf=16, branch=4),
    dict(leaf=4, branch=16),
    dict(leaf=32, branch=2),
    dict(leaf=2, branch=32),
    dict(traversal_kernel=1),                # generic kernel on the 64-ray shape
    dict(traversal_kernel=1, algorithm=1),
    dict(algorithm=1, bounces=2, bounce_shadows=1),
    dict(traversal_chunk=1),                 # one cluster per work item, merged by atomicMin
    dict(traversal_chunk=3),                 # items that start and end inside objects
    dict(traversal_chunk=3, algorithm=1),
    dict(traversal_chunk=1, bounces=2, bounce_shadows=1),
    dict(traversal_chunk=1 << 30),           # one item per top node
    dict(traversal_splits=3, traversal_kernel=1),
    dict(traversal_splits=5, traversal_kernel=1, bounces=2, bounce_shadows=1),
    dict(traversal_splits=4, depth=3, branch=4),
    dict(key_layout=1),
    dict(key_layout=2),                      # bit-interleaved cell / direction key
    dict(traversal_prefilter=1),             # without the ray pre-tests
    dict(traversal_prefilter=1, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=1, leaf=4, branch=16, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=1, traversal_chunk=1, hit_mode=1),
    dict(traversal_prefilter=2, bounces=2, bounce_shadows=1),   # triangle pre-test only
    dict(traversal_prefilter=3),             # and cluster ray masks on every pass
    dict(traversal_prefilter=3, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=3, leaf=4, branch=16, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=3, leaf=32, branch=2),
    dict(traversal_prefilter=3, traversal_chunk=1, hit_mode=1),
    dict(triangle_batch=512),
    dict(bounces=2, bounce_shadows=1),
])
def test_config0_variants(engine, oracle, rsh, variant):
    scene, cam = c0(rsh, 96)
    cfg = dataclasses.replace(rsh.config.config0(96, 96), **variant)
    run_case(engine, oracle, rsh, scene, cam, cfg)


def test_config1_small(engine, oracle, rsh):
    scene = rsh.scenes.config1_scene(boxes=300)
    cam = rsh.scenes.config1_camera(128, 128)
    run_case(engine, oracle, rsh, scene, cam, rsh.config.config1(128, 128))


def test_config2_small(engine, oracle, rsh):
    scene = rsh.scenes.config2_scene(grid=3, col_slices=16, col_stacks=12, tiles=8, tile_sub=2, arch_around=6,
                                     arch_along=8)
    cam = rsh.scenes.config2_camera(160, 90)
    run_case(engine, oracle, rsh, scene, cam, rsh.config.config2(160, 90))


@pytest.fixture(scope="module")
def atrium(rsh):
    return rsh.scenes.config2_scene()


def test_config2_full_scene_low_res(engine, oracle, rsh, atrium):
    """The full 262k-triangle atrium: many top nodes whose 64 shadow rays are
    all occluded part-way through a batch of candidate blocks (regression:
    the survivor queue must not grow past its capacity after the wave exits)."""
    cam = rsh.scenes.config2_camera(192, 108)
    run_case(engine, oracle, rsh, atrium, cam, rsh.config.config2(192, 108))


@pytest.mark.parametrize("variant", [
    dict(),
    dict(algorithm=1),
    dict(traversal_prefilter=1),
    dict(traversal_prefilter=1, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=3, bounces=2, bounce_shadows=1),   # counting keeps every ray
    dict(traversal_chunk=1, hit_mode=1),
    dict(depth=3, branch=4),                 # generic kernel
    dict(leaf=16, branch=4, bounces=2, bounce_shadows=1),
])
def test_traversal_counts_match_oracle(engine, oracle, rsh, variant):
    """Candidate (top node, triangle) pairs, passing (leaf, triangle) pairs and
    ray-triangle tests equal the oracle's for every closest-hit pass (the
    cluster pre-test and the work split must not change the candidate sets).
    Shadow passes stop early once every ray of a wave is occluded, so their
    leaf pairs and tests may only be fewer."""
    scene, cam = c0(rsh, 96)
    cfg = dataclasses.replace(rsh.config.config0(96, 96), **variant).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.COUNTERS)
    gp = engine.passes()
    op = oracle.render(scene, cam, cfg, threads=0, keep_hits=False).passes
    assert len(gp) == len(op)
    for g, o in zip(gp, op):
        assert g["pairs_top"] == o["pairs_top"], (g, o)
        if g["kind"] == 1:
            assert g["pairs_leaf"] <= o["pairs_leaf"] and g["ray_tests"] <= o["ray_tests"], (g, o)
        else:
            assert (g["pairs_leaf"], g["ray_tests"]) == (o["pairs_leaf"], o["ray_tests"]), (g, o)
        assert g["tri_tests"] <= g["pairs_top"]
        assert g["mt_tests"] <= g["ray_tests"]


def test_config2_counts_match_oracle(engine, oracle, rsh):
    """Same on an atrium slice, where clusters do cull."""
    scene = rsh.scenes.config2_scene(grid=3, col_slices=16, col_stacks=12, tiles=8, tile_sub=2, arch_around=6,
                                     arch_along=8)
    cam = rsh.scenes.config2_camera(160, 90)
    cfg = rsh.config.config2(160, 90).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.COUNTERS)
    gp = engine.passes()
    op = oracle.render(scene, cam, cfg, threads=0, keep_hits=False).passes
    for g, o in zip(gp, op):
        assert g["pairs_top"] == o["pairs_top"]
        if g["kind"] != 1:
            assert (g["pairs_leaf"], g["ray_tests"]) == (o["pairs_leaf"], o["ray_tests"])
    assert sum(g["tri_tests"] for g in gp) < sum(g["pairs_top"] for g in gp)


def test_render_twice_is_deterministic(engine, rsh):
    scene, cam = c0(rsh, 128)
    cfg = rsh.config.config0(128, 128).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg)
    a = engine.pixels()
    engine.render(cam, cfg)
    np.testing.assert_array_equal(a, engine.pixels())


def test_invalid_arguments_are_rejected(engine, rsh):
    scene, cam = c0(rsh, 32)
    engine.load_scene(scene)
    bad = dataclasses.replace(rsh.config.config0(32, 32), leaf=0)
    with pytest.raises(rsh.RshError):
        engine.render(cam, bad)
    bad = dataclasses.replace(rsh.config.config0(32, 32), cell_bits=8, dir_bits=8)
    with pytest.raises(rsh.RshError):
        engine.render(cam, bad)
    broken = dataclasses.replace(scene, object_spheres=scene.object_spheres.copy())
    broken.object_spheres[1, 4] += 1.0        # gap in the object tiling
    with pytest.raises(rsh.RshError):
        engine.load_scene(broken)
