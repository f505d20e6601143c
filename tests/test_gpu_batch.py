"""Batched views (rsh_engine_render_views) and graph replay (render flag bit 3).

A batch traces every view's rays together, with the view id in the key bits
above the octahedral hash (segment_shift), so it is checked bit for bit
against the oracle's batch (oracle_render_views), and per view against
single-view frames.  Graph replays are checked against the eager frame, with
other eager work and synchronisation between the replays.
"""
import dataclasses

import numpy as np
import pytest

from gpu_common import compare_frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(rsh):
    e = rsh.Engine(0)
    yield e
    e.close()


def batch_cfg(rsh, scene, base, **kw):
    cfg = dataclasses.replace(base, cell_bits=4, dir_bits=6, segment_shift=26, **kw)
    return cfg.with_scene_bounds(scene)


def random_views(rsh, seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n):
        w, h = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
        views.append((rsh.scenes.random_view(rng, w, h), w, h))
    return views


@pytest.mark.parametrize("variant", [dict(), dict(algorithm=1), dict(bounces=2, bounce_shadows=1), dict(leaf=4, branch=16),
                                     dict(soft_shadows=1), dict(bounces=2, bounce_shadows=1, soft_shadows=1)])
def test_batch_matches_oracle(engine, oracle, rsh, variant):
    scene = rsh.scenes.config0_scene()
    views = random_views(rsh, 11, 7, 24, 80)
    cfg = batch_cfg(rsh, scene, rsh.config.config0(64, 64), **variant)
    engine.load_scene(scene)
    engine.render_views(views, cfg, flags=engine.KEEP_PASSES)
    ofr = oracle.render_views(scene, views, cfg, threads=0)
    compare_frames(engine, ofr, engine.cfg)


def test_batch_config1_matches_oracle(engine, oracle, rsh):
    scene = rsh.scenes.config1_scene(boxes=200)
    views = random_views(rsh, 12, 5, 32, 72)
    cfg = batch_cfg(rsh, scene, rsh.config.config1(64, 64))
    engine.load_scene(scene)
    engine.render_views(views, cfg, flags=engine.KEEP_PASSES)
    compare_frames(engine, oracle.render_views(scene, views, cfg, threads=0), engine.cfg)


def test_batch_equals_per_view_frames(engine, rsh):
    """Views traced together give the images of views traced alone (up to
    rays where the different hierarchy grouping moves a culling decision;
    none are expected, at most 1e-4 of pixels allowed)."""
    scene = rsh.scenes.config1_scene(boxes=300)
    views = random_views(rsh, 13, 9, 48, 160)
    cfg = batch_cfg(rsh, scene, rsh.config.config1(64, 64))
    engine.load_scene(scene)
    engine.render_views(views, cfg)
    parts = engine.split_views(engine.pixels(), views)
    bad = total = 0
    for (cam, w, h), img in zip(views, parts):
        engine.render(cam, dataclasses.replace(cfg, width=w, height=h))
        bad += int((engine.pixels() != img).sum())
        total += w * h
    assert bad <= 1e-4 * total, (bad, total)


def test_single_view_batch_equals_render(engine, rsh):
    scene = rsh.scenes.config0_scene()
    cam = rsh.scenes.config0_camera(96, 80)
    cfg = rsh.config.config0(96, 80).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg)
    a = engine.pixels()
    engine.render_views([(cam, 96, 80)], cfg)
    np.testing.assert_array_equal(engine.pixels(), a)


def test_graph_replay_matches_eager(engine, rsh):
    scene = rsh.scenes.config1_scene(boxes=300)
    views = random_views(rsh, 14, 6, 64, 128)
    cfg = batch_cfg(rsh, scene, rsh.config.config1(64, 64))
    engine.load_scene(scene)
    engine.render_views(views, cfg)
    eager = engine.pixels().copy()
    st = engine.render_views(views, cfg, flags=engine.GRAPH)          # eager frame + capture
    assert st.graph_replayed == 0
    np.testing.assert_array_equal(engine.pixels(), eager)
    small = rsh.scenes.config1_camera(32, 32)
    for i in range(3):
        engine.render(small, dataclasses.replace(cfg, width=32, height=32, segment_shift=0))   # eager work between
        engine.render_views(views, cfg)                               # eager batch: no reallocation
        np.testing.assert_array_equal(engine.pixels(), eager)
        st = engine.render_views(views, cfg, flags=engine.GRAPH)
        assert st.graph_replayed == 1, i
        np.testing.assert_array_equal(engine.pixels(), eager, err_msg="replay %d" % i)


def test_graph_replay_soft_shadows(engine, rsh):
    """Replays of a captured soft-shadow batch (16 shadow rays per slot) equal
    its eager frame, with other eager work between the replays."""
    scene = rsh.scenes.config0_scene()
    views = random_views(rsh, 15, 5, 40, 96)
    cfg = batch_cfg(rsh, scene, rsh.config.config0(64, 64), soft_shadows=1, bounces=2, bounce_shadows=1)
    engine.load_scene(scene)
    engine.render_views(views, cfg)
    eager = engine.pixels().copy()
    st = engine.render_views(views, cfg, flags=engine.GRAPH)
    assert st.graph_replayed == 0
    np.testing.assert_array_equal(engine.pixels(), eager)
    for i in range(3):
        engine.render(rsh.scenes.config0_camera(32, 32), rsh.config.config0(32, 32).with_scene_bounds(scene))
        engine.render_views(views, cfg)
        np.testing.assert_array_equal(engine.pixels(), eager)
        st = engine.render_views(views, cfg, flags=engine.GRAPH)
        assert st.graph_replayed == 1, i
        np.testing.assert_array_equal(engine.pixels(), eager, err_msg="replay %d" % i)


def test_graph_rejects_other_flags(engine, rsh):
    scene = rsh.scenes.config0_scene()
    engine.load_scene(scene)
    cam = rsh.scenes.config0_camera(32, 32)
    with pytest.raises(rsh.RshError):
        engine.render(cam, rsh.config.config0(32, 32), flags=engine.GRAPH | engine.TIMING)


def test_batch_rejects_view_ids_that_do_not_fit(engine, rsh):
    scene = rsh.scenes.config0_scene()
    engine.load_scene(scene)
    views = random_views(rsh, 15, 3, 16, 24)
    bad = dataclasses.replace(rsh.config.config0(64, 64), segment_shift=26).with_scene_bounds(scene)   # 29-bit hash
    with pytest.raises(rsh.RshError):
        engine.render_views(views, bad)


@pytest.mark.parametrize("variant", [dict(), dict(bounces=2, bounce_shadows=1)])
def test_64_views_match_oracle(engine, oracle, rsh, variant):
    """64 views in one batch, so view ids 0..63 all occur in key bits 26-31
    (the batching workload's view count), at small sides (8-40 px) so that
    the oracle renders them in seconds."""
    scene = rsh.scenes.config0_scene()
    views = random_views(rsh, 64 + len(variant), 64, 8, 40)
    cfg = batch_cfg(rsh, scene, rsh.config.config0(64, 64), **variant)
    engine.load_scene(scene)
    engine.render_views(views, cfg, flags=engine.KEEP_PASSES)
    compare_frames(engine, oracle.render_views(scene, views, cfg, threads=0), engine.cfg)


def test_views_at_workload_sides_match_oracle(engine, oracle, rsh):
    """Views at the batching workload's sides (128-1024 px), fewer of them."""
    scene = rsh.scenes.config1_scene(boxes=300)
    rng = np.random.default_rng(65)
    views = [(rsh.scenes.random_view(rng, s, s), s, s) for s in (128, 1024, 301)]
    cfg = batch_cfg(rsh, scene, rsh.config.config1(64, 64))
    engine.load_scene(scene)
    engine.render_views(views, cfg, flags=engine.KEEP_PASSES)
    compare_frames(engine, oracle.render_views(scene, views, cfg, threads=0), engine.cfg)


def test_full_64_view_batch_equals_plain_frames(engine, rsh):
    """The batching workload itself (64 square views of side 128-1024 at
    random poses, 23.5M pixels, office-scale scene) against the plain
    pipeline, one frame per view: pixels identical but for at most 1e-4 of
    them, and camera-pass hit triangles likewise."""
    scene = rsh.scenes.config1_scene()
    views = rsh.scenes.batch_views()
    cfg = batch_cfg(rsh, scene, rsh.config.config1(64, 64))
    engine.load_scene(scene)
    engine.render_views(views, cfg, flags=engine.KEEP_PASSES)
    batch_px = engine.pixels().reshape(-1).copy()
    batch_hits = engine.pass_hits(0)
    bad_px = bad_hit = total = 0
    off = 0
    for cam, w, h in views:
        engine.render(cam, dataclasses.replace(cfg, width=w, height=h), flags=engine.KEEP_PASSES)
        bad_px += int((engine.pixels().reshape(-1) != batch_px[off:off + w * h]).sum())
        a = engine.pass_hits(0) & np.uint64(0xFFFFFFFF)
        bad_hit += int((a != (batch_hits[off:off + w * h] & np.uint64(0xFFFFFFFF))).sum())
        total += w * h
        off += w * h
    assert total == 23_543_814
    assert bad_px <= 1e-4 * total, bad_px
    assert bad_hit <= 1e-4 * total, bad_hit


@pytest.mark.parametrize("variant", [dict(), dict(bounces=2, bounce_shadows=1, soft_shadows=1)])
def test_graph_replay_follows_the_camera(engine, rsh, variant):
    """A frame graph whose shadow and reflection passes keep their sizes in
    device memory replays for any camera of the same size: each replay equals
    the eager frame of its own camera (or, when the camera's work list
    outgrows the captured capacities, is rendered eagerly instead), and the
    returned ray counts are the camera's."""
    scene = rsh.scenes.config1_scene(boxes=300)
    cfg = dataclasses.replace(rsh.config.config1(96, 80), **variant).with_scene_bounds(scene)
    engine.load_scene(scene)
    rng = np.random.default_rng(77)
    cams = [rsh.scenes.random_view(rng, 96, 80) for _ in range(5)]
    st = engine.render(cams[0], cfg, flags=engine.GRAPH)          # eager frame + capture
    assert st.graph_replayed == 0
    replayed = 0
    for i, cam in enumerate(cams[1:] + cams[:1]):
        eager = engine.render(cam, cfg)
        want = engine.pixels().copy()
        st = engine.render(cam, cfg, flags=engine.GRAPH)
        replayed += st.graph_replayed
        np.testing.assert_array_equal(engine.pixels(), want, err_msg=f"camera {i}")
        assert st.rays_total == eager.rays_total, (i, st.rays_total, eager.rays_total)
        for k in range(eager.pass_count):
            assert st.pass_rays[k] == eager.pass_rays[k], (i, k)
    assert replayed >= 3


def test_async_pixel_readback_beside_the_next_frame(engine, rsh):
    """rsh_engine_read_pixels_async: the copy of frame A, queued before frame
    B renders, holds A's image (B clears its pixels only once the copy has
    ended); the copy of B holds B's; a graph-replayed frame after an
    asynchronous read waits for it too."""
    scene = rsh.scenes.config1_scene(boxes=300)
    cfg = rsh.config.config1(96, 80).with_scene_bounds(scene)
    engine.load_scene(scene)
    rng = np.random.default_rng(5)
    a, b = rsh.scenes.random_view(rng, 96, 80), rsh.scenes.random_view(rng, 96, 80)
    engine.render(a, cfg)
    want_a = engine.pixels().copy()
    engine.render(b, cfg)
    want_b = engine.pixels().copy()
    assert not np.array_equal(want_a, want_b)
    for flags in (0, engine.GRAPH):
        for _ in range(2):
            engine.render(a, cfg, flags=flags)
            engine.read_pixels_async()
            engine.render(b, cfg, flags=flags)
            np.testing.assert_array_equal(engine.wait_pixels(), want_a, err_msg=f"flags {flags}")
            engine.read_pixels_async()
            np.testing.assert_array_equal(engine.wait_pixels(), want_b, err_msg=f"flags {flags}")
    assert engine.wait_pixels() is not None      # no copy pending: returns at once
