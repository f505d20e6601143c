"""The oracle's batch of views (oracle_render_views) against its single-view
frames: the batch lays the views out as one image of sum(w*h) pixels and, with
segment_shift, keeps each view's rays apart in the key, so every view's
pixels and per-pass hits must equal the frame of that view traced alone
(up to rays where a different hierarchy grouping moves one of the
reference's non-conservative node decisions: at most 1e-4 of them).  This
pins the view table (offsets, per-view camera and eye), the slot layout of a
batch and the view id in the keys, independently of the GPU."""
import dataclasses

import numpy as np
import pytest


def views_for(rsh, seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        w, h = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
        out.append((rsh.scenes.random_view(rng, w, h), w, h))
    return out


@pytest.mark.parametrize("variant", [dict(), dict(bounces=2, bounce_shadows=1)])
def test_batch_equals_single_view_frames(oracle, rsh, variant):
    scene = rsh.scenes.config0_scene(sphere_slices=16, sphere_stacks=9)
    views = views_for(rsh, 31, 5, 12, 40)
    cfg = dataclasses.replace(rsh.config.config0(8, 8), cell_bits=4, dir_bits=6, segment_shift=26,
                              **variant).with_scene_bounds(scene)
    batch = oracle.render_views(scene, views, cfg)
    px = batch.pixels.reshape(-1)
    off = 0
    bad = total = 0
    for cam, w, h in views:
        single = oracle.render(scene, cam, dataclasses.replace(cfg, width=w, height=h))
        bad += int((px[off:off + w * h] != single.pixels.reshape(-1)).sum())
        # the primary G-buffer is per pixel and does not depend on grouping
        for k in ("pos", "nrm", "dif", "spec"):
            assert np.array_equal(batch.gbuffer[k][off:off + w * h], single.gbuffer[k])
        # the camera pass: slot = pixel, view after view
        assert np.array_equal(batch.passes[0]["hits"][off:off + w * h], single.passes[0]["hits"][:w * h])
        total += w * h
        off += w * h
    assert bad <= 1e-4 * total, (bad, total)
    assert batch.passes[0]["rays"] == total
    # the live rays of every pass add up over the views
    for i, p in enumerate(batch.passes):
        s = 0
        for cam, w, h in views:
            s += oracle.render(scene, cam, dataclasses.replace(cfg, width=w, height=h),
                               keep_hits=False, keep_gbuffer=False).passes[i]["rays"]
        assert p["rays"] == s, i
