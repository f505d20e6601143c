"""Autotuner of the batched-views form.

Searches the octahedral hash split (cell bits x direction bits, at most 26
bits below the view id), the hierarchy shape (leaf size x branching factor,
depth 2), the traversal work-item size (candidate clusters per item) and the
traversal mode (the engine's per-pass rule, or one mode for every pass) by
coordinate descent: each dimension in turn, keeping the best value found so
far for the others.  A candidate is timed as the median of graph replays of
the whole batch (render flag bit 3), after the eager frame that captures it.
Every candidate's image is compared with the starting configuration's: the
hierarchy grouping may move a culling decision for a few rays, and the
fraction of identical pixels is reported with each timing.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import RshError

HASH_SPLITS = [(4, 6), (5, 4), (3, 7), (6, 3), (2, 9)]        # 2 + 3c + 2d <= 26
SHAPES = [(8, 8), (4, 16), (16, 4), (32, 2), (8, 4)]          # (leaf, branch); leaf * branch != 64 -> generic kernel
CHUNKS = [64, 16, 256]                                        # candidate clusters per work item
PREFILTERS = [0, 2, 3]          # traversal mode: the per-pass rule, ray pre-test everywhere, + cluster masks


@dataclasses.dataclass
class Trial:
    cell_bits: int
    dir_bits: int
    leaf: int
    branch: int
    traversal_chunk: int
    traversal_prefilter: int
    ms: float
    mrays: float
    same_pixels: float


def time_batch(engine, views, cfg, reps=3):
    """Median wall time (ms) of graph replays of the batch, rays per batch."""
    st = engine.render_views(views, cfg, flags=engine.GRAPH)          # eager frame + capture
    rays = st.rays_total
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        engine.render_views(views, cfg, flags=engine.GRAPH)
        times.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(times)), rays


def tune(engine, scene, views, base_cfg, reps=3, log=None):
    """Coordinate-descent search from base_cfg (cell grid re-fitted to the
    scene for every hash split); returns (best_cfg, [Trial])."""
    engine.render_views(views, base_cfg)
    ref = engine.pixels().copy()
    trials = []
    best = base_cfg
    best_ms = None

    def run(cfg):
        nonlocal best, best_ms
        try:
            ms, rays = time_batch(engine, views, cfg, reps)
        except RshError as err:      # e.g. a work list over capacity for a small chunk: not a candidate
            if log:
                log(f"skipped {cfg.cell_bits}/{cfg.dir_bits} {cfg.leaf}x{cfg.branch} chunk {cfg.traversal_chunk} "
                f"mode {cfg.traversal_prefilter}: {err}")
            return
        same = float((engine.pixels() == ref).mean())
        t = Trial(cfg.cell_bits, cfg.dir_bits, cfg.leaf, cfg.branch, cfg.traversal_chunk, cfg.traversal_prefilter, ms,
                  rays / ms / 1e3, same)
        trials.append(t)
        if log:
            log(t)
        if best_ms is None or ms < best_ms:
            best, best_ms = cfg, ms

    run(base_cfg)
    for c, d in HASH_SPLITS:
        if (c, d) != (best.cell_bits, best.dir_bits):
            run(dataclasses.replace(best, cell_bits=c, dir_bits=d).with_scene_bounds(scene))
    for leaf, branch in SHAPES:
        if (leaf, branch) != (best.leaf, best.branch):
            run(dataclasses.replace(best, leaf=leaf, branch=branch))
    for ch in CHUNKS:
        if ch != best.traversal_chunk:
            run(dataclasses.replace(best, traversal_chunk=ch))
    for pf in PREFILTERS:
        if pf != best.traversal_prefilter:
            run(dataclasses.replace(best, traversal_prefilter=pf))
    return best, trials
