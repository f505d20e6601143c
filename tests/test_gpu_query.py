"""Ray queries (rsh_engine_trace) against the oracle's restatement
(oracle_trace_rays): closest-hit records and any-hit flags bit for bit, with
and without segment ids, for camera, shadow and cosine-diffuse rays."""
import dataclasses

import numpy as np
import pytest

from rsh_amd import queries

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(rsh):
    e = rsh.Engine(0)
    yield e
    e.close()


def qcfg(rsh, scene, **kw):
    return dataclasses.replace(rsh.config.config0(64, 64), cell_bits=4, dir_bits=6, segment_shift=26,
                               **kw).with_scene_bounds(scene)


_SCENES = {}


def named_scene(rsh, name):
    if name not in _SCENES:
        _SCENES[name] = {"config0": lambda: rsh.scenes.config0_scene(),
                         "config1": lambda: rsh.scenes.config1_scene(boxes=200),
                         "soup": lambda: rsh.scenes.soup_scene()}[name]()
    return _SCENES[name]


@pytest.mark.parametrize("mix", [queries.PRIMARY, queries.SHADOW, queries.DIFFUSE])
@pytest.mark.parametrize("scene_name", ["config0", "config1", "soup"])
def test_query_matches_oracle(engine, oracle, rsh, mix, scene_name):
    """Every query kind of the serving mix, including the serving workload's
    100k-triangle random soup."""
    scene = named_scene(rsh, scene_name)
    rng = np.random.default_rng(100 + mix)
    rays, tmax, kind = queries.make_query(scene, mix, 20000, rng)
    cfg = qcfg(rsh, scene)
    engine.load_scene(scene)
    got = engine.trace(rays, kind, cfg, tmax=tmax)
    want = oracle.trace_rays(scene, rays, kind, cfg, tmax=tmax)
    np.testing.assert_array_equal(got.astype(np.uint64), want)
    if kind == 0:
        assert (want != np.uint64(0xFFFFFFFFFFFFFFFF)).mean() > 0.5      # most queries hit something


@pytest.mark.parametrize("variant", [dict(), dict(algorithm=1), dict(leaf=4, branch=16), dict(hit_mode=1),
                                     dict(key_layout=2)])
def test_segmented_query_matches_oracle(engine, oracle, rsh, variant):
    scene = rsh.scenes.config0_scene()
    rng = np.random.default_rng(7)
    rays, _, _ = queries.make_query(scene, queries.DIFFUSE, 15000, rng)
    seg = np.repeat(np.arange(5, dtype=np.uint32), 3000)
    cfg = qcfg(rsh, scene, **variant)
    engine.load_scene(scene)
    got = engine.trace(rays, 0, cfg, segments=seg)
    np.testing.assert_array_equal(got, oracle.trace_rays(scene, rays, 0, cfg, segments=seg))


def test_query_rejects_bad_arguments(engine, rsh):
    scene = rsh.scenes.config0_scene()
    engine.load_scene(scene)
    rays = np.zeros((4, 6), np.float32)
    with pytest.raises(rsh.RshError):
        engine.trace(rays, 1, qcfg(rsh, scene))                      # any hit without tmax
    with pytest.raises(rsh.RshError):
        engine.trace(rays, 0, dataclasses.replace(rsh.config.config0(8, 8), segment_shift=0),
                     segments=np.zeros(4, np.uint32))                # segments without a shift
