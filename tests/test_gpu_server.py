"""Serving loop (rsh_server_*): concurrent requests against resident scenes
are merged into shared batches, and every request's result is the same
whether it runs alone or in the mix (closest-hit ids and any-hit flags on at
least 99.99% of rays; identical is expected).  Admission holds requests back
under a small HBM budget and still completes them."""
import dataclasses

import numpy as np
import pytest

from rsh_amd import queries

pytestmark = pytest.mark.gpu


def scfg(rsh, scene):
    return dataclasses.replace(rsh.config.config0(64, 64), cell_bits=4, dir_bits=6).with_scene_bounds(scene)


def make_requests(rsh, scenes, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        s = int(rng.integers(len(scenes)))
        mix = int(rng.choice([queries.PRIMARY, queries.SHADOW, queries.DIFFUSE], p=[0.4, 0.3, 0.3]))
        n = int(np.exp(rng.uniform(np.log(1000), np.log(60000))))
        rays, tmax, kind = queries.make_query(scenes[s], mix, n, rng)
        reqs.append((s, rays, tmax, kind))
    return reqs


def hit_ids(res, kind):
    return res if kind == 1 else (res & np.uint64(0xFFFFFFFF))


def test_alone_equals_mixed(rsh):
    from rsh_amd.server import Server
    scenes = [rsh.scenes.config0_scene(), rsh.scenes.config1_scene(boxes=300)]
    srv = Server(0, hbm_budget=8 << 30, max_batch_rays=1 << 22)
    sids = [srv.add_scene(s, scfg(rsh, s)) for s in scenes]
    reqs = make_requests(rsh, scenes, 24, 1)
    alone = []
    for s, rays, tmax, kind in reqs:
        res, info = srv.wait(srv.submit(sids[s], rays, kind, tmax))
        assert info.batch_requests == 1
        alone.append(res.copy())
    before = srv.stats()
    ids = [srv.submit(sids[s], rays, kind, tmax) for s, rays, tmax, kind in reqs]
    merged = 0
    for rid, (s, rays, tmax, kind), a in zip(ids, reqs, alone):
        res, info = srv.wait(rid)
        merged = max(merged, info.batch_requests)
        same = (hit_ids(res, kind) == hit_ids(a, kind)).mean()
        assert same >= 0.9999, (rid, same)
    after = srv.stats()
    assert merged > 1
    assert after.batches - before.batches < len(reqs)
    srv.close()


def test_server_results_match_engine_trace(rsh, oracle):
    """A request served alone is the engine's (and the oracle's) ray query."""
    from rsh_amd.server import Server
    scene = rsh.scenes.config0_scene()
    cfg = scfg(rsh, scene)
    srv = Server(0, hbm_budget=1 << 30, max_batch_rays=1 << 20)
    sid = srv.add_scene(scene, cfg)
    rng = np.random.default_rng(3)
    rays, tmax, kind = queries.make_query(scene, queries.SHADOW, 5000, rng)
    res, _ = srv.wait(srv.submit(sid, rays, kind, tmax))
    seg = np.zeros(len(rays), np.uint32)
    want = oracle.trace_rays(scene, rays, kind, dataclasses.replace(cfg, segment_shift=26), tmax=tmax, segments=seg)
    np.testing.assert_array_equal(res.astype(np.uint64), want)
    srv.close()


def test_admission_under_small_budget(rsh):
    from rsh_amd.server import Server
    scene = rsh.scenes.config0_scene()
    srv = Server(0, hbm_budget=256 * 40000, max_batch_rays=40000)     # room for one 40k-ray request at a time
    sid = srv.add_scene(scene, scfg(rsh, scene))
    rng = np.random.default_rng(4)
    reqs = [queries.make_query(scene, queries.DIFFUSE, 30000, rng) for _ in range(6)]
    ids = [srv.submit(sid, r, k, t) for r, t, k in reqs]
    for rid in ids:
        res, info = srv.wait(rid)
        assert info.batch_requests == 1          # the budget admits one at a time
        assert (res != np.uint64(0xFFFFFFFFFFFFFFFF)).mean() > 0.5
    with pytest.raises(rsh.RshError):
        srv.submit(sid, np.zeros((50000, 6), np.float32), 0)          # larger than the budget
    srv.close()


def test_alone_equals_mixed_four_scenes(rsh):
    """The serving workload's four resident scenes (512 scene, office-scale
    boxes, the 262k-triangle atrium, the 100k-triangle soup), requests of
    1k-60k rays plus one of 1.2M rays on the atrium: each result is the same
    alone and in the mix."""
    from rsh_amd.server import Server
    scenes = [rsh.scenes.config0_scene(), rsh.scenes.config1_scene(), rsh.scenes.config2_scene(),
              rsh.scenes.soup_scene()]
    srv = Server(0, hbm_budget=16 << 30, max_batch_rays=1 << 22)
    sids = [srv.add_scene(s, scfg(rsh, s)) for s in scenes]
    reqs = make_requests(rsh, scenes, 20, 2)
    rng = np.random.default_rng(22)
    big = queries.make_query(scenes[2], queries.DIFFUSE, 1_200_000, rng)
    reqs.insert(7, (2, big[0], big[1], big[2]))
    assert {s for s, _, _, _ in reqs} == {0, 1, 2, 3}
    alone = []
    for s, rays, tmax, kind in reqs:
        res, info = srv.wait(srv.submit(sids[s], rays, kind, tmax))
        alone.append(res.copy())
    ids = [srv.submit(sids[s], rays, kind, tmax) for s, rays, tmax, kind in reqs]
    merged = 0
    for rid, (s, rays, tmax, kind), a in zip(ids, reqs, alone):
        res, info = srv.wait(rid)
        merged = max(merged, info.batch_requests)
        same = (hit_ids(res, kind) == hit_ids(a, kind)).mean()
        assert same >= 0.9999, (rid, s, len(rays), same)
    assert merged > 1
    srv.close()


def test_sliced_requests_equal_alone(rsh):
    """With a slice of 8k rays, a 50k-ray request is traced over several
    batches shared with the small requests queued behind it; every result
    equals the request traced alone in one slice."""
    from rsh_amd.server import Server
    scene = rsh.scenes.config1_scene(boxes=300)
    cfg = scfg(rsh, scene)
    srv = Server(0, hbm_budget=8 << 30, max_batch_rays=1 << 22)
    sid = srv.add_scene(scene, cfg)
    rng = np.random.default_rng(31)
    reqs = [queries.make_query(scene, queries.DIFFUSE, 50_000, rng)] + \
           [queries.make_query(scene, queries.DIFFUSE, int(n), rng) for n in rng.integers(1000, 6000, 12)]
    alone = [srv.wait(srv.submit(sid, r, k, t))[0].copy() for r, t, k in reqs]
    srv.set_slice(8192)
    ids = [srv.submit(sid, r, k, t) for r, t, k in reqs]
    infos = []
    for rid, (r, t, k), a in zip(ids, reqs, alone):
        res, info = srv.wait(rid)
        infos.append(info)
        assert (hit_ids(res, k) == hit_ids(a, k)).mean() >= 0.9999, rid
    assert infos[0].batch_requests > 1          # the big request shared batches
    with pytest.raises(rsh.RshError):
        srv.set_slice(0)
    srv.close()
