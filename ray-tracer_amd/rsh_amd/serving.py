"""Serving workload: Poisson-arriving ray-query requests against resident scenes.

Four scenes stay on the device (the plain 512x512 scene, the office-scale
boxes, the atrium and a 100k-triangle soup).  A request picks a scene
uniformly, a query kind from the primary / shadow / cosine-diffuse mix
40/30/30 and a size log-uniform in [1k, 4M] rays.  Its rays are a slice of a
per-(scene, kind) pool in page-locked host memory (1M distinct rays tiled to
4M), so request generation costs no time inside the measurement.  Latency is
the server's submit -> results-in-host-memory time.
"""
from __future__ import annotations

import dataclasses
import threading
import time

import numpy as np

from . import config as cf
from . import queries, scenes
from .server import Server, pinned

MIX = ([queries.PRIMARY, queries.SHADOW, queries.DIFFUSE], [0.4, 0.3, 0.3])
MIN_RAYS, MAX_RAYS = 1_000, 4_000_000


def serving_scenes():
    return [scenes.config0_scene(), scenes.config1_scene(), scenes.config2_scene(), scenes.soup_scene()]


def query_config(scene):
    # 25-bit key (2 type + 3 x 5 cell + 2 x 4 direction), interleaved: the best
    # measured split for incoherent queries that leaves 6 bits of request id
    return dataclasses.replace(cf.FrameConfig(), cell_bits=5, dir_bits=4, key_layout=2).with_scene_bounds(scene)


class Pools:
    """Per (scene, query kind) ray pools of MAX_RAYS rays in pinned memory."""

    def __init__(self, scene_list, distinct=1 << 20, seed=0xC0FFEE + 4):
        rng = np.random.default_rng(seed)
        self.rays, self.tmax = {}, {}
        for si, sc in enumerate(scene_list):
            for mix in MIX[0]:
                r, t, _ = queries.make_query(sc, mix, distinct, rng)
                reps = -(-MAX_RAYS // distinct)
                buf = pinned((MAX_RAYS, 6), np.float32)
                buf[:] = np.tile(r, (reps, 1))[:MAX_RAYS]
                self.rays[si, mix] = buf
                if t is not None:
                    tb = pinned((MAX_RAYS,), np.float32)
                    tb[:] = np.tile(t, reps)[:MAX_RAYS]
                    self.tmax[si, mix] = tb

    def request(self, si, mix, n, offset):
        off = min(offset, MAX_RAYS - n)
        r = self.rays[si, mix][off:off + n]
        t = self.tmax[si, mix][off:off + n] if (si, mix) in self.tmax else None
        return r, t, queries.KIND_OF[mix]


def draw_requests(count, nscenes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        si = int(rng.integers(nscenes))
        mix = int(rng.choice(MIX[0], p=MIX[1]))
        n = int(np.exp(rng.uniform(np.log(MIN_RAYS), np.log(MAX_RAYS))))
        out.append((si, mix, n, int(rng.integers(MAX_RAYS))))
    return out


class OutRing:
    """Reusable pinned result buffers (MAX_RAYS uint64 each)."""

    def __init__(self, count):
        self.free = [pinned((MAX_RAYS,), np.uint64) for _ in range(count)]
        self.cv = threading.Condition()

    def get(self):
        with self.cv:
            while not self.free:
                self.cv.wait()
            return self.free.pop()

    def put(self, b):
        with self.cv:
            self.free.append(b)
            self.cv.notify()


def _run(server, sids, pools, ring, reqs, arrivals):
    """Submit reqs at the given arrival offsets (seconds); returns latencies (ms),
    batch sizes, total rays and the wall span (s)."""
    done = []
    q = []
    qcv = threading.Condition()

    def waiter():
        for _ in range(len(reqs)):
            with qcv:
                while not q:
                    qcv.wait()
                rid, buf, n = q.pop(0)
            _, info = server.wait(rid)
            ring.put(buf)
            done.append((info.latency_ms, info.batch_requests, n, time.perf_counter()))

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for (si, mix, n, off), at in zip(reqs, arrivals):
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        rays, tmax, kind = pools.request(si, mix, n, off)
        buf = ring.get()
        out = buf[:n] if kind == 0 else buf.view(np.uint32)[:n]
        rid = server.submit(sids[si], rays, kind, tmax, out=out)
        with qcv:
            q.append((rid, buf, n))
            qcv.notify()
    th.join()
    span = max(d[3] for d in done) - t0
    lat = np.array([d[0] for d in done])
    return lat, np.array([d[1] for d in done]), sum(d[2] for d in done), span


def run(duration_s=8.0, load=0.7, budget=64 << 30, max_batch_rays=16 << 20, slice_rays=1 << 20, seed=0xC0FFEE + 4,
        log=None):
    scene_list = serving_scenes()
    server = Server(0, hbm_budget=budget, max_batch_rays=max_batch_rays)
    server.set_slice(slice_rays)
    sids = [server.add_scene(s, query_config(s)) for s in scene_list]
    pools = Pools(scene_list)
    ring = OutRing(48)
    # warm-up: every scene and kind once
    warm = [(si, mix, 50_000, 0) for si in range(len(scene_list)) for mix in MIX[0]]
    _run(server, sids, pools, ring, warm, [0.0] * len(warm))
    # capacity: a closed burst of requests submitted together
    burst = draw_requests(48, len(scene_list), seed)
    _, bsz, brays, bspan = _run(server, sids, pools, ring, burst, [0.0] * len(burst))
    capacity = brays / bspan
    mean_rays = (MAX_RAYS - MIN_RAYS) / np.log(MAX_RAYS / MIN_RAYS)  # mean of the log-uniform size
    qps = load * capacity / mean_rays
    if log:
        log(f"capacity {capacity / 1e6:.1f} Mrays/s -> {qps:.1f} requests/s at load {load}")
    rng = np.random.default_rng(seed + 1)
    count = max(16, int(qps * duration_s))
    reqs = draw_requests(count, len(scene_list), seed + 2)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, size=count))
    lat, bsz, rays, span = _run(server, sids, pools, ring, reqs, arrivals)
    info = server.stats()
    server.close()
    return {
        "scenes": ["plain 512 (2,060 tris)", "office boxes (30,012 tris)", "atrium (262,604 tris)",
                   "soup (100,000 tris)"],
        "requests": int(count), "qps": float(qps), "load": load,
        "capacity_mrays": capacity / 1e6,
        "mrays": rays / span / 1e6,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_batch_requests": float(bsz.mean()), "max_batch_requests": int(bsz.max()),
        "hbm_budget_gb": budget / (1 << 30),
        "batches": int(info.batches),
        "slice_rays": int(slice_rays),
        "p50_ms_by_size": _by_size(reqs, lat, 50), "p99_ms_by_size": _by_size(reqs, lat, 99),
    }


def _by_size(reqs, lat, q):
    """Latency percentile per request-size decade (1k-10k, ..., 1M-4M)."""
    n = np.array([r[2] for r in reqs])
    out = {}
    for lo, hi, name in ((1e3, 1e4, "1k-10k"), (1e4, 1e5, "10k-100k"), (1e5, 1e6, "100k-1M"), (1e6, 5e6, "1M-4M")):
        sel = (n >= lo) & (n < hi)
        if sel.any():
            out[name] = float(np.percentile(lat[sel], q))
    return out
