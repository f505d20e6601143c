"""Mrays/s of the full RSH frame on one MI355X.

Workload: the atrium frame -- 1920x1080 primary rays over a 262k-triangle
procedural atrium (tessellated columns, arches, tiled floor), 30% reflective
materials, 3 mirror bounces with shadow rays to 1 light at every hit.  One
step = one frame (primary, shadow and bounce passes, each: ray creation +
octahedral hashing, trimming, compression into packets, radix sort,
decompression, hierarchy build, fused traversal + intersection, shading).
Scene and camera are resident in HBM before timing starts.

Beside the headline it measures, on the same GPU:
  "batching"  64 views (square, side uniform 128..1024, random poses) of the
              office-scale and atrium scenes: the plain pipeline (one frame
              per view) against the batched form (all views in one pass per
              stage, view id in the key) eager and as a replayed HIP graph;
  "serving"   Poisson-arriving ray-query requests (1k-4M rays, primary /
              shadow / diffuse 40/30/30) against 4 resident scenes at 70% of
              the measured capacity: p50/p99 latency and Mrays/s;
  "soft_shadows", "object_update", "rah_vs_crsh"
              the atrium frame with 16 shadow rays per slot, with every
              object moving each frame, and with the unsorted hierarchy;
  "office_obj", "cornell_obj" (with --office / --cornell)
              the reference's thesis scenes read from its model files.

Prints one JSON line.  "roofline" describes the dominant kernel, the fused
traversal (k_traverse64, one launch per pass): its launch durations come from
HIP events recorded around every traversal launch of three frames rendered
with serial passes (the timed frames run the shadow passes concurrently with
the reflection chain, so their launches share the GPU); its algorithmic FLOPs
from one extra frame with the traversal
counters on (pair and test counts are deterministic); its HBM traffic from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed under profiles/.
"cpu_baseline" is the C++ oracle on a band of rows of the same frame.

--copy-probe runs a plain device copy of known size (the calibration kernel
for the FETCH_SIZE / WRITE_SIZE counters) and exits.
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ray-tracer_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
FLOP_NODE_TEST = 25      # SphereNodeIntersection, RayTracer.cu:283-298
FLOP_RAY_TEST = 46       # RayTriangleIntersection full path, RayTracer.cu:257-280
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_traverse64_atrium1080p_traffic.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=0, help="rows of the CPU sample (0: automatic)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-views leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-update", action="store_true", help="skip the object-update leg")
    ap.add_argument("--no-rah", action="store_true", help="skip the RAH / CRSH A/B leg")
    ap.add_argument("--no-soft", action="store_true", help="skip the soft-shadow leg")
    ap.add_argument("--office", default="", metavar="OBJ",
                    help="also time the reference's office scene (its models/office.obj) through its camera")
    ap.add_argument("--cornell", default="", metavar="DIR",
                    help="also time the reference's Cornell box (Cube.obj, its wall MTL files and "
                         "sphere/ from the reference's models folder DIR) through its camera")
    ap.add_argument("--serve-seconds", type=float, default=6.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="copy MIB MiB device-to-device 3 times with a plain copy kernel and exit")
    ap.add_argument("--out", default="")
    return ap.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert bool((dst[-4:] == src[-4:]).all())
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}))


def traffic_per_launch():
    """HBM bytes per launch of the traversal kernel on this workload, from the
    committed rocprofv3 summary (None when it is absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, cam, cfg, rows):
    """The oracle (C++ restatement, OpenMP over top nodes and pixels) on a
    band of rows of the same frame at full resolution, on every core the
    process may use and on one core (a narrower band)."""
    from oracle import pyoracle
    threads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        threads = min(threads, int(omp))

    def band(r, th):
        y0 = (cfg.height - r) // 2
        t0 = time.perf_counter()
        fr = pyoracle.render(scene, cam, cfg, threads=th, keep_hits=False, keep_gbuffer=False, row_range=(y0, y0 + r))
        dt = time.perf_counter() - t0
        return y0, sum(p["rays"] for p in fr.passes), dt

    y0, rays, dt = band(rows, threads)
    r1 = max(4, rows // max(1, threads))
    y1, rays1, dt1 = band(r1, 1)
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"rows {y0}-{y0 + rows} of the 1920x1080 atrium frame (all passes), "
                      f"{rays} rays in {dt:.1f} s",
            "one_core": {"value": rays1 / dt1 / 1e6, "cores": 1,
                         "sample": f"rows {y1}-{y1 + r1}, {rays1} rays in {dt1:.1f} s"}}


def sort_passes(cfg):
    """8-bit digit passes of the radix sort over the key width of cfg."""
    bits = 32 if cfg.key_mode == 0 or cfg.segment_shift else min(32, cfg.hash_bits())
    return (bits + 7) // 8


def hsc_bytes(first, S, N, C, digit_passes):
    """Algorithmic HBM bytes of one pass's hash, trim, compress, sort and
    decompress stages as the reference structures them (per stage: read its
    inputs, write its outputs, at the reference's buffer layouts): create +
    hash 36 B per slot for primary rays, 68 B per slot where rays come from a
    G-buffer; trim 12 B per slot + 8 B per live ray; compress 4 B per ray + 16 B
    per chunk; sort 4 + 16 B per chunk and digit pass; decompress 16 B per ray
    + 16 B per chunk.  S slots, N live rays, C chunks."""
    create = (36 if first else 68) * S
    return create + (12 * S + 8 * N) + (4 * N + 16 * C) + C * (4 + 16 * digit_passes) + (16 * N + 16 * C)


def hsc_figure(passes, stages, digit_passes):
    """Hash + sort + compress effective bandwidth: the algorithmic bytes of those
    stages over their summed time (events around the stages' kernels; trimming
    and compression run fused, so the compress stage time is inside trim's).
    The engine moves fewer bytes than the reference's stage structure, so this
    is an effective figure."""
    nbytes = sum(hsc_bytes(i == 0, p["slots"], p["rays"], p["chunks"], digit_passes) for i, p in enumerate(passes))
    ms = stages["create"] + stages["trim"] + stages["compress"] + stages["sort"] + stages["decompress"]
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "bytes": int(nbytes), "ms": ms, "kind": "effective (reference stage bytes / time)"}


def soft_leg(rsh_amd, eng, scene, cam, cfg, reps=5):
    """Soft shadows on the atrium frame: 16 rays per shadow slot from a 4x4
    grid about the light, each an exact any-hit ray, shaded by the open
    cells' weights (the reference's compiled-out RayTracer.cu:1550-1603,
    1680-1697).  Median frame time, rays per frame and Mrays/s."""
    import dataclasses
    c = dataclasses.replace(cfg, soft_shadows=1)
    eng.render(cam, c)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        st = eng.render(cam, c)
        times.append(1e3 * (time.perf_counter() - t0))
    ms = statistics.median(times)
    shadow = sum(int(st.pass_rays[i]) for i in range(st.pass_count) if st.pass_kind[i] == 1)
    return {"ms": ms, "rays": int(st.rays_total), "shadow_rays": shadow, "mrays": st.rays_total / ms / 1e3}


def update_leg(rsh_amd, eng, scene, cam, cfg, reps=5):
    """Object bounding spheres and per-frame object updates on the atrium:
    rsh_object_spheres (host, the smallest enclosing balls of all objects),
    rsh_engine_update_objects (every object but the first turned, scaled and
    moved; positions, normals and spheres recomputed from the loaded scene,
    the derived triangle / cluster data rebuilt), and the frame rate with the
    smallest balls as object spheres, and with objects moving every frame."""
    import dataclasses
    import numpy as np
    rg = scene.object_spheres[:, 4:6].astype(np.uint32)
    t0 = time.perf_counter()
    tight = rsh_amd.object_spheres(scene.tri_pos, rg)
    spheres_ms = 1e3 * (time.perf_counter() - t0)
    K, T = scene.object_total, scene.triangle_total
    rng = np.random.default_rng(7)
    mats = []
    for _ in range(reps):
        model, normal, scale = (np.tile(np.eye(4), (K, 1, 1)) for _ in range(3))
        for k in range(1, K):
            c = scene.object_spheres[k, :3].astype(np.float64)
            th = rng.uniform(-0.05, 0.05)
            R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
            model[k, :3, :3] = R
            model[k, :3, 3] = c + rng.uniform(-0.2, 0.2, 3) - R @ c
            normal[k, :3, :3] = R
        mats.append([m.astype(np.float32).reshape(K, 16) for m in (model, normal, scale)])
    times = []
    for m in mats:
        t0 = time.perf_counter()
        eng.update_objects(*m)
        times.append(time.perf_counter() - t0)
    update_ms = 1e3 * statistics.median(times)
    # frames of a scene whose objects move every frame (update + render)
    rays = 0
    t0 = time.perf_counter()
    for m in mats:
        eng.update_objects(*m)
        rays += eng.render(cam, cfg).rays_total
    moving = rays / (time.perf_counter() - t0) / 1e6
    # the same frame with the smallest balls as object spheres
    eng.load_scene(dataclasses.replace(scene, object_spheres=tight))
    eng.render(cam, cfg)
    rays = 0
    t0 = time.perf_counter()
    for _ in range(reps):
        rays += eng.render(cam, cfg).rays_total
    tight_mrays = rays / (time.perf_counter() - t0) / 1e6
    eng.load_scene(scene)
    # algorithmic bytes of the update kernels: vertices (position + normal in
    # and out, object id), triangles (positions in, sphere and edge vectors out)
    upd_bytes = 3 * T * (16 + 16 + 16 + 16 + 4) + T * (48 + 16 + 48)
    return {"object_spheres_ms": spheres_ms, "objects": K, "triangles": T,
            "update_ms": update_ms, "update_bytes": upd_bytes,
            "update_GBps": upd_bytes / (update_ms * 1e-3) / 1e9,
            "moving_objects_mrays": moving, "smallest_ball_spheres_mrays": tight_mrays,
            "sphere_radius_sum_ratio": float(tight[:, 3].sum() / scene.object_spheres[:, 3].sum()),
            "note": "update_ms is the host call (three 16-float matrices per object uploaded, kernels, sync)"}


def office_leg(rsh_amd, eng, path, reps=10):
    """The reference's thesis office scene read from its OBJ / MTL files
    (rsh_amd.objfile) through the reference's camera and light, 1024x1024,
    2 mirror bounces with shadow rays at every hit and reflections from every
    covered pixel (the reference's rule): CRSH and RAH frames."""
    import dataclasses
    from rsh_amd import config as cf, objfile
    sc = objfile.office_scene(path)
    cam = objfile.office_camera(1024, 1024)
    base = dataclasses.replace(cf.config1(1024, 1024), reflect_gate=0).with_scene_bounds(sc)
    eng.load_scene(sc)
    out = {"triangles": sc.triangle_total, "objects": sc.object_total}
    for alg, tag in ((0, "crsh"), (1, "rah")):
        c = dataclasses.replace(base, algorithm=alg)
        eng.render(cam, c)
        n = reps if alg == 0 else 2
        t0 = time.perf_counter()
        for _ in range(n):
            st = eng.render(cam, c)
        ms = 1e3 * (time.perf_counter() - t0) / n
        out[tag + "_ms"], out[tag + "_mrays"], out[tag + "_rays"] = ms, st.rays_total / ms / 1e3, st.rays_total
    # the thesis's office frame: 512x512, shadow rays only (its primary
    # visibility is rasterised; here the primary pass is traced as well)
    c = dataclasses.replace(cf.config1(512, 512), bounces=0).with_scene_bounds(sc)
    cam5 = objfile.office_camera(512, 512)
    eng.render(cam5, c)
    t0 = time.perf_counter()
    for _ in range(reps):
        st = eng.render(cam5, c)
    ms = 1e3 * (time.perf_counter() - t0) / reps
    shadow = sum(st.pass_rays[i] for i in range(st.pass_count) if st.pass_kind[i] == 1)
    out["thesis_512_shadow_frame"] = {"ms": ms, "shadow_rays": int(shadow), "rays": int(st.rays_total),
                                      "mrays": st.rays_total / ms / 1e3}
    return out


def cornell_leg(rsh_amd, eng, models, reps=20):
    """The reference's Cornell box (rsh_amd.objfile.cornell_scene) through its
    camera, as its recorded run traced it: 512x512, one shadow pass, then two
    mirror bounces from every covered pixel with no bounce shadows (its
    primary visibility is rasterised; the primary pass is traced here as
    well).  CRSH and RAH, each the median of `reps` frames."""
    import dataclasses
    from rsh_amd import config as cf, objfile
    sc = objfile.cornell_scene(models)
    cam = objfile.cornell_camera(512, 512)
    base = dataclasses.replace(cf.config1(512, 512), reflect_gate=0, bounces=2,
                               bounce_shadows=0).with_scene_bounds(sc)
    eng.load_scene(sc)
    out = {"triangles": sc.triangle_total, "objects": sc.object_total}
    for alg, tag in ((0, "crsh"), (1, "rah")):
        c = dataclasses.replace(base, algorithm=alg)
        eng.render(cam, c)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            st = eng.render(cam, c)
            times.append(1e3 * (time.perf_counter() - t0))
        ms = statistics.median(times)
        rays = [int(st.pass_rays[i]) for i in range(st.pass_count)]
        out[tag] = {"ms": ms, "mrays": st.rays_total / ms / 1e3, "pass_rays": rays,
                    "non_primary_rays": int(sum(rays[1:]))}
    return out


def rah_leg(rsh_amd, eng, scene2, cam2, cfg2, reps=3):
    """The reference's two algorithms on the same frames: CRSH (sorted rays,
    object culling) against RAH (rays in slot order, trimmed only, every
    object a candidate; main.cpp:1016-1037), on the office-scale 1024^2 frame
    and the atrium frame.  Reports ms, Mrays/s and the share of identical
    pixels: the two group rays into different nodes, and the reference's node
    tests are not exactly conservative, so a few pixels can differ (each
    algorithm matches the oracle's version of itself)."""
    import dataclasses
    from rsh_amd import config as cf, scenes
    out = {}
    office = scenes.config1_scene()
    for name, scene, cam, cfg in (("office 1024x1024", office, scenes.config1_camera(1024, 1024),
                                   cf.config1(1024, 1024).with_scene_bounds(office)),
                                  ("atrium 1920x1080", scene2, cam2, cfg2)):
        eng.load_scene(scene)
        row, px = {}, None
        for alg, tag in ((0, "crsh"), (1, "rah")):
            c = dataclasses.replace(cfg, algorithm=alg)
            eng.render(cam, c)
            t0 = time.perf_counter()
            for _ in range(reps):
                st = eng.render(cam, c)
            ms = 1e3 * (time.perf_counter() - t0) / reps
            row[tag + "_ms"] = ms
            row[tag + "_mrays"] = st.rays_total / ms / 1e3
            p = eng.pixels()
            if px is None:
                px = p.copy()
            else:
                row["identical_pixels"] = float((p == px).mean())
        row["crsh_speedup"] = row["rah_ms"] / row["crsh_ms"]
        out[name] = row
    eng.load_scene(scene2)
    return out


def batching_leg(rsh_amd, eng, scene2):
    """Plain per-view frames against the batched form on the 64-view batch."""
    import dataclasses
    from rsh_amd import autotune, config as cf, scenes
    views = scenes.batch_views()
    out = {"views": len(views), "pixels": int(sum(w * h for _, w, h in views)),
           "key": "2 type + 3x4 cell + 2x6 direction bits interleaved, view id in bits 26-31"}
    for name, scene, base in (("office", scenes.config1_scene(), cf.config1()),
                              ("atrium", scene2, cf.config2())):
        eng.load_scene(scene)
        cfg = dataclasses.replace(base, cell_bits=4, dir_bits=6, segment_shift=26).with_scene_bounds(scene)
        for cam, w, h in views[:2]:
            eng.render(cam, dataclasses.replace(cfg, width=w, height=h))
        t0 = time.perf_counter()
        rays, imgs = 0, []
        for cam, w, h in views:
            rays += eng.render(cam, dataclasses.replace(cfg, width=w, height=h)).rays_total
            imgs.append(eng.pixels().copy())
        plain = time.perf_counter() - t0
        eng.render_views(views, cfg)
        t0 = time.perf_counter()
        st = eng.render_views(views, cfg)
        eager = time.perf_counter() - t0
        parts = eng.split_views(eng.pixels(), views)
        same = sum(int((a == b).sum()) for a, b in zip(parts, imgs)) / out["pixels"]
        graph_ms, brays = autotune.time_batch(eng, views, cfg, reps=3)
        st = eng.render_views(views, cfg, flags=eng.TIMING | eng.SERIAL)
        stages = {n: st.stage_ms[i] for i, n in enumerate(rsh_amd.STAGES)}
        bp = [{"slots": st.pass_slots[i], "rays": st.pass_rays[i], "chunks": st.pass_chunks[i]}
              for i in range(st.pass_count)]
        hsc = hsc_figure(bp, stages, sort_passes(cfg))
        out[name] = {"plain_mrays": rays / plain / 1e6, "plain_ms": plain * 1e3,
                     "batched_mrays": st.rays_total / eager / 1e6, "batched_ms": eager * 1e3,
                     "graph_mrays": brays / graph_ms / 1e3, "graph_ms": graph_ms,
                     "speedup": (brays / graph_ms / 1e3) / (rays / plain / 1e6),
                     "rays": int(brays), "identical_pixels_vs_plain": same,
                     "hash_sort_compress": hsc}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    import rsh_amd
    from rsh_amd import config as cf, scenes

    scene = scenes.config2_scene()
    cam = scenes.config2_camera(a.width, a.height)
    cfg = cf.config2(a.width, a.height).with_scene_bounds(scene)
    eng = rsh_amd.Engine(0)
    eng.load_scene(scene)
    for _ in range(a.warmup):
        eng.render(cam, cfg)
    rsh_amd.lib().rsh_device_synchronize()
    times, rays = [], []
    stage_sum = [0.0] * len(rsh_amd.STAGES)
    trav_ms, trav_launches = 0.0, 0
    t_all = time.perf_counter()
    for _ in range(a.steps):
        t0 = time.perf_counter()
        # the frame as served: shadow passes concurrent with the reflection chain
        st = eng.render(cam, cfg)
        times.append(time.perf_counter() - t0)
        rays.append(st.rays_total)
    rsh_amd.lib().rsh_device_synchronize()
    total = time.perf_counter() - t_all
    value = sum(rays) / total / 1e6
    # The traversal kernel's launch time (the roofline): HIP events around each
    # traversal launch only, in frames whose passes run one at a time, so no
    # other kernel shares the GPU with the timed launch.
    for _ in range(3):
        st = eng.render(cam, cfg, flags=eng.TRAV_TIMING | eng.SERIAL)
        for i in range(st.pass_count):
            if st.pass_rays[i]:
                trav_ms += st.pass_traverse_ms[i]
                trav_launches += 1
    # Stage breakdown: events around every stage (their own overhead stays out
    # of the timed frames), serial passes as well.
    for _ in range(3):
        st = eng.render(cam, cfg, flags=eng.TIMING | eng.SERIAL)
        for i in range(len(stage_sum)):
            stage_sum[i] += st.stage_ms[i]
    stages = {name: stage_sum[i] / 3 for i, name in enumerate(rsh_amd.STAGES)}

    # Algorithmic work of the traversal: one frame with the counters on.
    eng.render(cam, cfg, flags=eng.COUNTERS)
    passes = eng.passes()
    n_tests = sum(p["ray_tests"] for p in passes)
    n_top = sum(p["pairs_top"] for p in passes)
    n_surv = sum(p["top_survivors"] for p in passes)
    launches_per_frame = sum(1 for p in passes if p["rays"])
    flops_frame = FLOP_RAY_TEST * n_tests + FLOP_NODE_TEST * (n_top + cfg.branch * n_surv)
    avg_launch_ms = trav_ms / trav_launches if trav_launches else 0.0
    flops_launch = flops_frame / launches_per_frame if launches_per_frame else 0.0
    achieved = flops_launch / (avg_launch_ms * 1e-3) / 1e12 if avg_launch_ms > 0 else 0.0
    roof = {"bound": "valu", "kernel": "k_traverse64<8, MODE> (fused traversal + intersection, one launch per pass)",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_per_launch(),
            "avg_launch_ms": avg_launch_ms, "launches_per_frame": launches_per_frame,
            "flops_per_launch": flops_launch}
    hsc = hsc_figure(passes, stages, sort_passes(cfg))
    hsc["workload"] = "the timed atrium frames (8 passes, %d rays per frame)" % int(statistics.median(rays))

    progress(f"frames: {value:.1f} Mrays/s")
    soft = None if a.no_soft else soft_leg(rsh_amd, eng, scene, cam, cfg)
    if soft:
        progress("soft-shadow leg done")
    dynamic = None if a.no_update else update_leg(rsh_amd, eng, scene, cam, cfg)
    if dynamic:
        progress("object-update leg done")
    rah = None if a.no_rah else rah_leg(rsh_amd, eng, scene, cam, cfg)
    if rah:
        progress("RAH A/B leg done")
    office = office_leg(rsh_amd, eng, a.office) if a.office else None
    cornell = cornell_leg(rsh_amd, eng, a.cornell) if a.cornell else None
    eng.load_scene(scene)
    if not a.no_batch:
        batching = batching_leg(rsh_amd, eng, scene)
        progress("batching leg done")
    out = {
        "metric": "Mrays/s end-to-end at 262k tris 1080p",
        "value": value,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * total / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural atrium)",
        "config": {"workload": f"atrium {a.width}x{a.height}, {scene.triangle_total} tris, "
                               f"{scene.object_total} objects, 3 bounces + shadows, 30% reflective",
                   "rays_per_frame": int(statistics.median(rays)),
                   "hash": "octahedral 3x5+2x6+2 bits, cell and direction bits interleaved",
                   "hierarchy": "leaf 8, branch 8, depth 2"},
        "roofline": roof,
        "hash_sort_compress": hsc,
        "stage_ms": stages,
        "frame_ms_median": 1e3 * statistics.median(times),
    }
    eng.close()
    if soft:
        out["soft_shadows"] = soft
    if dynamic:
        out["object_update"] = dynamic
    if rah:
        out["rah_vs_crsh"] = rah
    if office:
        out["office_obj"] = office
    if cornell:
        out["cornell_obj"] = cornell
    if not a.no_batch:
        out["batching"] = batching
    if not a.no_serve:
        from rsh_amd import serving
        out["serving"] = serving.run(duration_s=a.serve_seconds)
        progress("serving leg done")
    if not a.no_cpu:
        rows = a.cpu_rows or a.height      # the whole frame on all cores (~6 s on 16), a band on one (~10 s)
        out["cpu_baseline"] = cpu_baseline(scene, cam, cfg, rows)
        progress("cpu baseline done")
    line = json.dumps(out)
    print(line)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
