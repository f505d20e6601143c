"""Mrays/s of the full RSH frame on one MI355X.

Workload: the atrium frame -- 1920x1080 primary rays over a 262k-triangle
procedural atrium (tessellated columns, arches, tiled floor), 30% reflective
materials, 3 mirror bounces with shadow rays to 1 light at every hit.  One
step = one frame (primary, shadow and bounce passes, each: ray creation +
octahedral hashing, trimming, compression into packets, radix sort,
decompression, hierarchy build, fused traversal + intersection, shading).
Scene and camera are resident in HBM before timing starts.

Prints one JSON line.  "roofline" describes the dominant kernel, the fused
traversal (k_traverse64, one launch per pass): its launch durations come from
HIP events recorded on the engine's stream around every launch of the timed
frames; its algorithmic FLOPs from one extra frame with the traversal
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


def cpu_baseline(scene, cam, cfg, rows):
    """The oracle (C++ restatement, OpenMP over top nodes and pixels) on a
    band of rows of the same frame at full resolution."""
    from oracle import pyoracle
    threads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        threads = min(threads, int(omp))
    y0 = (cfg.height - rows) // 2
    t0 = time.perf_counter()
    fr = pyoracle.render(scene, cam, cfg, threads=threads, keep_hits=False, keep_gbuffer=False,
                         row_range=(y0, y0 + rows))
    dt = time.perf_counter() - t0
    rays = sum(p["rays"] for p in fr.passes)
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {y0}-{y0 + rows} of the 1920x1080 atrium frame (all passes), "
                      f"{rays} rays in {dt:.1f} s"}


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
        st = eng.render(cam, cfg, flags=eng.TIMING)
        times.append(time.perf_counter() - t0)
        rays.append(st.rays_total)
        for i in range(len(stage_sum)):
            stage_sum[i] += st.stage_ms[i]
        for i in range(st.pass_count):
            if st.pass_rays[i]:
                trav_ms += st.pass_traverse_ms[i]
                trav_launches += 1
    rsh_amd.lib().rsh_device_synchronize()
    total = time.perf_counter() - t_all
    value = sum(rays) / total / 1e6
    stages = {name: stage_sum[i] / a.steps for i, name in enumerate(rsh_amd.STAGES)}

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
    roof = {"bound": "valu", "kernel": "k_traverse64<8> (fused traversal + intersection, one launch per pass)",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_per_launch(),
            "avg_launch_ms": avg_launch_ms, "launches_per_frame": launches_per_frame,
            "flops_per_launch": flops_launch}
    # hash + sort + compress bandwidth (algorithmic bytes / stage time)
    hsc_bytes = 0
    for p in passes:
        N, C, S = p["rays"], p["chunks"], p["slots"]
        hsc_bytes += 12 * S + 8 * N + 4 * N + 16 * C + C * (4 + 16 * 4) + 16 * N + 16 * C
    hsc_ms = stages["trim"] + stages["compress"] + stages["sort"] + stages["decompress"]
    hsc = {"achieved": hsc_bytes / (hsc_ms * 1e-3) / 1e9 if hsc_ms > 0 else 0.0, "peak": PEAK_HBM_GBS,
           "unit": "GB/s"}
    hsc["frac"] = hsc["achieved"] / PEAK_HBM_GBS

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
                   "hash": "octahedral 3x5+2x6+2 bits", "hierarchy": "leaf 8, branch 8, depth 2"},
        "roofline": roof,
        "hash_sort_compress": hsc,
        "stage_ms": stages,
        "frame_ms_median": 1e3 * statistics.median(times),
    }
    eng.close()
    if not a.no_cpu:
        rows = a.cpu_rows or max(8, a.height // 4)     # about 10-15 s of oracle work at 1080p
        out["cpu_baseline"] = cpu_baseline(scene, cam, cfg, rows)
    line = json.dumps(out)
    print(line)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
