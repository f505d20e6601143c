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
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner run in the batching leg")
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
    ap.add_argument("--hsc-profile", default="", choices=["", "frame", "office", "atrium"],
                    help="render only this workload --steps times (for a rocprofv3 kernel trace or counter run "
                         "of the hash, sort and compress kernels); write its pass sizes as JSON to --out")
    ap.add_argument("--hsc-summary", nargs=2, default=None, metavar=("KERNEL_STATS_CSV", "SIZES_JSON"),
                    help="(no GPU) summarise a rocprofv3 --stats kernel_stats.csv of a --hsc-profile run")
    ap.add_argument("--fetch-csv", default="", help="with --hsc-summary: FETCH_SIZE counter_collection.csv")
    ap.add_argument("--write-csv", default="", help="with --hsc-summary: WRITE_SIZE counter_collection.csv")
    ap.add_argument("--fetch-scale", type=float, default=2.0,
                    help="FETCH_SIZE calibration: bytes read per counted byte on a plain copy")
    ap.add_argument("--counter-renders", type=int, default=0,
                    help="with --fetch-csv: renders in the counter runs (default: those of SIZES_JSON)")
    ap.add_argument("--trav-summary", nargs=2, default=None, metavar=("KERNEL_STATS_CSV", "KERNEL_TRACE_CSV"),
                    help="(no GPU) summarise a rocprofv3 kernel trace of this bench run with the same --steps and "
                         "--warmup (and the --no-* legs off): kernel table, traversal launches per frame and their "
                         "average over the serial frames the roofline times")
    ap.add_argument("--trav-traffic", nargs=4, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="(no GPU) HBM traffic per k_traverse64 launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs "
                         "of this bench (--steps 1 --warmup 0, frame legs only), scaled by the same counters on "
                         "--copy-probe 512")
    ap.add_argument("--trav-valu", default="", metavar="SQ_CSV",
                    help="(no GPU) SQ counters per k_traverse64 launch of the first serial frame")
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
        return y0, sum(p["rays"] for p in fr.passes), dt, fr.stage_ms

    y0, rays, dt, stage_ms = band(rows, threads)
    r1 = max(4, rows // max(1, threads))
    y1, rays1, dt1, stage_ms1 = band(r1, 1)
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"rows {y0}-{y0 + rows} of the 1920x1080 atrium frame (all passes), "
                      f"{rays} rays in {dt:.1f} s",
            "stage_ms": stage_ms,
            "stage_clock": "std::chrono::steady_clock per stage inside the oracle (TestManager.cpp:284-309's "
                           "per-stage timers); compare with stage_ms of the GPU frame",
            "one_core": {"value": rays1 / dt1 / 1e6, "cores": 1,
                         "sample": f"rows {y1}-{y1 + r1}, {rays1} rays in {dt1:.1f} s", "stage_ms": stage_ms1}}


def sort_digit_passes(cfg):
    """Radix digit passes the engine's sort runs on a shadow or reflection
    pass (run_radix_sort): octahedral keys without a view id sort the bits
    below their common type bits, reference keys and keys with a view id all
    32; 9-bit digits where they save a pass (27 bits: three passes)."""
    if cfg.key_mode == 1 and not cfg.segment_shift:
        bits = min(32, cfg.hash_bits()) - 2
    else:
        bits = 32
    rb = 9 if -(-bits // 9) < -(-bits // 8) else 8
    return -(-bits // rb)


def hsc_bytes(p, first, cfg, digit_passes):
    """Algorithmic HBM bytes of one pass's hash, trim, compress, sort and
    decompress stages (the survey's per-stage definitions, counted only for
    the stages the engine runs).  The camera pass of a CRSH frame creates its
    rays in key order straight into the hierarchy records: 36 B per slot and
    nothing else.  Other passes: create + hash 68 B per slot (from the
    G-buffer); trim 12 B per slot + 8 B per live ray; compress 4 B per ray +
    16 B per chunk; sort 4 + 16 B per chunk and digit pass actually run;
    decompress 16 B per ray + 16 B per chunk.  S slots, N live rays, C chunks."""
    S, N, C = int(p["slots"]), int(p["rays"]), int(p["chunks"])
    if first and cfg.algorithm == 0:
        return 36 * S
    return 68 * S + (12 * S + 8 * N) + (4 * N + 16 * C) + C * (4 + 16 * digit_passes) + (16 * N + 16 * C)


def hsc_total_bytes(passes, cfg):
    P = sort_digit_passes(cfg)
    return sum(hsc_bytes(p, i == 0, cfg, P) for i, p in enumerate(passes))


def hsc_figure(passes, stages, cfg, name):
    """Hash + sort + compress bandwidth: the algorithmic bytes of those stages
    over their summed time (HIP events around the stages' kernels, passes
    serial; trimming and compression run fused, so the compress stage time is
    inside trim's)."""
    nbytes = hsc_total_bytes(passes, cfg)
    ms = stages["create"] + stages["trim"] + stages["compress"] + stages["sort"] + stages["decompress"]
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "bytes": int(nbytes), "ms": ms, "digit_passes": sort_digit_passes(cfg),
            "kind": "algorithmic bytes of the stages that run / HIP-event stage time",
            "frac_rocprof": hsc_profile_frac(name)}


HSC_KERNELS = ("k_create_primary_sorted", "k_create_shadow", "k_create_reflection", "k_create_bounce",
               "k_pack_tiles", "k_pack_scan", "k_pack_write", "k_radix_upsweep", "k_radix_rowscan",
               "k_radix_downsweep", "k_unpack_sizes", "k_unpack_scan", "k_unpack_offsets", "k_unpack_expand")
HSC_PROFILES = {"frame": "hsc_atrium1080p_frame.txt", "office": "hsc_office_64views.txt",
                "atrium": "hsc_atrium_64views.txt"}


def hsc_profile_frac(name):
    """The rocprofv3-based fraction committed under profiles/ for this
    workload: the same algorithmic bytes over the summed kernel durations of
    a kernel trace (None when absent)."""
    try:
        with open(os.path.join(ROOT, "profiles", HSC_PROFILES[name])) as f:
            m = re.search(r"^hsc_frac_rocprof\s*=\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except (OSError, KeyError):
        return None


def hsc_workload(name, rsh_amd):
    """(engine, render callable, cfg) of a hash / sort / compress workload:
    the atrium 1080p frame, or the 64-view batch of the office-scale or
    atrium scene."""
    import dataclasses
    from rsh_amd import config as cf, scenes
    eng = rsh_amd.Engine(0)
    if name == "frame":
        scene = scenes.config2_scene()
        cam = scenes.config2_camera(1920, 1080)
        cfg = cf.config2(1920, 1080).with_scene_bounds(scene)
        eng.load_scene(scene)
        # serial passes: no kernel of the other lane shares the GPU with a timed one
        return eng, (lambda: eng.render(cam, cfg, flags=eng.SERIAL)), cfg
    views = scenes.batch_views()
    scene, base = (scenes.config1_scene(), cf.config1()) if name == "office" else (scenes.config2_scene(), cf.config2())
    cfg = dataclasses.replace(base, cell_bits=4, dir_bits=6, segment_shift=26).with_scene_bounds(scene)
    eng.load_scene(scene)
    return eng, (lambda: eng.render_views(views, cfg, flags=eng.SERIAL)), cfg


def hsc_profile_run(a):
    """--hsc-profile: the workload rendered --steps times and nothing else, so
    that a kernel trace holds exactly those renders; pass sizes to --out."""
    import dataclasses
    import rsh_amd
    eng, render, cfg = hsc_workload(a.hsc_profile, rsh_amd)
    st = None
    for i in range(a.steps):
        st = render()
        progress(f"{a.hsc_profile} render {i + 1}/{a.steps}")
    passes = [{"kind": int(st.pass_kind[i]), "slots": int(st.pass_slots[i]), "rays": int(st.pass_rays[i]),
               "chunks": int(st.pass_chunks[i])} for i in range(st.pass_count)]
    out = {"workload": a.hsc_profile, "renders": a.steps, "passes": passes,
           "cfg": {k: v for k, v in dataclasses.asdict(cfg).items() if isinstance(v, (int, float))},
           "bytes_per_render": hsc_total_bytes(passes, cfg), "digit_passes": sort_digit_passes(cfg)}
    eng.close()
    line = json.dumps(out)
    print(line)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


def _kernel_base(name):
    """The kernel's own name (k_...) inside a demangled rocprofv3 name such as
    'void rsh::(anonymous namespace)::k_radix_upsweep<9u>(unsigned int const*, ...)'."""
    m = re.search(r"\b(k_[A-Za-z0-9_]+)", name)
    return m.group(1) if m else name.strip()


def trav_summary(a):
    """--trav-summary: the kernel table of a rocprofv3 --stats run of this
    bench (frame legs only), the k_traverse64 launches grouped by frame in
    start order, and their average over the three serial frames with
    traversal timers, which the roofline's avg_launch_ms times with HIP
    events.  Frames in a run: --warmup warm-up, --steps timed, one eager
    frame recorded for the graph, 1 + --steps graph replays, then 3 serial
    frames with traversal timers, 3 with stage timers, 1 with counters."""
    import csv
    stats_csv, trace_csv = a.trav_summary
    rows = []
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            m = re.search(r"\b(k_[A-Za-z0-9_]+(<[^>]*>)?)", r["Name"])
            rows.append((m.group(1) if m else r["Name"][:40], int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                         float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3,
                         float(r["Percentage"])))
    rows.sort(key=lambda t: -t[2])
    print(f"{'kernel':28s} {'calls':>6s} {'total_us':>12s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'pct':>7s}")
    for name, c, tot, avg, mn, mx, pct in rows[:30]:
        print(f"{name[:28]:28s} {c:6d} {tot:12.1f} {avg:10.1f} {mn:10.1f} {mx:10.1f} {pct:7.2f}")
    launches = []
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            if "k_traverse64" in r["Kernel_Name"]:
                launches.append((int(r["Start_Timestamp"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6))
    launches.sort()
    counts, first = _frame_layout(len(launches), a.steps, a.warmup)
    frames = _split_frames(launches, counts)
    print()
    print("k_traverse64<8, MODE> per launch (ms), one row per frame, in start order")
    for i, fr in enumerate(frames):
        tag = "  <- serial, traversal timers" if first <= i < first + 3 else ""
        print("  " + " ".join(f"{d:6.2f}" for _, d in fr) + f"   sum {sum(d for _, d in fr):5.2f}{tag}")
    serial = [sum(d for _, d in ps) for fr in frames[first:first + 3] for ps in _passes_of(fr)]
    print()
    print(f"k_traverse64 average launch: {sum(d for _, d in launches) / max(len(launches), 1):.3f} ms over "
          f"{len(launches)} launches")
    print(f"k_traverse64 time per traced pass over the 3 serial frames with traversal timers (the camera pass's "
          f"two launches together): {sum(serial) / max(len(serial), 1):.3f} ms ({len(serial)} passes)")


def _counter_rows(path, kernel=None):
    """{dispatch: {counter: value}} and {dispatch: kernel name} of a
    rocprofv3 counter_collection.csv (optionally one kernel only)."""
    import csv
    vals, names = {}, {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if kernel and kernel not in r["Kernel_Name"]:
                continue
            d = int(r["Dispatch_Id"])
            vals.setdefault(d, {})[r["Counter_Name"]] = vals.get(d, {}).get(r["Counter_Name"], 0.0) + \
                float(r["Counter_Value"])
            names[d] = r["Kernel_Name"]
    return vals, names


PASS_NAMES = ["primary", "shadow1", "refl1", "shadow2", "refl2", "shadow3", "refl3", "shadow4"]
FRAME_PASSES = len(PASS_NAMES)


def _frame_layout(n_launches, steps, warmup):
    """k_traverse64 launches per frame of a bench run (the --no-* legs off)
    and the index of its first serial frame with traversal timers.  Frames,
    in order: --warmup warm-up, --steps timed, one eager frame recorded for
    the graph, 1 + --steps graph replays, 3 serial frames with traversal
    timers, 3 with stage timers, 1 with counters.  The camera pass runs as
    two launches when it goes front to back (every frame but the counting
    one), so the other frames have 8 or 9 launches alike."""
    frames = warmup + steps + 1 + (1 + steps) + 3 + 3 + 1
    per = (n_launches - FRAME_PASSES) // (frames - 1)
    return [per] * (frames - 1) + [FRAME_PASSES], warmup + 2 * steps + 2


def _split_frames(items, counts):
    out, i = [], 0
    for c in counts:
        out.append(items[i:i + c])
        i += c
    return out


def _passes_of(frame):
    """A frame's launches grouped by pass: a front-to-back camera pass's two
    launches together, then one launch per pass."""
    if len(frame) == FRAME_PASSES + 1:
        return [frame[0:2]] + [[x] for x in frame[2:]]
    return [[x] for x in frame]


def _first_serial_frame(dispatches, steps=1, warmup=0):
    """The k_traverse64 dispatches of the first serial frame with traversal
    timers of a bench run (frame order as in trav_summary), grouped by pass
    in pass order: primary, then shadow and reflection passes alternating
    (in the timed frames the two lanes' launches interleave differently)."""
    d = sorted(dispatches)
    counts, first = _frame_layout(len(d), steps, warmup)
    return _passes_of(_split_frames(d, counts)[first])


def trav_traffic(a):
    fetch_csv, write_csv, cf_csv, cw_csv = a.trav_traffic
    fv, _ = _counter_rows(fetch_csv, "k_traverse64")
    wv, _ = _counter_rows(write_csv, "k_traverse64")
    cfv, _ = _counter_rows(cf_csv, "copyBuffer")
    cwv, _ = _counter_rows(cw_csv, "copyBuffer")
    copy_kib = 512 * 1024
    big_f = [v["FETCH_SIZE"] for v in cfv.values() if v.get("FETCH_SIZE", 0) > copy_kib / 4]
    big_w = [v["WRITE_SIZE"] for v in cwv.values() if v.get("WRITE_SIZE", 0) > copy_kib / 4]
    fscale = copy_kib / (sum(big_f) / len(big_f))
    wscale = copy_kib / (sum(big_w) / len(big_w))
    print(f"copy FETCH_SIZE reads {sum(big_f) / len(big_f):.0f} KiB for {copy_kib} KiB read  -> fetch scale {fscale:.4f}")
    print(f"copy WRITE_SIZE reads {sum(big_w) / len(big_w):.0f} KiB for {copy_kib} KiB written -> write scale "
          f"{wscale:.4f}")
    print()
    fd = _first_serial_frame(fv)
    wd = _first_serial_frame(wv)
    print("per pass (KiB as counted; the camera pass's two launches summed), the first serial frame of the run:")
    print(f"{'pass':10s} {'FETCH_KiB':>12s} {'WRITE_KiB':>12s} {'scaled_MB':>12s}")
    tf = tw = 0.0
    for name, df, dw in zip(PASS_NAMES, fd, wd):
        f_ = sum(fv[x]["FETCH_SIZE"] for x in df)
        w_ = sum(wv[x]["WRITE_SIZE"] for x in dw)
        tf += f_ * 1024 * fscale
        tw += w_ * 1024 * wscale
        print(f"{name:10s} {f_:12.0f} {w_:12.0f} {(f_ * fscale + w_ * wscale) * 1024 / 1e6:12.1f}")
    n = max(len(fd), 1)
    print()
    print(f"mean per pass: fetch {tf / n:.4g} B, write {tw / n:.4g} B")
    print(f"traffic_bytes_per_launch = {(tf + tw) / n:.6g}   (per traced pass)")


def trav_valu(a):
    vals, _ = _counter_rows(a.trav_valu, "k_traverse64")
    d = _first_serial_frame(vals)
    keys = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM", "SQ_WAVE_CYCLES",
            "SQ_BUSY_CYCLES", "SQ_ACTIVE_INST_VALU"]
    print("Counters summed over the chip per pass (the first serial frame of the run, passes in order; the camera "
          "pass's two launches summed).")
    print(f"{'pass':10s} {'waves':>10s} {'VALU_insts':>12s} {'SALU_insts':>12s} {'LDS_insts':>12s} "
          f"{'VMEM_insts':>12s} {'VALU/wave':>10s}")
    tot = 0.0
    for name, ds in zip(PASS_NAMES, d):
        v = {}
        for di in ds:
            for k_, x in vals[di].items():
                v[k_] = v.get(k_, 0.0) + x
        w = max(v.get("SQ_WAVES", 0.0), 1.0)
        tot += v.get("SQ_INSTS_VALU", 0.0)
        print(f"{name:10s} {v.get('SQ_WAVES', 0):10.0f} {v.get('SQ_INSTS_VALU', 0):12.4g} "
              f"{v.get('SQ_INSTS_SALU', 0):12.4g} {v.get('SQ_INSTS_LDS', 0):12.4g} {v.get('SQ_INSTS_VMEM', 0):12.4g} "
              f"{v.get('SQ_INSTS_VALU', 0) / w:10.0f}")
    print()
    print(f"VALU instructions per frame: {tot:.4g}")


def hsc_summary(a):
    """--hsc-summary: the hash / sort / compress kernels of a --hsc-profile
    kernel trace, their summed durations, the algorithmic bytes of the
    renders and the fraction of peak HBM; with counter files, FETCH_SIZE +
    WRITE_SIZE of the same kernels against the algorithmic bytes."""
    import csv
    from rsh_amd import config as cf
    stats_csv, sizes_json = a.hsc_summary
    with open(sizes_json) as f:
        sizes = json.loads(f.read().strip().splitlines()[-1])
    cfg = cf.FrameConfig(**{k: v for k, v in sizes["cfg"].items() if k in cf.FrameConfig.__dataclass_fields__})
    rows = {}
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            base = _kernel_base(r["Name"])
            if base in HSC_KERNELS:
                m = re.search(r"\b(k_[A-Za-z0-9_]+(<[^>]*>)?)", r["Name"])
                d = rows.setdefault(m.group(1) if m else base, [0, 0.0])
                d[0] += int(r["Calls"])
                d[1] += float(r["TotalDurationNs"]) / 1e3
    total_us = sum(v[1] for v in rows.values())
    R = sizes["renders"]
    nbytes = R * hsc_total_bytes(sizes["passes"], cfg)
    gbs = nbytes / (total_us * 1e-6) / 1e9
    lines = [f"kernel{'':40s} calls    total_us   avg_us"]
    for k, (c, t) in sorted(rows.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{k:46s} {c:6d} {t:11.1f} {t / max(c, 1):8.2f}")
    lines.append("")
    lines.append(f"renders in the trace: {R}; passes per render: {len(sizes['passes'])}; "
                 f"sort digit passes: {sizes['digit_passes']}")
    lines.append("per pass (kind 0 camera / 1 shadow / 2 reflection): slots, live rays, chunks, algorithmic bytes")
    P = sort_digit_passes(cfg)
    for i, p in enumerate(sizes["passes"]):
        lines.append(f"  {i}: kind {p['kind']}  S {p['slots']:>10d}  N {p['rays']:>10d}  C {p['chunks']:>10d}  "
                     f"{hsc_bytes(p, i == 0, cfg, P):>13d} B")
    lines.append(f"algorithmic bytes: {nbytes} ({R} renders)")
    lines.append(f"summed kernel time: {total_us:.1f} us  ({total_us / R:.1f} us per render)")
    lines.append(f"achieved: {gbs:.1f} GB/s of {PEAK_HBM_GBS:.0f}")
    lines.append(f"hsc_frac_rocprof = {gbs / PEAK_HBM_GBS:.4f}")

    def counter(path):
        tot = 0.0
        with open(path) as f:
            for r in csv.DictReader(f):
                if _kernel_base(r.get("Kernel_Name", "")) in HSC_KERNELS:
                    tot += float(r["Counter_Value"])
        return tot * 1024.0
    if a.fetch_csv and a.write_csv:
        rc = a.counter_renders or R
        fb = counter(a.fetch_csv) * a.fetch_scale / rc
        wb = counter(a.write_csv) / rc
        per = nbytes / R
        lines.append(f"per render ({rc} in the counter runs): FETCH_SIZE x {a.fetch_scale:.4f}: {fb:.4e} B; "
                     f"WRITE_SIZE: {wb:.4e} B; sum {fb + wb:.4e} B = {(fb + wb) / per:.3f} x the algorithmic "
                     f"{per:.4e} B")
    print("\n".join(lines))


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


def pass_slot_maps(views, L):
    """Per view, the batch slot of each of its slots, by pass kind: camera
    slots are pixels; reflection slot x + y w L of a view is pixel off + x +
    y w of the batch (a P x 1 image); shadow slot x + w (l + L y) is batch
    slot (off + x + y w) + P l."""
    P = sum(w * h for _, w, h in views)
    import numpy as np
    maps, off = [], 0
    for _, w, h in views:
        q = np.arange(w * h)
        x, y = q % w, q // w
        m = {0: (q, off + q), 2: (x + y * w * L, off + q)}
        sl = [x + w * (l + L * y) for l in range(L)]
        bl = [off + q + P * l for l in range(L)]
        m[1] = (np.concatenate(sl), np.concatenate(bl))
        maps.append(m)
        off += w * h
    return maps


def hit_agreement(eng, views, cfg, rsh_amd):
    """Per-ray results of the batched form against the plain per-view frames,
    pass by pass: closest-hit triangle ids (and, where they differ, the
    relative difference of the hit distances) and shadow flags."""
    import dataclasses
    import numpy as np
    L = eng._scene.light_total
    eng.render_views(views, cfg, flags=eng.KEEP_PASSES)
    nb = eng.stats.pass_count
    kinds = [int(eng.stats.pass_kind[i]) for i in range(nb)]
    batch = [eng.pass_hits(i) for i in range(nb)]
    maps = pass_slot_maps(views, L)
    same = [0] * nb
    total = [0] * nb
    worst = 0.0
    hit_vs_miss = 0
    for v, (cam, w, h) in enumerate(views):
        eng.render(cam, dataclasses.replace(cfg, width=w, height=h), flags=eng.KEEP_PASSES)
        for i in range(min(nb, eng.stats.pass_count)):
            plain = eng.pass_hits(i)
            ps, bs = maps[v][kinds[i]]
            a, b = plain[ps], batch[i][bs]
            if kinds[i] == 1:
                same[i] += int((a == b).sum())
            else:
                ta, tb = a & np.uint64(0xFFFFFFFF), b & np.uint64(0xFFFFFFFF)
                eq = ta == tb
                same[i] += int(eq.sum())
                d = ~eq
                if d.any():
                    miss = np.uint64(0xFFFFFFFFFFFFFFFF)
                    both = d & (a != miss) & (b != miss)
                    hit_vs_miss += int((d & ~both).sum())
                    fa = (a[both] >> np.uint64(32)).astype(np.uint32).view(np.float32).astype(np.float64)
                    fb = (b[both] >> np.uint64(32)).astype(np.uint32).view(np.float32).astype(np.float64)
                    if both.any():
                        worst = max(worst, float(np.max(np.abs(fa - fb) / np.maximum(np.abs(fa), 1e-30))))
            total[i] += len(ps)
    rays = sum(total)
    agree = sum(same)
    return {"slots_compared": rays, "identical": agree, "fraction": agree / rays,
            "per_pass_differing": [t - s for s, t in zip(same, total)],
            "max_rel_hit_distance_among_differing_hits": worst,
            "differing_hit_vs_miss": hit_vs_miss}


def batching_leg(rsh_amd, eng, scene2, tune=True):
    """Plain per-view frames against the batched form on the 64-view batch:
    throughput (eager and graph-replayed), per-ray hit agreement, the hash /
    sort / compress bandwidth, and the autotuner's choice for the atrium."""
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
            imgs.append(eng.read_pixels_pinned().copy())
        plain = time.perf_counter() - t0
        eng.render_views(views, cfg)
        t0 = time.perf_counter()
        st = eng.render_views(views, cfg)
        eng.read_pixels_pinned()
        eager = time.perf_counter() - t0
        parts = eng.split_views(eng.pixels(), views)
        same = sum(int((a == b).sum()) for a, b in zip(parts, imgs)) / out["pixels"]
        graph_ms, brays = autotune.time_batch(eng, views, cfg, reps=3)
        st = eng.render_views(views, cfg, flags=eng.TIMING | eng.SERIAL)
        stages = {n: st.stage_ms[i] for i, n in enumerate(rsh_amd.STAGES)}
        bp = [{"slots": st.pass_slots[i], "rays": st.pass_rays[i], "chunks": st.pass_chunks[i]}
              for i in range(st.pass_count)]
        hsc = hsc_figure(bp, stages, cfg, name)
        out[name] = {"plain_mrays": rays / plain / 1e6, "plain_ms": plain * 1e3,
                     "batched_mrays": st.rays_total / eager / 1e6, "batched_ms": eager * 1e3,
                     "graph_mrays": brays / graph_ms / 1e3, "graph_ms": graph_ms,
                     "speedup": (brays / graph_ms / 1e3) / (rays / plain / 1e6),
                     "rays": int(brays), "identical_pixels_vs_plain": same,
                     "hits_vs_plain": hit_agreement(eng, views, cfg, rsh_amd),
                     "hash_sort_compress": hsc}
        progress(f"batching: {name} done")
        if tune and name == "atrium":
            t0 = time.perf_counter()
            best, trials = autotune.tune(eng, scene, views, cfg, reps=3)
            out[name]["autotune"] = {
                "seconds": time.perf_counter() - t0,
                "chosen": {"cell_bits": best.cell_bits, "dir_bits": best.dir_bits, "leaf": best.leaf,
                           "branch": best.branch, "traversal_chunk": best.traversal_chunk,
                           "traversal_prefilter": best.traversal_prefilter},
                "chosen_graph_mrays": max(t.mrays for t in trials),
                "start_graph_mrays": trials[0].mrays,
                "trials": [[t.cell_bits, t.dir_bits, t.leaf, t.branch, t.traversal_chunk, t.traversal_prefilter,
                            round(t.ms, 2), round(t.mrays, 1), round(t.same_pixels, 6)] for t in trials],
                "trial_fields": ["cell_bits", "dir_bits", "leaf", "branch", "traversal_chunk",
                                 "traversal_prefilter", "graph_ms", "mrays", "same_pixels_as_start"]}
            progress("batching: autotune done")
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
    if a.hsc_summary:
        hsc_summary(a)
        return
    if a.trav_summary:
        trav_summary(a)
        return
    if a.trav_traffic:
        trav_traffic(a)
        return
    if a.trav_valu:
        trav_valu(a)
        return
    if a.hsc_profile:
        hsc_profile_run(a)
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
        eng.read_pixels_async()
    eng.wait_pixels()
    rsh_amd.lib().rsh_device_synchronize()
    times, rays = [], []
    stage_sum = [0.0] * len(rsh_amd.STAGES)
    trav_ms, trav_launches = 0.0, 0
    t_all = time.perf_counter()
    for _ in range(a.steps):
        t0 = time.perf_counter()
        # the frame as served: shadow passes concurrent with the reflection
        # chain, then the image read back into page-locked host memory, the
        # copy running beside the next frame's first passes (the last one is
        # waited for inside the timed region)
        st = eng.render(cam, cfg)
        eng.read_pixels_async()
        times.append(time.perf_counter() - t0)
        rays.append(st.rays_total)
    async_view = eng.wait_pixels()
    rsh_amd.lib().rsh_device_synchronize()
    total = time.perf_counter() - t_all
    value = sum(rays) / total / 1e6
    async_img = async_view.copy()
    eager_img = eng.pixels()
    async_same = bool((async_img == eager_img).all())     # the pipelined copy is the frame's image
    # The same frame replayed from a captured HIP graph (valid while the
    # camera, configuration and scene stay the same), image read back too.
    eng.render(cam, cfg, flags=eng.GRAPH)
    eng.render(cam, cfg, flags=eng.GRAPH)
    graph_same = bool((eng.read_pixels_pinned() == eager_img).all())
    g_times = []
    g0 = time.perf_counter()
    for _ in range(a.steps):
        t0 = time.perf_counter()
        st = eng.render(cam, cfg, flags=eng.GRAPH)
        eng.read_pixels_pinned()
        g_times.append(time.perf_counter() - t0)
    g_total = time.perf_counter() - g0
    graph = {"mrays": a.steps * statistics.median(rays) / g_total / 1e6, "ms_per_frame": 1e3 * g_total / a.steps,
             "frame_ms_median": 1e3 * statistics.median(g_times), "identical_image": graph_same,
             "note": "HIP graph of the whole frame replayed; valid while camera, configuration and scene repeat"}
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
    passes_per_frame = sum(1 for p in passes if p["rays"])
    flops_frame = FLOP_RAY_TEST * n_tests + FLOP_NODE_TEST * (n_top + cfg.branch * n_surv)
    avg_launch_ms = trav_ms / trav_launches if trav_launches else 0.0
    flops_launch = flops_frame / passes_per_frame if passes_per_frame else 0.0
    achieved = flops_launch / (avg_launch_ms * 1e-3) / 1e12 if avg_launch_ms > 0 else 0.0
    roof = {"bound": "valu",
            "kernel": "k_traverse64<8, MODE> (fused traversal + intersection; one launch per pass, two for the "
                      "front-to-back camera pass)",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_per_launch(),
            "avg_launch_ms": avg_launch_ms, "per": "traced pass (the camera pass's two launches together)",
            "passes_per_frame": passes_per_frame,
            "flops_per_launch": flops_launch}
    hsc = hsc_figure(passes, stages, cfg, "frame")
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
        batching = batching_leg(rsh_amd, eng, scene, tune=not a.no_tune)
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
                   "hierarchy": "leaf 8, branch 8, depth 2",
                   "cone_union": "guarded (normalised, clamped dots, parallel and antipodal rays handled); "
                                 "the reference's CreateHierarchyCone0 union is cone_mode 0"},
        "roofline": roof,
        "hash_sort_compress": hsc,
        "stage_ms": stages,
        "frame_ms_median": 1e3 * statistics.median(times),
        "frame_ms": [round(1e3 * t, 3) for t in times],
        "timed_region": "render + read-back of the 1920x1080 image into page-locked host memory (each frame's copy "
                        "queued behind it on a copy stream and run beside the next frame's first passes; the last "
                        "copy waited for)",
        "readback_identical": async_same,
        "graph_replay": graph,
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
