"""ctypes bindings of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

u32, i32, f32, u64 = C.c_uint32, C.c_int32, C.c_float, C.c_uint64
PF, PU, PI, PU64 = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)


class OScene(C.Structure):
    _fields_ = [("triangle_total", u32), ("tri_pos", PF), ("tri_nrm", PF), ("tri_mat", PI),
                ("material_total", u32), ("mat_diffuse", PF), ("mat_specular", PF),
                ("object_total", u32), ("object_spheres", PF),
                ("light_total", u32), ("light_pos", PF), ("light_col", PF), ("light_int", PF)]


class OCamera(C.Structure):
    _fields_ = [("eye", f32 * 3), ("direction", f32 * 3), ("right", f32 * 3), ("up", f32 * 3)]


class OView(C.Structure):
    _fields_ = [("camera", OCamera), ("width", u32), ("height", u32)]


class OConfig(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("algorithm", u32), ("key_mode", u32),
                ("hit_mode", u32), ("cone_mode", u32), ("leaf", u32), ("branch", u32),
                ("depth", u32), ("triangle_batch", u32), ("bounces", u32), ("reflect_gate", u32),
                ("bounce_shadows", u32), ("primary_rays", u32), ("cell_bits", u32), ("dir_bits", u32),
                ("key_layout", u32), ("scene_min", f32 * 3), ("cell_scale", f32 * 3),
                ("segment_shift", u32), ("initial_radius", f32), ("initial_spread", f32),
                ("row_begin", u32), ("row_end", u32), ("soft_shadows", u32)]


class OFrameOut(C.Structure):
    _fields_ = [("pixels", PU), ("gbuffer_pos", PF), ("gbuffer_nrm", PF), ("gbuffer_dif", PF),
                ("gbuffer_spec", PF), ("pass_hits", PU64 * 16), ("pass_count", u32),
                ("pass_kind", u32 * 16), ("pass_slots", u32 * 16), ("pass_rays", u32 * 16),
                ("pass_chunks", u32 * 16), ("pass_pairs_top", u64 * 16), ("pass_pairs_leaf", u64 * 16),
                ("pass_ray_tests", u64 * 16), ("stage_ms", C.c_double * 8),
                ("pass_keys", PU * 16), ("pass_flags", PU * 16), ("pass_order", PU * 16),
                ("pass_bases", PU * 16), ("pass_ckeys", PU * 16), ("pass_nodes", PF * 16)]


REF_SRC = "/root/reference/projects/CUDA-Engine/source"
REF_MINIBALL = os.path.join(HERE, "_ref", "libref_miniball.so")


def build(quiet: bool = True) -> str:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off), and the
    reference's Miniball behind a C entry point (oracle/_ref) where the
    reference sources are present."""
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if os.path.exists(os.path.join(REF_SRC, "Miniball.hpp")):
        r = subprocess.run(["make", "-s", "-C", HERE, "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference Miniball build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def ref_miniball(points):
    """(centre, squared radius) of the reference's Miniball over float32
    points [n, 3]; None when oracle/_ref was not built."""
    if not os.path.exists(REF_MINIBALL):
        return None
    L = C.CDLL(REF_MINIBALL)
    L.ref_miniball.restype = None
    L.ref_miniball.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    c = np.zeros(3, np.float32)
    r2 = np.zeros(1, np.float32)
    L.ref_miniball(p.ctypes.data, p.shape[0], c.ctypes.data, r2.ctypes.data)
    return c, float(r2[0])


_lib = None
STAGES = ("create", "trim", "compress", "sort", "decompress", "hierarchy", "traversal", "shading")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name in ("oracle_acos", "oracle_cos", "oracle_sin", "oracle_tan"):
            getattr(L, name).restype = f32
            getattr(L, name).argtypes = [f32]
        L.oracle_pow.restype = f32
        L.oracle_pow.argtypes = [f32, f32]
        L.oracle_f2u_sat.restype = u32
        L.oracle_f2u_sat.argtypes = [f32]
        L.oracle_shadow_key_ref.restype = u32
        L.oracle_shadow_key_ref.argtypes = [PF, PF, u32]
        L.oracle_reflection_key_ref.restype = u32
        L.oracle_reflection_key_ref.argtypes = [PF, PF]
        L.oracle_refraction_key_ref.restype = u32
        L.oracle_refraction_key_ref.argtypes = [PF, PF]
        L.oracle_refract.argtypes = [PF, PF, f32, PF]
        L.oracle_key_oct.restype = u32
        L.oracle_key_oct.argtypes = [C.POINTER(OConfig), PF, PF, u32]
        L.oracle_ray_triangle.restype = f32
        L.oracle_ray_triangle.argtypes = [PF] * 5
        L.oracle_soft_grid.argtypes = [PF, PF, PF]
        L.oracle_soft_factor.restype = f32
        L.oracle_soft_factor.argtypes = [u32]
        L.oracle_sphere_node.restype = C.c_int
        L.oracle_sphere_node.argtypes = [PF] * 3
        L.oracle_triangle_sphere.argtypes = [PF] * 4
        L.oracle_cone0.argtypes = [PF, PF, u32, PF]
        L.oracle_coneN.argtypes = [PF, PF, PF]
        L.oracle_sphere_union.argtypes = [PF, PF, PF]
        L.oracle_rgb_to_int.restype = u32
        L.oracle_rgb_to_int.argtypes = [f32, f32, f32]
        L.oracle_primary_key.restype = u32
        L.oracle_primary_key.argtypes = [u32] * 5
        L.oracle_light_term.restype = None
        L.oracle_light_term.argtypes = [PF] * 8 + [f32, C.c_int, PF]
        L.oracle_render.restype = C.c_int
        L.oracle_render.argtypes = [C.POINTER(OScene), C.POINTER(OCamera), C.POINTER(OConfig), PF,
                                    C.POINTER(OFrameOut), C.c_int]
        L.oracle_trace_rays.restype = C.c_int
        L.oracle_trace_rays.argtypes = [C.POINTER(OScene), PF, PF, PU, u32, u32, C.POINTER(OConfig), PU64, C.c_int]
        L.oracle_render_views.restype = C.c_int
        L.oracle_render_views.argtypes = [C.POINTER(OScene), C.POINTER(OView), u32, C.POINTER(OConfig),
                                          C.POINTER(OFrameOut), C.c_int]
        for n in ("oracle_leaf_pairs_literal", "oracle_leaf_pairs_closed"):
            getattr(L, n).restype = u64
            getattr(L, n).argtypes = [PF, PU, u32, u32, C.POINTER(OScene), u32, u32, PU64, u64]
        _lib = L
    return _lib


def fp(a):
    return a.ctypes.data_as(PF)


def up(a):
    return a.ctypes.data_as(PU)


def ip(a):
    return a.ctypes.data_as(PI)


def u64p(a):
    return a.ctypes.data_as(PU64)


def make_scene(s):
    """OScene over the arrays of a Scene (keeps references alive in .keep)."""
    arrs = [np.ascontiguousarray(x) for x in (s.tri_pos, s.tri_nrm, s.tri_mat, s.mat_diffuse,
                                               s.mat_specular, s.object_spheres, s.light_pos,
                                               s.light_col, s.light_int)]
    o = OScene(s.tri_pos.shape[0] // 3, fp(arrs[0]), fp(arrs[1]), ip(arrs[2]), arrs[3].shape[0],
               fp(arrs[3]), fp(arrs[4]), arrs[5].shape[0], fp(arrs[5]), arrs[6].shape[0],
               fp(arrs[6]), fp(arrs[7]), fp(arrs[8]))
    o.keep = arrs
    return o


def make_camera(cam):
    c = OCamera()
    c.eye[:] = [float(v) for v in cam.eye]
    c.direction[:] = [float(v) for v in cam.direction]
    c.right[:] = [float(v) for v in cam.right]
    c.up[:] = [float(v) for v in cam.up]
    return c


def make_config(cfg):
    o = OConfig()
    for name, _ in OConfig._fields_:
        if name in ("row_begin", "row_end"):
            continue
        v = getattr(cfg, name)
        if name in ("scene_min", "cell_scale"):
            getattr(o, name)[:] = [float(x) for x in v]
        else:
            setattr(o, name, v)
    return o


def call(name, *args):
    """Call an oracle_<Wrapper> restatement: numpy arrays become pointers,
    Python ints become uint32, Python floats float32, None a null pointer."""
    fn = getattr(lib(), name)
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            assert a.flags.c_contiguous
            cargs.append(C.c_void_p(a.ctypes.data))
        elif a is None:
            cargs.append(C.c_void_p(0))
        elif isinstance(a, bool):
            cargs.append(C.c_int(int(a)))
        elif isinstance(a, int):
            cargs.append(C.c_uint32(a))
        elif isinstance(a, float):
            cargs.append(C.c_float(a))
        else:
            cargs.append(a)
    fn.restype = None
    return fn(*cargs)


class Frame:
    """Outputs of one oracle frame."""

    def __init__(self, pixels, gbuffer, passes, stage_ms=None):
        self.pixels = pixels
        self.gbuffer = gbuffer
        self.passes = passes
        self.stage_ms = stage_ms


def _order_buffers(out, S):
    """Per-pass keys, flags, sorted order, packets and nodes (capacity S)."""
    bufs = []
    for i in range(16):
        b = {k: np.zeros(S, np.uint32) for k in ("keys", "flags", "order", "bases", "ckeys")}
        b["nodes"] = np.zeros((2 * S, 4), np.float32)
        out.pass_keys[i], out.pass_flags[i], out.pass_order[i] = up(b["keys"]), up(b["flags"]), up(b["order"])
        out.pass_bases[i], out.pass_ckeys[i], out.pass_nodes[i] = up(b["bases"]), up(b["ckeys"]), fp(b["nodes"])
        bufs.append(b)
    return bufs


def _attach_order(passes, bufs):
    for p, b in zip(passes, bufs):
        p["keys"], p["flags"] = b["keys"][:p["slots"]], b["flags"][:p["slots"]]
        p["order"] = b["order"][:p["rays"]]
        p["bases"], p["ckeys"] = b["bases"][:p["chunks"]], b["ckeys"][:p["chunks"]]
        p["nodes"] = b["nodes"]


def render(scene, camera, cfg, gbuffer=None, threads=0, keep_hits=True, keep_gbuffer=True, row_range=None,
           keep_order=False) -> Frame:
    """One oracle frame.  keep_order: every pass's slot keys and flags, sorted
    order, packets (chunk bases and keys) and hierarchy nodes in
    passes[i]["keys" / "flags" / "order" / "bases" / "ckeys" / "nodes"]."""
    L = lib()
    W, H = cfg.width, cfg.height
    S = W * H * max(1, scene.light_pos.shape[0]) * (16 if getattr(cfg, "soft_shadows", 0) else 1)
    pixels = np.zeros(W * H, np.uint32)
    out = OFrameOut()
    out.pixels = up(pixels)
    obufs = _order_buffers(out, S) if keep_order else None
    gb = None
    if keep_gbuffer:
        gb = {k: np.zeros((W * H, 4), np.float32) for k in ("pos", "nrm", "dif", "spec")}
        out.gbuffer_pos, out.gbuffer_nrm = fp(gb["pos"]), fp(gb["nrm"])
        out.gbuffer_dif, out.gbuffer_spec = fp(gb["dif"]), fp(gb["spec"])
    hits = []
    if keep_hits:
        for i in range(16):
            h = np.zeros(S, np.uint64)
            hits.append(h)
            out.pass_hits[i] = u64p(h)
    os_ = make_scene(scene)
    oc = make_camera(camera)
    ocfg = make_config(cfg)
    if row_range is not None:
        ocfg.row_begin, ocfg.row_end = int(row_range[0]), int(row_range[1])
    gin = None
    if gbuffer is not None:
        gin = np.ascontiguousarray(np.concatenate([gbuffer["dif"], gbuffer["spec"], gbuffer["pos"], gbuffer["nrm"]]),
                                   dtype=np.float32)
    rc = L.oracle_render(C.byref(os_), C.byref(oc), C.byref(ocfg), fp(gin) if gin is not None else None,
                         C.byref(out), threads)
    if rc != 0:
        raise RuntimeError("oracle_render failed: %d" % rc)
    passes = []
    for i in range(out.pass_count):
        passes.append(dict(kind=out.pass_kind[i], slots=out.pass_slots[i], rays=out.pass_rays[i],
                           chunks=out.pass_chunks[i], pairs_top=out.pass_pairs_top[i],
                           pairs_leaf=out.pass_pairs_leaf[i], ray_tests=out.pass_ray_tests[i],
                           hits=hits[i][:out.pass_slots[i]].copy() if keep_hits else None))
    if obufs:
        _attach_order(passes, obufs)
    return Frame(pixels.reshape(H, W), gb, passes, dict(zip(STAGES, list(out.stage_ms))))


def render_views(scene, views, cfg, threads=0, keep_hits=True, keep_gbuffer=True) -> Frame:
    """A batch of views, each (camera, width, height), rendered as one image of
    sum(w*h) pixels (returned as a 1 x P frame; a single view keeps H x W)."""
    L = lib()
    n = len(views)
    P = sum(int(w) * int(h) for _, w, h in views)
    W, H = (int(views[0][1]), int(views[0][2])) if n == 1 else (P, 1)
    S = P * max(1, scene.light_pos.shape[0]) * (16 if getattr(cfg, "soft_shadows", 0) else 1)
    pixels = np.zeros(P, np.uint32)
    out = OFrameOut()
    out.pixels = up(pixels)
    gb = None
    if keep_gbuffer:
        gb = {k: np.zeros((P, 4), np.float32) for k in ("pos", "nrm", "dif", "spec")}
        out.gbuffer_pos, out.gbuffer_nrm = fp(gb["pos"]), fp(gb["nrm"])
        out.gbuffer_dif, out.gbuffer_spec = fp(gb["dif"]), fp(gb["spec"])
    hits = []
    if keep_hits:
        for i in range(16):
            h = np.zeros(S, np.uint64)
            hits.append(h)
            out.pass_hits[i] = u64p(h)
    ov = (OView * n)()
    for i, (cam, w, h) in enumerate(views):
        ov[i].camera = make_camera(cam)
        ov[i].width, ov[i].height = int(w), int(h)
    os_ = make_scene(scene)
    ocfg = make_config(cfg)
    rc = L.oracle_render_views(C.byref(os_), ov, n, C.byref(ocfg), C.byref(out), threads)
    if rc != 0:
        raise RuntimeError("oracle_render_views failed: %d" % rc)
    passes = []
    for i in range(out.pass_count):
        passes.append(dict(kind=out.pass_kind[i], slots=out.pass_slots[i], rays=out.pass_rays[i],
                           chunks=out.pass_chunks[i], pairs_top=out.pass_pairs_top[i],
                           pairs_leaf=out.pass_pairs_leaf[i], ray_tests=out.pass_ray_tests[i],
                           hits=hits[i][:out.pass_slots[i]].copy() if keep_hits else None))
    return Frame(pixels.reshape(H, W), gb, passes)


def trace_rays(scene, rays, kind, cfg, tmax=None, segments=None, threads=0):
    """Ray query: rays float32 [n, 6]; returns uint64 [n] (closest-hit records,
    or 0xFFFFFFFF occluded / 0 clear for kind 1)."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
    sg = None if segments is None else np.ascontiguousarray(segments, np.uint32)
    out = np.zeros(n, np.uint64)
    os_ = make_scene(scene)
    rc = lib().oracle_trace_rays(C.byref(os_), fp(rays), fp(tm) if tm is not None else None,
                                 up(sg) if sg is not None else None, n, kind, C.byref(make_config(cfg)),
                                 u64p(out), threads)
    if rc != 0:
        raise RuntimeError("oracle_trace_rays failed: %d" % rc)
    return out
