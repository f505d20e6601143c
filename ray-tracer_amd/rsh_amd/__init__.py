"""Python access to the gfx950 ray-space-hierarchy library (lib/librsh.so).

Only ctypes plumbing lives here; every stage runs in the HIP library.  If the
library is missing, :func:`lib` raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import config as config  # noqa: F401
from . import scenes as scenes  # noqa: F401

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(ROOT, "lib", "librsh.so")

u32, i32, f32, u64 = C.c_uint32, C.c_int32, C.c_float, C.c_uint64
PF, PU, PI, PU64, PV = (C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                        C.POINTER(C.c_uint64), C.c_void_p)

MAX_PASSES = 16


class SceneDesc(C.Structure):
    _fields_ = [("triangle_total", u32), ("tri_pos", PF), ("tri_nrm", PF), ("tri_mat", PI),
                ("material_total", u32), ("mat_diffuse", PF), ("mat_specular", PF),
                ("object_total", u32), ("object_spheres", PF),
                ("light_total", u32), ("light_pos", PF), ("light_col", PF), ("light_int", PF)]


class CameraDesc(C.Structure):
    _fields_ = [("eye", f32 * 3), ("direction", f32 * 3), ("right", f32 * 3), ("up", f32 * 3)]


class ViewDesc(C.Structure):
    _fields_ = [("camera", CameraDesc), ("width", u32), ("height", u32)]


class FrameConfigDesc(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("algorithm", u32), ("key_mode", u32),
                ("hit_mode", u32), ("cone_mode", u32), ("leaf", u32), ("branch", u32),
                ("depth", u32), ("triangle_batch", u32), ("bounces", u32), ("reflect_gate", u32),
                ("bounce_shadows", u32), ("primary_rays", u32), ("cell_bits", u32), ("dir_bits", u32),
                ("key_layout", u32), ("scene_min", f32 * 3), ("cell_scale", f32 * 3),
                ("segment_shift", u32), ("initial_radius", f32), ("initial_spread", f32),
                ("traversal_kernel", u32), ("traversal_splits", u32), ("traversal_chunk", u32),
                ("traversal_prefilter", u32), ("soft_shadows", u32)]


class FrameStats(C.Structure):
    _fields_ = [("pass_count", u32), ("pass_kind", u32 * MAX_PASSES), ("pass_slots", u32 * MAX_PASSES),
                ("pass_rays", u32 * MAX_PASSES), ("pass_chunks", u32 * MAX_PASSES),
                ("pass_pairs_top", u64 * MAX_PASSES), ("pass_pairs_leaf", u64 * MAX_PASSES),
                ("pass_ray_tests", u64 * MAX_PASSES), ("stage_ms", f32 * 8), ("rays_total", u32),
                ("pass_top_survivors", u64 * MAX_PASSES), ("pass_traverse_ms", f32 * MAX_PASSES),
                ("pass_splits", u32 * MAX_PASSES), ("pass_items", u32 * MAX_PASSES),
                ("pass_chunk", u32 * MAX_PASSES), ("pass_tri_tests", u64 * MAX_PASSES),
                ("pass_cluster_tests", u64 * MAX_PASSES), ("pass_leaf_steps", u64 * MAX_PASSES),
                ("pass_ray_steps", u64 * MAX_PASSES), ("graph_replayed", u32),
                ("pass_mt_tests", u64 * MAX_PASSES), ("pass_mode", u32 * MAX_PASSES)]


class Float3(C.Structure):
    _fields_ = [("x", f32), ("y", f32), ("z", f32)]


STAGES = ("create", "trim", "compress", "sort", "decompress", "hierarchy", "traversal", "shading")

# Reference wrapper signatures (include/rsh_engine.h); P = device pointer.
_P = PV
_U = C.c_uint
_F3 = Float3
COMPAT_SIGNATURES = {
    "TriangleUpdateWrapper": [_P, _P, _U, _P, _P],
    "BoundingSphereUpdateWrapper": [_P, _P, _U, _P],
    "MemoryPreparationWrapper": [_U, _U, _U, _P, _P, _P, _P, _P, _P],
    "ShadowRayCreationWrapper": [_U, _U, _U, _P, _P, _P, _P],
    "ReflectionRayCreationWrapper": [_U, _U, _F3, _P, _P, _P, _P],
    "RefractionRayCreationWrapper": [_U, _U, _F3, _P, _P, _P, _P],
    "RayTrimmingWrapper": [_P, _P, _U, _U, _U, _P, _P, _P, _P, PU],
    "RayCompressionWrapper": [_P, _P, _U, _P, _P, _P, _P, _P, _P, PU],
    "RaySortingWrapper": [_P, _P, _U, _P, _P],
    "rsh_sort_pairs": [_P, _P, _U, _U, _P, _P],
    "RayDecompressionWrapper": [_P, _P, _P, _P, _P, _P, _P, _P, _U, _P, _P],
    "HierarchyCreationWrapper": [_P, _P, _P, _U, f32, f32, _P],
    "HierarchyTraversalWarmUpWrapper": [_P, _P, _P, _U, _U, _U, _U, _P, _P, _P, _P, PU, PU],
    "HierarchyTraversalWrapper": [_P, _P, _U, _U, _U, _P, _P, _P, _P, PU, PU],
    "ShadowRayPreparationWrapper": [_U, _U, _U, _P],
    "ShadowRayIntersectionWrapper": [_P, _P, _P, _P, _P, _U, _U, _U, _U, _U, _F3, _P],
    "ShadowRayColoringWrapper": [_U, _U, _U, _F3, _P, _P],
    "ReflectionRayPreparationWrapper": [_U, _U, _P],
    "ReflectionRayIntersectionWrapper": [_P, _P, _P, _P, _P, _U, _U, _U, _U, _U, _U, _F3, _P],
    "ReflectionRayColoringWrapper": [_P, _P, _U, _U, _U, _F3, C.c_bool, _P, _P, _P, _P, _P, _P],
}
# The per-context stage API (include/rsh_ctx.h): (ctx, stream) and then the
# wrapper's parameters, with device-count words (P) after the counts.
CTX_STAGES = {
    "rsh_triangle_update": "TriangleUpdateWrapper", "rsh_bounding_sphere_update": "BoundingSphereUpdateWrapper",
    "rsh_shadow_ray_creation": "ShadowRayCreationWrapper", "rsh_reflection_ray_creation": "ReflectionRayCreationWrapper",
    "rsh_refraction_ray_creation": "RefractionRayCreationWrapper",
    "rsh_hierarchy_traversal_warmup": "HierarchyTraversalWarmUpWrapper",
    "rsh_hierarchy_traversal": "HierarchyTraversalWrapper", "rsh_shadow_ray_preparation": "ShadowRayPreparationWrapper",
    "rsh_shadow_ray_intersection": "ShadowRayIntersectionWrapper", "rsh_shadow_ray_coloring": "ShadowRayColoringWrapper",
    "rsh_reflection_ray_preparation": "ReflectionRayPreparationWrapper",
    "rsh_reflection_ray_intersection": "ReflectionRayIntersectionWrapper",
    "rsh_reflection_ray_coloring": "ReflectionRayColoringWrapper",
}
CTX_SIGNATURES = {name: COMPAT_SIGNATURES[w] for name, w in CTX_STAGES.items()}
CTX_SIGNATURES.update({
    "rsh_memory_preparation": [_U, _U, _U],
    "rsh_ray_trimming": [_P, _P, _U, _U, _U, _P, _P, _P, _P, PU, _P],
    "rsh_ray_compression": [_P, _P, _U, _P, _P, _P, _P, _P, _P, _P, PU, _P],
    "rsh_ray_sorting": [_P, _P, _U, _P, _P, _P],
    "rsh_ctx_sort_pairs": [_P, _P, _U, _P, _U, _P, _P],
    "rsh_ray_decompression": [_P, _P, _P, _P, _P, _P, _P, _P, _U, _P, _P, _P],
    "rsh_hierarchy_creation": [_P, _P, _P, _U, _P, f32, f32, _P],
})
BINDERS = ["bindDiffuseTextureArray", "bindSpecularTextureArray", "bindFragmentPositionArray",
           "bindFragmentNormalArray", "bindTrianglePositions", "bindTriangleNormals",
           "bindTriangleTextureCoordinates", "bindTriangleObjectIDs", "bindTriangleMaterialIDs",
           "bindMaterialDiffuseProperties", "bindMaterialSpecularProperties", "bindBoundingSpheres",
           "bindLightPositions", "bindLightColors", "bindLightIntensities"]
NATIVE_SYMBOLS = ["rsh_engine_create", "rsh_engine_destroy", "rsh_last_error", "rsh_engine_set_stream",
                  "rsh_engine_load_scene", "rsh_engine_render", "rsh_engine_render_views", "rsh_engine_trace",
                  "rsh_engine_read_pixels", "rsh_engine_read_pixels_async", "rsh_engine_wait_pixels",
                  "rsh_engine_read_gbuffer", "rsh_engine_read_pass", "rsh_engine_pixels_device",
                  "rsh_device_alloc", "rsh_device_free", "rsh_memcpy_htod", "rsh_memcpy_dtoh",
                  "rsh_memset_d32", "rsh_device_synchronize", "rsh_device_count",
                  "rsh_compat_configure", "rsh_compat_set_stream", "rsh_compat_set_exit_on_error",
                  "rsh_compat_status", "rsh_compat_error", "rsh_compat_reset_status", "rsh_selfcheck_rcp",
                  "rsh_object_spheres", "rsh_engine_update_objects", "rsh_engine_set_wait_timeout",
                  "rsh_selfcheck_wait_deadline", "rsh_engine_read_pass_order"]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load lib/librsh.so (raises LibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} not built: run __graft_entry__.build() or rsh_amd/build.py")
    L = C.CDLL(LIB_PATH)
    L.rsh_engine_create.restype = C.c_int
    L.rsh_engine_create.argtypes = [C.POINTER(PV), C.c_int]
    L.rsh_engine_destroy.restype = None
    L.rsh_engine_destroy.argtypes = [PV]
    L.rsh_last_error.restype = C.c_char_p
    L.rsh_last_error.argtypes = [PV]
    L.rsh_engine_set_stream.argtypes = [PV, PV]
    L.rsh_engine_load_scene.restype = C.c_int
    L.rsh_engine_load_scene.argtypes = [PV, C.POINTER(SceneDesc)]
    L.rsh_engine_render.restype = C.c_int
    L.rsh_engine_render.argtypes = [PV, C.POINTER(CameraDesc), C.POINTER(FrameConfigDesc), u32,
                                    C.POINTER(FrameStats)]
    L.rsh_engine_render_views.restype = C.c_int
    L.rsh_engine_render_views.argtypes = [PV, C.POINTER(ViewDesc), u32, C.POINTER(FrameConfigDesc), u32,
                                          C.POINTER(FrameStats)]
    L.rsh_engine_trace.restype = C.c_int
    L.rsh_engine_trace.argtypes = [PV, PV, PV, PV, u32, u32, C.POINTER(FrameConfigDesc), PV, u32,
                                   C.POINTER(FrameStats)]
    L.rsh_engine_read_pixels.restype = C.c_int
    L.rsh_engine_read_pixels.argtypes = [PV, PU, C.c_size_t]
    L.rsh_engine_read_pixels_async.restype = C.c_int
    L.rsh_engine_read_pixels_async.argtypes = [PV, PU, C.c_size_t]
    L.rsh_engine_wait_pixels.restype = C.c_int
    L.rsh_engine_wait_pixels.argtypes = [PV]
    L.rsh_engine_read_gbuffer.restype = C.c_int
    L.rsh_engine_read_gbuffer.argtypes = [PV, PF, PF, PF, PF, C.c_size_t]
    L.rsh_engine_read_pass.restype = C.c_int
    L.rsh_engine_read_pass.argtypes = [PV, u32, PU64, C.c_size_t]
    L.rsh_engine_pixels_device.restype = PV
    L.rsh_engine_pixels_device.argtypes = [PV]
    L.rsh_device_alloc.restype = PV
    L.rsh_device_alloc.argtypes = [C.c_size_t]
    L.rsh_device_free.argtypes = [PV]
    L.rsh_memcpy_htod.restype = C.c_int
    L.rsh_memcpy_htod.argtypes = [PV, PV, C.c_size_t]
    L.rsh_memcpy_dtoh.restype = C.c_int
    L.rsh_memcpy_dtoh.argtypes = [PV, PV, C.c_size_t]
    L.rsh_memset_d32.restype = C.c_int
    L.rsh_memset_d32.argtypes = [PV, u32, C.c_size_t]
    L.rsh_device_synchronize.restype = C.c_int
    L.rsh_device_count.restype = C.c_int
    L.rsh_compat_configure.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_int]
    L.rsh_compat_set_stream.argtypes = [PV]
    L.rsh_compat_set_exit_on_error.argtypes = [C.c_int]
    L.rsh_compat_status.restype = C.c_int
    L.rsh_compat_error.restype = C.c_char_p
    L.rsh_selfcheck_rcp.restype = C.c_int
    L.rsh_selfcheck_rcp.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.rsh_engine_set_wait_timeout.restype = C.c_int
    L.rsh_engine_set_wait_timeout.argtypes = [PV, u32]
    L.rsh_selfcheck_wait_deadline.restype = C.c_int
    L.rsh_selfcheck_wait_deadline.argtypes = [u32, u32, C.POINTER(u32)]
    for name, args in COMPAT_SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = args
    for name, args in CTX_SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = [PV, PV] + args
    L.rsh_ctx_create.restype = C.c_int
    L.rsh_ctx_create.argtypes = [C.c_int, C.POINTER(PV)]
    L.rsh_ctx_destroy.restype = None
    L.rsh_ctx_destroy.argtypes = [PV]
    L.rsh_ctx_configure.restype = C.c_int
    L.rsh_ctx_configure.argtypes = [PV, C.c_uint, C.c_uint, C.c_uint, C.c_int]
    L.rsh_ctx_last_error.restype = C.c_char_p
    L.rsh_ctx_last_error.argtypes = [PV]
    L.rsh_set_gbuffer.restype = C.c_int
    L.rsh_set_gbuffer.argtypes = [PV, PV, PV, PV, PV, C.c_uint, C.c_uint]
    L.rsh_set_scene.restype = C.c_int
    L.rsh_set_scene.argtypes = [PV, PV, PV, PV, PV, C.c_uint, PV, PV, C.c_uint, PV, C.c_uint, PV, PV, PV, C.c_uint]
    for name in BINDERS:
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = [PV] if "TextureArray" in name or name in ("bindFragmentPositionArray",
                                                                   "bindFragmentNormalArray") else [PV, C.c_uint]
    _lib = L
    return L


def selfcheck_rcp(device=0):
    """(tested, mismatches) of the device reciprocal check (rsh_selfcheck_rcp)."""
    t, m = C.c_uint64(0), C.c_uint64(0)
    rc = lib().rsh_selfcheck_rcp(device, C.byref(t), C.byref(m))
    if rc != 0:
        raise RshError(f"rsh_selfcheck_rcp failed with HIP error {rc}")
    return t.value, m.value


def selfcheck_wait_deadline(wait_ms, block_ms):
    """(status, elapsed ms) of the bounded host wait on a word that is never
    published, the stream held busy by a host function for block_ms
    (rsh_selfcheck_wait_deadline)."""
    el = u32(0)
    st = lib().rsh_selfcheck_wait_deadline(wait_ms, block_ms, C.byref(el))
    return st, el.value


SPHERE_MINIMAL, SPHERE_REFERENCE = 0, 1


def object_spheres(tri_pos, ranges, mode=SPHERE_MINIMAL):
    """Object bounding spheres (rsh_object_spheres, host code): tri_pos float32
    [3T, 4] (or [T, 3, 4]), ranges uint32 [K, 2] (first, last triangle) ->
    float32 [K, 6] (centre | radius, first, last), the smallest ball around
    each object's vertices (SPHERE_MINIMAL) or with the reference's radius
    (SPHERE_REFERENCE)."""
    pos = np.ascontiguousarray(tri_pos, np.float32).reshape(-1, 4)
    rg = np.ascontiguousarray(ranges, np.uint32).reshape(-1, 2)
    out = np.empty((rg.shape[0], 6), np.float32)
    L = lib()
    L.rsh_object_spheres.restype = C.c_int
    L.rsh_object_spheres.argtypes = [PV, u32, PV, u32, u32, PV]
    rc = L.rsh_object_spheres(pos.ctypes.data, pos.shape[0] // 3, rg.ctypes.data, rg.shape[0], mode,
                              out.ctypes.data)
    if rc != 0:
        raise RshError(f"rsh_object_spheres failed ({rc})")
    return out


# ------------------------------------------------------------ device memory --
class DeviceArray:
    """A device allocation holding a numpy-shaped array (owned; freed on close)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().rsh_device_alloc(max(self.nbytes, 4))
        if not self.ptr:
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes <= self.nbytes
        if lib().rsh_memcpy_htod(self.ptr, arr.ctypes.data, arr.nbytes) != 0:
            raise RuntimeError("host->device copy failed")

    def download(self, count=None):
        out = np.empty(self.shape, self.dtype)
        if count is not None:
            out = out.reshape(-1)[:count]
        if lib().rsh_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes) != 0:
            raise RuntimeError("device->host copy failed")
        return out

    def fill_u32(self, value):
        if lib().rsh_memset_d32(self.ptr, value, self.nbytes // 4) != 0:
            raise RuntimeError("fill failed")

    def close(self):
        if self.ptr:
            lib().rsh_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    return lib().rsh_device_count()


# ------------------------------------------------------------------ engine --
def _fp(a):
    return a.ctypes.data_as(PF)


def scene_desc(s):
    arrs = [np.ascontiguousarray(x) for x in (s.tri_pos, s.tri_nrm, s.tri_mat, s.mat_diffuse, s.mat_specular,
                                               s.object_spheres, s.light_pos, s.light_col, s.light_int)]
    d = SceneDesc(arrs[0].shape[0] // 3, _fp(arrs[0]), _fp(arrs[1]), arrs[2].ctypes.data_as(PI), arrs[3].shape[0],
                  _fp(arrs[3]), _fp(arrs[4]), arrs[5].shape[0], _fp(arrs[5]), arrs[6].shape[0], _fp(arrs[6]),
                  _fp(arrs[7]), _fp(arrs[8]))
    d.keep = arrs
    return d


def camera_desc(cam):
    c = CameraDesc()
    c.eye[:] = [float(v) for v in cam.eye]
    c.direction[:] = [float(v) for v in cam.direction]
    c.right[:] = [float(v) for v in cam.right]
    c.up[:] = [float(v) for v in cam.up]
    return c


def config_desc(cfg):
    o = FrameConfigDesc()
    for name, _ in FrameConfigDesc._fields_:
        v = getattr(cfg, name)
        if name in ("scene_min", "cell_scale"):
            getattr(o, name)[:] = [float(x) for x in v]
        else:
            setattr(o, name, v)
    return o


class Context:
    """A stage context (include/rsh_ctx.h): ctx.stage(name, stream, *args)
    calls rsh_<name>(ctx, stream, *args) and raises RshError on failure."""

    def __init__(self, device=0):
        L = lib()
        p = PV()
        rc = L.rsh_ctx_create(device, C.byref(p))
        if rc != 0:
            raise RshError(f"rsh_ctx_create({device}) failed ({rc})")
        self.ptr = p

    def _ok(self, rc, what):
        if rc != 0:
            raise RshError(f"{what}: {lib().rsh_ctx_last_error(self.ptr).decode()} (status {rc})")

    def configure(self, light_max, depth, sub, algorithm):
        self._ok(lib().rsh_ctx_configure(self.ptr, light_max, depth, sub, algorithm), "rsh_ctx_configure")

    def set_gbuffer(self, dif, spec, pos, nrm, w, h):
        self._ok(lib().rsh_set_gbuffer(self.ptr, dif, spec, pos, nrm, w, h), "rsh_set_gbuffer")

    def set_scene(self, *args):
        self._ok(lib().rsh_set_scene(self.ptr, *args), "rsh_set_scene")

    def stage(self, name, stream, *args):
        fn = name if name.startswith("rsh_") else "rsh_" + name
        self._ok(getattr(lib(), fn)(self.ptr, stream, *args), fn)

    def close(self):
        if self.ptr:
            lib().rsh_ctx_destroy(self.ptr)
            self.ptr = None


class RshError(RuntimeError):
    pass


class Engine:
    """Device-resident scene + frame renderer (include/rsh.h)."""

    KEEP_PASSES, COUNTERS, TIMING, GRAPH, TRAV_TIMING, SERIAL, KEEP_ORDER = 1, 2, 4, 8, 16, 32, 64

    def __init__(self, device: int = 0):
        self.L = lib()
        h = PV()
        rc = self.L.rsh_engine_create(C.byref(h), device)
        if rc != 0:
            raise RshError(f"rsh_engine_create failed ({rc})")
        self.h = h
        self.cfg = None
        self.stats = None

    def _check(self, rc, what):
        if rc != 0:
            raise RshError(f"{what} failed ({rc}): {self.L.rsh_last_error(self.h).decode()}")

    def load_scene(self, scene):
        self._scene = scene
        self._check(self.L.rsh_engine_load_scene(self.h, C.byref(scene_desc(scene))), "load_scene")

    def update_objects(self, model, normal, scale):
        """Move the loaded scene's objects (rsh_engine_update_objects): float32
        [K, 16] row-major model, normal and scale matrices, applied to the scene
        as loaded."""
        K = self._scene.object_total
        m, n, s = (np.ascontiguousarray(a, np.float32).reshape(K, 16) for a in (model, normal, scale))
        self.L.rsh_engine_update_objects.restype = C.c_int
        self.L.rsh_engine_update_objects.argtypes = [PV, PV, PV, PV]
        self._check(self.L.rsh_engine_update_objects(self.h, m.ctypes.data, n.ctypes.data, s.ctypes.data),
                    "update_objects")

    def render(self, camera, cfg, flags=0) -> FrameStats:
        self.cfg = cfg
        st = FrameStats()
        self._check(self.L.rsh_engine_render(self.h, C.byref(camera_desc(camera)), C.byref(config_desc(cfg)),
                                             flags, C.byref(st)), "render")
        self.stats = st
        return st

    def render_views(self, views, cfg, flags=0) -> FrameStats:
        """Render a batch of (camera, width, height) views as one image of
        sum(w*h) pixels; pixels() / gbuffer() / pass_hits() then address that
        1 x P image (a single view keeps its H x W shape)."""
        n = len(views)
        vd = (ViewDesc * n)()
        for i, (cam, w, h) in enumerate(views):
            vd[i].camera = camera_desc(cam)
            vd[i].width, vd[i].height = int(w), int(h)
        P = sum(int(w) * int(h) for _, w, h in views)
        W, H = (int(views[0][1]), int(views[0][2])) if n == 1 else (P, 1)
        self.cfg = dataclasses.replace(cfg, width=W, height=H)
        st = FrameStats()
        self._check(self.L.rsh_engine_render_views(self.h, vd, n, C.byref(config_desc(cfg)), flags, C.byref(st)),
                    "render_views")
        self.stats = st
        return st

    def trace(self, rays, kind, cfg, tmax=None, segments=None, flags=0):
        """Ray query through the GPU pipeline: rays float32 [n, 6] (origin,
        direction); kind 0 closest hit -> uint64 records, kind 1 any hit with
        per-ray tmax -> uint32 (0xFFFFFFFF occluded)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        bufs = [DeviceArray.from_host(rays)]
        tm = sg = None
        if tmax is not None:
            tm = DeviceArray.from_host(np.ascontiguousarray(tmax, np.float32))
            bufs.append(tm)
        if segments is not None:
            sg = DeviceArray.from_host(np.ascontiguousarray(segments, np.uint32))
            bufs.append(sg)
        out = DeviceArray((n,), np.uint32 if kind == 1 else np.uint64)
        bufs.append(out)
        try:
            st = FrameStats()
            self._check(self.L.rsh_engine_trace(self.h, bufs[0].ptr, tm.ptr if tm else None, sg.ptr if sg else None,
                                                n, kind, C.byref(config_desc(cfg)), out.ptr, flags, C.byref(st)),
                        "trace")
            self.stats = st
            return out.download()
        finally:
            for b in bufs:
                b.close()

    @staticmethod
    def split_views(pixels, views):
        """Per-view H x W images from the pixels of a batch."""
        flat = np.asarray(pixels).reshape(-1)
        out, off = [], 0
        for _, w, h in views:
            out.append(flat[off:off + w * h].reshape(h, w))
            off += w * h
        return out

    def pixels(self):
        W, H = self.cfg.width, self.cfg.height
        out = np.empty(W * H, np.uint32)
        self._check(self.L.rsh_engine_read_pixels(self.h, out.ctypes.data_as(PU), out.size), "read_pixels")
        return out.reshape(H, W)

    def _pinned_pixels(self):
        W, H = self.cfg.width, self.cfg.height
        n = W * H
        if getattr(self, "_pinned_n", 0) < n:
            self.wait_pixels()
            self._free_pinned()
            self.L.rsh_host_alloc.restype = PV
            self.L.rsh_host_alloc.argtypes = [C.c_size_t]
            p = self.L.rsh_host_alloc(4 * n)
            if not p:
                raise MemoryError("page-locked host allocation failed")
            self._pinned, self._pinned_n = p, n
        return n, np.ctypeslib.as_array(C.cast(self._pinned, PU), shape=(n,)).reshape(H, W)

    def read_pixels_pinned(self):
        """The frame's pixels copied into a page-locked host buffer owned by
        this engine (rsh_host_alloc; reused while the size fits): one DMA
        transfer on the engine's stream.  Returns an H x W view of that buffer,
        overwritten by the next call."""
        n, view = self._pinned_pixels()
        self._check(self.L.rsh_engine_read_pixels(self.h, C.cast(self._pinned, PU), n), "read_pixels")
        return view

    def read_pixels_async(self):
        """Queue the frame's pixels for copying into the page-locked buffer of
        read_pixels_pinned and return at once (rsh_engine_read_pixels_async);
        wait_pixels() returns the H x W view once the copy has ended.  The next
        frame's passes run beside the copy."""
        n, _ = self._pinned_pixels()
        self._check(self.L.rsh_engine_read_pixels_async(self.h, C.cast(self._pinned, PU), n), "read_pixels_async")

    def wait_pixels(self):
        """Wait for read_pixels_async's copy; the H x W view of its buffer."""
        self._check(self.L.rsh_engine_wait_pixels(self.h), "wait_pixels")
        if not getattr(self, "_pinned", None):
            return None
        W, H = self.cfg.width, self.cfg.height
        return np.ctypeslib.as_array(C.cast(self._pinned, PU), shape=(W * H,)).reshape(H, W)

    def _free_pinned(self):
        if getattr(self, "_pinned", None):
            self.L.rsh_host_free.argtypes = [PV]
            self.L.rsh_host_free(self._pinned)
            self._pinned, self._pinned_n = None, 0

    def gbuffer(self):
        n = self.cfg.width * self.cfg.height
        g = {k: np.empty((n, 4), np.float32) for k in ("dif", "spec", "pos", "nrm")}
        self._check(self.L.rsh_engine_read_gbuffer(self.h, _fp(g["dif"]), _fp(g["spec"]), _fp(g["pos"]),
                                                   _fp(g["nrm"]), n), "read_gbuffer")
        return g

    def pass_hits(self, i):
        n = self.stats.pass_slots[i]
        out = np.empty(n, np.uint64)
        self._check(self.L.rsh_engine_read_pass(self.h, i, out.ctypes.data_as(PU64), n), "read_pass")
        return out

    def pass_order(self, i):
        """Pass i's keys, flags, sorted order, packets and hierarchy nodes
        (render flag KEEP_ORDER): a dict with "order" (slot of each sorted
        ray), "nodes" (float32 [2 * nodes, 4]), and "keys" / "flags" (per slot)
        and "bases" / "ckeys" (per chunk) where the pass has them."""
        st = self.stats
        S, N, Cn = int(st.pass_slots[i]), int(st.pass_rays[i]), int(st.pass_chunks[i])
        nn = 0
        n = (N + self.cfg.leaf - 1) // self.cfg.leaf
        for _ in range(self.cfg.depth):
            nn += n
            n = (n + self.cfg.branch - 1) // self.cfg.branch
        L = self.L
        L.rsh_engine_read_pass_order.restype = C.c_int
        L.rsh_engine_read_pass_order.argtypes = [PV, u32, PV, PV, C.c_size_t, PV, C.c_size_t, PV, PV, C.c_size_t,
                                                 PV, C.c_size_t]
        out = {"order": np.empty(N, np.uint32), "nodes": np.empty((2 * nn, 4), np.float32)}
        self._check(L.rsh_engine_read_pass_order(self.h, i, None, None, 0, out["order"].ctypes.data, N, None, None, 0,
                                                 out["nodes"].ctypes.data, 2 * nn), "read_pass_order")
        keys, flags = np.empty(S, np.uint32), np.empty(S, np.uint32)
        if L.rsh_engine_read_pass_order(self.h, i, keys.ctypes.data, flags.ctypes.data, S, None, 0, None, None, 0,
                                        None, 0) == 0:
            out["keys"], out["flags"] = keys, flags
        bases, ckeys = np.empty(Cn, np.uint32), np.empty(Cn, np.uint32)
        if L.rsh_engine_read_pass_order(self.h, i, None, None, 0, None, 0, bases.ctypes.data, ckeys.ctypes.data, Cn,
                                        None, 0) == 0:
            out["bases"], out["ckeys"] = bases, ckeys
        return out

    def passes(self):
        st = self.stats
        return [dict(kind=st.pass_kind[i], slots=st.pass_slots[i], rays=st.pass_rays[i], chunks=st.pass_chunks[i],
                     pairs_top=st.pass_pairs_top[i], pairs_leaf=st.pass_pairs_leaf[i],
                     ray_tests=st.pass_ray_tests[i], top_survivors=st.pass_top_survivors[i],
                     traverse_ms=st.pass_traverse_ms[i], splits=st.pass_splits[i], items=st.pass_items[i],
                     chunk=st.pass_chunk[i], tri_tests=st.pass_tri_tests[i],
                     cluster_tests=st.pass_cluster_tests[i], leaf_steps=st.pass_leaf_steps[i],
                     ray_steps=st.pass_ray_steps[i], mt_tests=st.pass_mt_tests[i], mode=st.pass_mode[i])
                for i in range(st.pass_count)]

    def close(self):
        if getattr(self, "h", None):
            self.L.rsh_engine_wait_pixels(self.h)      # no copy left writing the page-locked buffer
        self._free_pinned()
        if getattr(self, "h", None):
            self.L.rsh_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def image_rgb(pixels: np.ndarray) -> np.ndarray:
    """RGBA8-packed pixels (R | G<<8 | B<<16) -> uint8 [H, W, 3], top row first."""
    p = pixels.astype(np.uint32)
    rgb = np.stack([p & 255, (p >> 8) & 255, (p >> 16) & 255], -1).astype(np.uint8)
    return rgb[::-1]


def write_ppm(path: str, pixels: np.ndarray):
    rgb = image_rgb(pixels)
    with open(path, "wb") as f:
        f.write(b"P6 %d %d 255\n" % (rgb.shape[1], rgb.shape[0]))
        f.write(rgb.tobytes())
