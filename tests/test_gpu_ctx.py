"""Per-context stage API (include/rsh_ctx.h) against the oracle.

The reference frame of test_gpu_compat (shadow rays from a G-buffer, trim,
compress, sort, decompress, hierarchy, warm-up and lower-level traversal,
shadow intersection and shading, then a mirror bounce; main.cpp:630-1054,
1234-1303) runs through rsh_<stage>(ctx, stream, ...) on two contexts, each
on its own host thread and HIP stream at the same time, with the ray and
chunk totals kept in device words from trimming to hierarchy creation (no
host count until the traversal).  Every buffer the reference's driver reads
afterwards is compared with the oracle bit for bit.  The device-count chain
is also captured in a HIP graph and replayed on new rays.
"""
import ctypes as C
import os
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIGHT_MAX, DEPTH, SUB, TB = 1, 2, 8, 5000


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    for fn in ("hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize", "hipStreamBeginCapture",
               "hipStreamEndCapture", "hipGraphInstantiate", "hipGraphLaunch", "hipGraphExecDestroy",
               "hipGraphDestroy"):
        getattr(h, fn).restype = C.c_int
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
    h.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    h.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
    h.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
    h.hipGraphExecDestroy.argtypes = [C.c_void_p]
    h.hipGraphDestroy.argtypes = [C.c_void_p]
    return h


def new_stream(h):
    s = C.c_void_p()
    assert h.hipStreamCreate(C.byref(s)) == 0
    return s


def f3(v):
    import rsh_amd
    return rsh_amd.Float3(*[float(x) for x in v])


@pytest.fixture(scope="module")
def frame(rsh, oracle):
    """Scene, G-buffer and the oracle's buffers after each step of the frame."""
    scene = rsh.scenes.config0_scene()
    W = H = 96
    cam = rsh.scenes.config0_camera(W, H)
    cfg = rsh.config.config0(W, H).with_scene_bounds(scene)
    gb = oracle.render(scene, cam, cfg, keep_hits=False).gbuffer
    return scene, cam, gb, W, H


class Frame:
    """Device buffers of one run of the reference frame (a context's own)."""

    def __init__(self, rsh, scene, gb, W, H):
        D = rsh.DeviceArray.from_host
        T, K = scene.triangle_total, scene.object_total
        self.S = S = W * H * LIGHT_MAX
        self.W, self.H, self.T, self.K = W, H, T, K
        self.pos, self.nrm = D(scene.tri_pos), D(scene.tri_nrm)
        self.obj = D(np.repeat(scene.triangle_object(), 3).astype(np.int32))
        self.mat = D(np.ascontiguousarray(scene.tri_mat, np.int32))
        self.mdif, self.mspec = D(scene.mat_diffuse), D(scene.mat_specular)
        sph = np.zeros((2 * K, 3), np.float32)
        sph[0::2] = scene.object_spheres[:, :3]
        sph[1::2] = scene.object_spheres[:, 3:6]
        self.sph = D(sph)
        self.lpos, self.lcol, self.lint = D(scene.light_pos), D(scene.light_col), D(scene.light_int)
        self.gdif, self.gspec, self.gpos, self.gnrm = D(gb["dif"]), D(gb["spec"]), D(gb["pos"]), D(gb["nrm"])
        n0 = (S + SUB - 1) // SUB
        self.hit_mem = (n0 // 2) * 10000
        U = lambda n: rsh.DeviceArray((n,), np.uint32)
        self.rays = rsh.DeviceArray((2 * S, 3), np.float32)
        self.pkeys, self.pvals, self.skeys, self.svals = U(S), U(S), U(S), U(S)
        self.bases, self.sizes, self.ckeys, self.cvals, self.sckeys, self.scvals = (U(S) for _ in range(6))
        self.flags, self.scan, self.hits, self.thits = (U(self.hit_mem) for _ in range(4))
        self.nodes_total = n0 + (n0 + SUB - 1) // SUB
        self.hier = rsh.DeviceArray((2 * self.nodes_total, 4), np.float32)
        self.sflags, self.pixels = U(S), D(np.zeros(W * H, np.uint32))
        self.counts = U(4)            # device words: ray total, chunk total


def bind(ctx, fr, light_total):
    ctx.configure(LIGHT_MAX, DEPTH, SUB, 0)
    ctx.set_gbuffer(fr.gdif.ptr, fr.gspec.ptr, fr.gpos.ptr, fr.gnrm.ptr, fr.W, fr.H)
    ctx.set_scene(fr.pos.ptr, fr.nrm.ptr, fr.obj.ptr, fr.mat.ptr, fr.T, fr.mdif.ptr, fr.mspec.ptr,
                  fr.mdif.shape[0], fr.sph.ptr, fr.K, fr.lpos.ptr, fr.lcol.ptr, fr.lint.ptr, light_total)


def cast_device_counts(ctx, s, fr):
    """Trim -> compress -> sort -> decompress -> hierarchy with the ray and
    chunk totals in fr.counts (device); nothing waits for the host."""
    S, W, H = fr.S, fr.W, fr.H
    n_dev, c_dev = fr.counts.ptr, fr.counts.ptr + 4
    ctx.stage("ray_trimming", s, fr.pkeys.ptr, fr.pvals.ptr, W, H, LIGHT_MAX, fr.flags.ptr, fr.scan.ptr,
              fr.skeys.ptr, fr.svals.ptr, None, n_dev)
    ctx.stage("ray_compression", s, fr.skeys.ptr, fr.svals.ptr, S, n_dev, fr.flags.ptr, fr.scan.ptr, fr.bases.ptr,
              fr.sizes.ptr, fr.ckeys.ptr, fr.cvals.ptr, None, c_dev)
    ctx.stage("ray_sorting", s, fr.ckeys.ptr, fr.cvals.ptr, S, c_dev, fr.sckeys.ptr, fr.scvals.ptr)
    ctx.stage("ray_decompression", s, fr.bases.ptr, fr.sizes.ptr, fr.sckeys.ptr, fr.scvals.ptr, fr.skeys.ptr,
              fr.svals.ptr, fr.flags.ptr, fr.scan.ptr, S, c_dev, fr.pkeys.ptr, fr.pvals.ptr)
    ctx.stage("hierarchy_creation", s, fr.rays.ptr, fr.pkeys.ptr, fr.pvals.ptr, S, n_dev, 0.0, 0.0, fr.hier.ptr)


def run_frame(rsh, ctx, s, fr, eye, light_total):
    """The frame's shadow iteration and first mirror bounce on (ctx, s);
    returns the buffers the checks read, in order."""
    h = hip()
    S, W, H, T, K = fr.S, fr.W, fr.H, fr.T, fr.K
    out = {}

    def totals():
        assert h.hipStreamSynchronize(s) == 0
        n, c = fr.counts.download()[:2]
        return int(n), int(c)

    def traverse(N, shadow):
        for i in range((T + TB - 1) // TB):
            off = i * TB
            tb = TB - max(TB * (i + 1) - T, 0)
            ht, mem = C.c_uint32(0), C.c_uint32(fr.hit_mem)
            ctx.stage("hierarchy_traversal_warmup", s, fr.hier.ptr, fr.sph.ptr, fr.pos.ptr, K, N, tb, off,
                      fr.flags.ptr, fr.scan.ptr, fr.hits.ptr, fr.thits.ptr, C.byref(ht), C.byref(mem))
            ctx.stage("hierarchy_traversal", s, fr.hier.ptr, fr.pos.ptr, N, tb, off, fr.flags.ptr, fr.scan.ptr,
                      fr.hits.ptr, fr.thits.ptr, C.byref(ht), C.byref(mem))
            if shadow:
                ctx.stage("shadow_ray_intersection", s, fr.rays.ptr, fr.pkeys.ptr, fr.pvals.ptr, fr.thits.ptr,
                          fr.pos.ptr, ht.value, N, off, W, H, f3(eye), fr.sflags.ptr)
            else:
                ctx.stage("reflection_ray_intersection", s, fr.rays.ptr, fr.pkeys.ptr, fr.pvals.ptr, fr.thits.ptr,
                          fr.pos.ptr, ht.value, N, off, W, H, 1, f3(eye), fr.sflags.ptr)

    ctx.stage("memory_preparation", s, W, H, fr.hit_mem)
    ctx.stage("shadow_ray_creation", s, W, H, LIGHT_MAX, fr.rays.ptr, fr.flags.ptr, fr.pkeys.ptr, fr.pvals.ptr)
    cast_device_counts(ctx, s, fr)
    N, Cn = totals()
    out["shadow_totals"] = (N, Cn)
    out["shadow_sorted"] = (fr.pkeys.download()[:N], fr.pvals.download()[:N])
    ctx.stage("shadow_ray_preparation", s, W, H, LIGHT_MAX, fr.sflags.ptr)
    traverse(N, True)
    ctx.stage("shadow_ray_coloring", s, W, H, light_total, f3(eye), fr.sflags.ptr, fr.pixels.ptr)
    assert h.hipStreamSynchronize(s) == 0
    out["shadow_flags"] = fr.sflags.download()
    out["pixels_shadow"] = fr.pixels.download()
    ctx.stage("reflection_ray_creation", s, W, H, f3(eye), fr.rays.ptr, fr.flags.ptr, fr.pkeys.ptr, fr.pvals.ptr)
    cast_device_counts(ctx, s, fr)
    N, Cn = totals()
    out["refl_totals"] = (N, Cn)
    ctx.stage("reflection_ray_preparation", s, W, H, fr.sflags.ptr)
    traverse(N, False)
    ctx.stage("reflection_ray_coloring", s, fr.pos.ptr, fr.nrm.ptr, W, H, light_total, f3(eye), True, fr.sflags.ptr,
              fr.rays.ptr, fr.flags.ptr, fr.pkeys.ptr, fr.pvals.ptr, fr.pixels.ptr)
    assert h.hipStreamSynchronize(s) == 0
    out["records"] = fr.sflags.download()
    out["pixels"] = fr.pixels.download()
    return out


def oracle_frame(oracle, scene, gb, W, H, eye):
    """The same steps through the oracle's wrapper restatements."""
    S = W * H * LIGHT_MAX
    T, K = scene.triangle_total, scene.object_total
    oracle.call("oracle_set_constants", LIGHT_MAX, DEPTH, SUB, 0)
    osc = oracle.make_scene(scene)
    oracle.lib().oracle_bind_scene(C.byref(osc))
    oracle.call("oracle_bind_gbuffer", gb["dif"], gb["spec"], gb["pos"], gb["nrm"])
    obj = np.repeat(scene.triangle_object(), 3).astype(np.int32)
    sph4 = np.zeros((2 * K, 4), np.float32)
    sph4[0::2, :3] = scene.object_spheres[:, :3]
    sph4[1::2, :3] = scene.object_spheres[:, 3:6]
    sph4[:, 3] = 1.0
    oracle.call("oracle_bind_original", scene.tri_pos, scene.tri_nrm, obj, sph4)
    sph = np.ascontiguousarray(sph4[:, :3])
    n0 = (S + SUB - 1) // SUB
    hit_mem = (n0 // 2) * 10000
    z = lambda n: np.zeros(n, np.uint32)
    rays = np.zeros((2 * S, 3), np.float32)
    pkeys, pvals, skeys, svals = z(S), z(S), z(S), z(S)
    bases, sizes, ckeys, cvals, sckeys, scvals = (z(S) for _ in range(6))
    flags, scan, hits, thits = (z(hit_mem) for _ in range(4))
    hier = np.zeros((2 * (n0 + (n0 + SUB - 1) // SUB), 4), np.float32)
    sflags, pixels = z(S), z(W * H)
    tri, nrm = np.ascontiguousarray(scene.tri_pos), np.ascontiguousarray(scene.tri_nrm)
    out = {}

    def cast():
        tot, ct = C.c_uint32(0), C.c_uint32(0)
        oracle.call("oracle_RayTrimmingWrapper", pkeys, pvals, W, H, LIGHT_MAX, flags, scan, skeys, svals,
                    C.byref(tot))
        N = tot.value
        oracle.call("oracle_RayCompressionWrapper", skeys, svals, N, flags, scan, bases, sizes, ckeys, cvals,
                    C.byref(ct))
        Cn = ct.value
        oracle.call("oracle_RaySortingWrapper", ckeys, cvals, Cn, sckeys, scvals)
        oracle.call("oracle_RayDecompressionWrapper", bases, sizes, sckeys, scvals, skeys, svals, flags, scan, Cn,
                    pkeys, pvals)
        oracle.call("oracle_HierarchyCreationWrapper", rays, pkeys, pvals, N, 0.0, 0.0, hier)
        return N, Cn

    def traverse(N, shadow):
        for i in range((T + TB - 1) // TB):
            off = i * TB
            tb = TB - max(TB * (i + 1) - T, 0)
            ht = C.c_uint32(0)
            oracle.call("oracle_HierarchyTraversalWarmUpWrapper", hier, sph, tri, K, N, tb, off, flags, scan, hits,
                        thits, C.byref(ht))
            oracle.call("oracle_HierarchyTraversalWrapper", hier, tri, N, tb, off, flags, scan, hits, thits,
                        C.byref(ht))
            if shadow:
                oracle.call("oracle_ShadowRayIntersectionWrapper", rays, pkeys, pvals, thits, tri, ht.value, N, off,
                            W, H, sflags)
            else:
                oracle.call("oracle_ReflectionRayIntersectionWrapper", rays, pkeys, pvals, thits, tri, ht.value, N,
                            off, W, H, sflags)

    oracle.call("oracle_ShadowRayCreationWrapper", W, H, LIGHT_MAX, rays, flags, pkeys, pvals)
    N, Cn = cast()
    out["shadow_totals"] = (N, Cn)
    out["shadow_sorted"] = (pkeys[:N].copy(), pvals[:N].copy())
    oracle.call("oracle_ShadowRayPreparationWrapper", W, H, LIGHT_MAX, sflags)
    traverse(N, True)
    oracle.call("oracle_ShadowRayColoringWrapper", W, H, scene.light_total, eye, sflags, pixels)
    out["shadow_flags"] = sflags.copy()
    out["pixels_shadow"] = pixels.copy()
    oracle.call("oracle_ReflectionRayCreationWrapper", W, H, eye, rays, flags, pkeys, pvals)
    out["refl_totals"] = cast()
    oracle.call("oracle_ReflectionRayPreparationWrapper", W, H, sflags)
    traverse(out["refl_totals"][0], False)
    oracle.call("oracle_ReflectionRayColoringWrapper", tri, nrm, W, H, scene.light_total, eye, True, sflags, rays,
                flags, pkeys, pvals, pixels)
    out["records"] = sflags.copy()
    out["pixels"] = pixels.copy()
    return out


def compare(got, want, who):
    assert got["shadow_totals"] == want["shadow_totals"], who
    assert got["refl_totals"] == want["refl_totals"], who
    for i in range(2):
        np.testing.assert_array_equal(got["shadow_sorted"][i], want["shadow_sorted"][i], err_msg=who)
    for k in ("shadow_flags", "pixels_shadow", "records", "pixels"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{who}: {k}")


def test_two_contexts_on_two_threads_match_the_oracle(rsh, oracle, frame):
    scene, cam, gb, W, H = frame
    eye = np.asarray(cam.eye, np.float32)
    want = oracle_frame(oracle, scene, gb, W, H, eye)
    assert want["shadow_totals"][0] > 0 and want["refl_totals"][0] > 0
    h = hip()
    ctxs, streams, frames, results, errors = [], [], [], [None, None], [None, None]
    for _ in range(2):
        ctx = rsh.Context(0)
        fr = Frame(rsh, scene, gb, W, H)
        bind(ctx, fr, scene.light_total)
        ctxs.append(ctx); frames.append(fr); streams.append(new_stream(h))

    def work(i):
        try:
            results[i] = run_frame(rsh, ctxs[i], streams[i], frames[i], eye, scene.light_total)
        except Exception as ex:          # reported below on the main thread
            errors[i] = ex

    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i in range(2):
        assert errors[i] is None, errors[i]
        compare(results[i], want, f"context {i}")
    for ctx, s in zip(ctxs, streams):
        ctx.close()
        h.hipStreamDestroy(s)


def test_errors_stay_in_their_context(rsh, frame):
    """A failing call reports through its own context (status and message)
    and leaves another context and the drop-in wrappers' status untouched."""
    scene, cam, gb, W, H = frame
    a, b = rsh.Context(0), rsh.Context(0)
    L = rsh.lib()
    L.rsh_compat_set_exit_on_error(0)
    L.rsh_compat_reset_status()
    d = rsh.DeviceArray((64,), np.uint32)
    rc = L.rsh_shadow_ray_creation(a.ptr, None, W, H, 1, d.ptr, d.ptr, d.ptr, d.ptr)     # nothing bound
    assert rc == 1 and b"not bound" in L.rsh_ctx_last_error(a.ptr)
    assert L.rsh_ctx_last_error(b.ptr) == b""
    assert L.rsh_compat_status() == 0
    assert L.rsh_ctx_configure(b.ptr, 0, 2, 8, 0) == 1
    with pytest.raises(rsh.RshError):
        a.stage("ctx_sort_pairs", None, d.ptr, d.ptr, 64, None, 0, d.ptr, d.ptr)        # bits 0
    a.close(); b.close()


def test_device_count_chain_replays_in_a_graph(rsh, oracle, frame):
    """Trim -> compress -> sort -> decompress -> hierarchy with device counts,
    captured once in a HIP graph (no host count, no allocation inside) and
    replayed on the shadow rays and then the reflection rays of the frame:
    each replay equals the oracle's chain on the same rays."""
    scene, cam, gb, W, H = frame
    eye = np.asarray(cam.eye, np.float32)
    h = hip()
    ctx = rsh.Context(0)
    fr = Frame(rsh, scene, gb, W, H)
    bind(ctx, fr, scene.light_total)
    s = new_stream(h)
    S = fr.S
    ctx.stage("memory_preparation", s, W, H, fr.hit_mem)
    assert h.hipStreamSynchronize(s) == 0
    # capture (relaxed mode) the device-count chain on s
    assert h.hipStreamBeginCapture(s, 2) == 0
    cast_device_counts(ctx, s, fr)
    g = C.c_void_p()
    assert h.hipStreamEndCapture(s, C.byref(g)) == 0
    ge = C.c_void_p()
    assert h.hipGraphInstantiate(C.byref(ge), g, None, None, 0) == 0

    # the oracle's chain on the same inputs
    oracle.call("oracle_set_constants", LIGHT_MAX, DEPTH, SUB, 0)
    oracle.call("oracle_bind_gbuffer", gb["dif"], gb["spec"], gb["pos"], gb["nrm"])
    osc = oracle.make_scene(scene)
    oracle.lib().oracle_bind_scene(C.byref(osc))
    z = lambda n: np.zeros(n, np.uint32)
    o_rays = np.zeros((2 * S, 3), np.float32)
    o = {k: z(S) for k in ("pk", "pv", "sk", "sv", "b", "sz", "ck", "cv", "sck", "scv")}
    o_flags, o_scan = z(fr.hit_mem), z(fr.hit_mem)
    o_hier = np.zeros((2 * fr.nodes_total, 4), np.float32)
    for kind in ("shadow", "reflection", "shadow"):
        if kind == "shadow":
            ctx.stage("shadow_ray_creation", s, W, H, LIGHT_MAX, fr.rays.ptr, fr.flags.ptr, fr.pkeys.ptr,
                      fr.pvals.ptr)
            oracle.call("oracle_ShadowRayCreationWrapper", W, H, LIGHT_MAX, o_rays, o_flags, o["pk"], o["pv"])
        else:
            ctx.stage("reflection_ray_creation", s, W, H, f3(eye), fr.rays.ptr, fr.flags.ptr, fr.pkeys.ptr,
                      fr.pvals.ptr)
            oracle.call("oracle_ReflectionRayCreationWrapper", W, H, eye, o_rays, o_flags, o["pk"], o["pv"])
        # outputs cleared outside the graph, so a replay that wrote nothing fails
        fr.sckeys.upload(np.full(S, 0xFFFFFFFF, np.uint32))
        fr.hier.upload(np.full(fr.hier.shape, np.nan, np.float32))
        fr.counts.upload(np.full(4, 0xFFFFFFFF, np.uint32))
        assert h.hipGraphLaunch(ge, s) == 0
        assert h.hipStreamSynchronize(s) == 0
        tot, ct = C.c_uint32(0), C.c_uint32(0)
        oracle.call("oracle_RayTrimmingWrapper", o["pk"], o["pv"], W, H, LIGHT_MAX, o_flags, o_scan, o["sk"],
                    o["sv"], C.byref(tot))
        N = tot.value
        oracle.call("oracle_RayCompressionWrapper", o["sk"], o["sv"], N, o_flags, o_scan, o["b"], o["sz"], o["ck"],
                    o["cv"], C.byref(ct))
        Cn = ct.value
        oracle.call("oracle_RaySortingWrapper", o["ck"], o["cv"], Cn, o["sck"], o["scv"])
        oracle.call("oracle_RayDecompressionWrapper", o["b"], o["sz"], o["sck"], o["scv"], o["sk"], o["sv"], o_flags,
                    o_scan, Cn, o["pk"], o["pv"])
        oracle.call("oracle_HierarchyCreationWrapper", o_rays, o["pk"], o["pv"], N, 0.0, 0.0, o_hier)
        n, c = fr.counts.download()[:2]
        assert (int(n), int(c)) == (N, Cn), kind
        np.testing.assert_array_equal(fr.pkeys.download()[:N], o["pk"][:N], err_msg=kind)
        np.testing.assert_array_equal(fr.pvals.download()[:N], o["pv"][:N], err_msg=kind)
        np.testing.assert_array_equal(fr.sckeys.download()[:Cn], o["sck"][:Cn], err_msg=kind)
        nn = (N + SUB - 1) // SUB
        nn += (nn + SUB - 1) // SUB
        got = fr.hier.download()[:2 * nn]
        assert np.array_equal(got.view(np.uint32), o_hier[:2 * nn].view(np.uint32)), kind
    assert h.hipGraphExecDestroy(ge) == 0
    assert h.hipGraphDestroy(g) == 0
    h.hipStreamDestroy(s)
    ctx.close()
