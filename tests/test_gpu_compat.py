"""Drop-in wrappers (include/rsh_engine.h) against the oracle's restatements.

The test follows the reference's frame driver for CRSH and RAH
(projects/CUDA-Engine/source/main.cpp:630-1054, 1234-1303): shadow rays from a
G-buffer, trim, compress, sort, decompress, hierarchy, per-triangle-batch
warm-up traversal, lower-level traversal, shadow intersection and shading,
then two mirror bounces, then refraction rays (the reference's unlaunched
CreateRefractionRays) traced for closest hits.  After every wrapper the device buffers are compared
with the oracle's host buffers (bit-exact, including the order of the
compacted hit records).
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIGHT_MAX, DEPTH, SUB, TB = 1, 2, 8, 5000


class Twin:
    """A buffer that exists on the device (for the wrappers) and on the host (for the oracle)."""

    def __init__(self, rsh, n, dtype=np.uint32, init=None):
        self.host = np.zeros(n, dtype) if init is None else np.ascontiguousarray(init, dtype).copy()
        self.dev = rsh.DeviceArray.from_host(self.host)

    def d(self):
        return self.dev.download().reshape(self.host.shape)


def f3(v):
    import rsh_amd
    return rsh_amd.Float3(*[float(x) for x in v])


@pytest.fixture(scope="module")
def setup(rsh, oracle):
    scene = rsh.scenes.config0_scene()
    W = H = 96
    cam = rsh.scenes.config0_camera(W, H)
    cfg = rsh.config.config0(W, H).with_scene_bounds(scene)
    ofr = oracle.render(scene, cam, cfg, keep_hits=False)
    return scene, cam, ofr.gbuffer, W, H


@pytest.mark.parametrize("alias", [False, True])
@pytest.mark.parametrize("algorithm", [0, 1])
def test_reference_frame_through_wrappers(rsh, oracle, setup, algorithm, alias):
    """alias: the shadow flags and closest-hit records live in the chunk-key
    buffer, as the reference's frame driver passes cudaPrimaryChunkKeysArrayDP
    to the preparation, intersection and coloring wrappers
    (main.cpp:674-677, 749, 823, 896) between two compressions."""
    L = rsh.lib()
    scene, cam, gb, W, H = setup
    T = scene.triangle_total
    K = scene.object_total
    S = W * H * LIGHT_MAX
    eye = np.asarray(cam.eye, np.float32)
    L.rsh_compat_set_exit_on_error(0)
    L.rsh_compat_reset_status()
    L.rsh_compat_configure(LIGHT_MAX, DEPTH, SUB, algorithm)
    oracle.call("oracle_set_constants", LIGHT_MAX, DEPTH, SUB, algorithm)

    # ---- scene binding (main.cpp:2498-2656) ----
    pos = Twin(rsh, scene.tri_pos.shape, np.float32, scene.tri_pos)
    nrm = Twin(rsh, scene.tri_nrm.shape, np.float32, scene.tri_nrm)
    obj = Twin(rsh, 3 * T, np.int32, np.repeat(scene.triangle_object(), 3))
    mat = Twin(rsh, 3 * T, np.int32, scene.tri_mat)
    mdif = Twin(rsh, scene.mat_diffuse.shape, np.float32, scene.mat_diffuse)
    mspec = Twin(rsh, scene.mat_specular.shape, np.float32, scene.mat_specular)
    sph4 = np.zeros((2 * K, 4), np.float32)
    sph4[0::2, :3] = scene.object_spheres[:, :3]
    sph4[0::2, 3] = 1.0
    sph4[1::2, :3] = scene.object_spheres[:, 3:6]
    sph4[1::2, 3] = 1.0
    sph_orig = Twin(rsh, sph4.shape, np.float32, sph4)
    lpos = Twin(rsh, scene.light_pos.shape, np.float32, scene.light_pos)
    lcol = Twin(rsh, scene.light_col.shape, np.float32, scene.light_col)
    lint = Twin(rsh, scene.light_int.shape, np.float32, scene.light_int)
    gdif = Twin(rsh, gb["dif"].shape, np.float32, gb["dif"])
    gspec = Twin(rsh, gb["spec"].shape, np.float32, gb["spec"])
    gpos = Twin(rsh, gb["pos"].shape, np.float32, gb["pos"])
    gnrm = Twin(rsh, gb["nrm"].shape, np.float32, gb["nrm"])
    L.bindTrianglePositions(pos.dev.ptr, T)
    L.bindTriangleNormals(nrm.dev.ptr, T)
    L.bindTriangleObjectIDs(obj.dev.ptr, T)
    L.bindTriangleMaterialIDs(mat.dev.ptr, T)
    L.bindMaterialDiffuseProperties(mdif.dev.ptr, scene.mat_diffuse.shape[0])
    L.bindMaterialSpecularProperties(mspec.dev.ptr, scene.mat_specular.shape[0])
    L.bindBoundingSpheres(sph_orig.dev.ptr, K)
    L.bindLightPositions(lpos.dev.ptr, scene.light_total)
    L.bindLightColors(lcol.dev.ptr, scene.light_total)
    L.bindLightIntensities(lint.dev.ptr, scene.light_total)
    L.bindDiffuseTextureArray(gdif.dev.ptr)
    L.bindSpecularTextureArray(gspec.dev.ptr)
    L.bindFragmentPositionArray(gpos.dev.ptr)
    L.bindFragmentNormalArray(gnrm.dev.ptr)
    osc = oracle.make_scene(scene)
    oracle.lib().oracle_bind_scene(C.byref(osc))
    oracle.call("oracle_bind_gbuffer", gb["dif"], gb["spec"], gb["pos"], gb["nrm"])
    oracle.call("oracle_bind_original", scene.tri_pos, scene.tri_nrm, obj.host, sph4)

    # ---- per-frame scene update with identity / translation matrices (RT.cu:2046-2091) ----
    eye16 = np.tile(np.eye(4, dtype=np.float32).reshape(-1), (K, 1))
    model = Twin(rsh, eye16.shape, np.float32, eye16)
    upos = Twin(rsh, scene.tri_pos.shape, np.float32)
    unrm = Twin(rsh, scene.tri_nrm.shape, np.float32)
    L.TriangleUpdateWrapper(model.dev.ptr, model.dev.ptr, T, upos.dev.ptr, unrm.dev.ptr)
    oracle.call("oracle_TriangleUpdateWrapper", model.host, model.host, T, upos.host, unrm.host)
    np.testing.assert_array_equal(upos.d(), upos.host)
    np.testing.assert_array_equal(unrm.d(), unrm.host)
    usph = Twin(rsh, (2 * K, 3), np.float32)
    L.BoundingSphereUpdateWrapper(model.dev.ptr, model.dev.ptr, K, usph.dev.ptr)
    oracle.call("oracle_BoundingSphereUpdateWrapper", model.host, model.host, K, usph.host)
    np.testing.assert_array_equal(usph.d(), usph.host)
    tri = upos

    # ---- frame buffers (main.cpp:1400-1497) ----
    n0 = (S + SUB - 1) // SUB
    hit_mem = (n0 // 2) * 10000
    rays = Twin(rsh, (2 * S, 3), np.float32)
    pkeys, pvals, skeys, svals = (Twin(rsh, S) for _ in range(4))
    bases, sizes, ckeys, cvals, sckeys, scvals = (Twin(rsh, S) for _ in range(6))
    flags, scan = Twin(rsh, hit_mem), Twin(rsh, hit_mem)
    hits, thits = Twin(rsh, hit_mem), Twin(rsh, hit_mem)
    nodes_total = n0 + (n0 + SUB - 1) // SUB
    hier = Twin(rsh, (2 * nodes_total, 4), np.float32)
    sflags = ckeys if alias else Twin(rsh, S)
    pixels = Twin(rsh, W * H)
    L.MemoryPreparationWrapper(W, H, hit_mem, flags.dev.ptr, scan.dev.ptr, ckeys.dev.ptr, cvals.dev.ptr,
                               sckeys.dev.ptr, scvals.dev.ptr)

    def eq(tw, n=None, what=""):
        got = tw.d()
        if n is not None:
            got, want = got.reshape(-1)[:n], tw.host.reshape(-1)[:n]
        else:
            want = tw.host
        np.testing.assert_array_equal(got, want, err_msg=what)

    def cast_rays(rk, rv):
        """main.cpp:951-1054: trim, (compress, sort, decompress), hierarchy."""
        tot, otot = C.c_uint32(0), C.c_uint32(0)
        if algorithm == 0:
            L.RayTrimmingWrapper(pkeys.dev.ptr, pvals.dev.ptr, W, H, LIGHT_MAX, flags.dev.ptr, scan.dev.ptr,
                                 skeys.dev.ptr, svals.dev.ptr, C.byref(tot))
            oracle.call("oracle_RayTrimmingWrapper", pkeys.host, pvals.host, W, H, LIGHT_MAX, flags.host, scan.host,
                        skeys.host, svals.host, C.byref(otot))
            assert tot.value == otot.value
            N = tot.value
            if N == 0:
                return 0
            eq(skeys, N, "trimmed keys"); eq(svals, N, "trimmed values")
            ct, oct_ = C.c_uint32(0), C.c_uint32(0)
            L.RayCompressionWrapper(skeys.dev.ptr, svals.dev.ptr, N, flags.dev.ptr, scan.dev.ptr, bases.dev.ptr,
                                    sizes.dev.ptr, ckeys.dev.ptr, cvals.dev.ptr, C.byref(ct))
            oracle.call("oracle_RayCompressionWrapper", skeys.host, svals.host, N, flags.host, scan.host, bases.host,
                        sizes.host, ckeys.host, cvals.host, C.byref(oct_))
            assert ct.value == oct_.value
            Cn = ct.value
            for tw, nm in ((bases, "bases"), (sizes, "sizes"), (ckeys, "chunk keys"), (cvals, "chunk values")):
                eq(tw, Cn, nm)
            L.RaySortingWrapper(ckeys.dev.ptr, cvals.dev.ptr, Cn, sckeys.dev.ptr, scvals.dev.ptr)
            oracle.call("oracle_RaySortingWrapper", ckeys.host, cvals.host, Cn, sckeys.host, scvals.host)
            eq(sckeys, Cn, "sorted chunk keys"); eq(scvals, Cn, "sorted chunk values")
            L.RayDecompressionWrapper(bases.dev.ptr, sizes.dev.ptr, sckeys.dev.ptr, scvals.dev.ptr, skeys.dev.ptr,
                                      svals.dev.ptr, flags.dev.ptr, scan.dev.ptr, Cn, pkeys.dev.ptr, pvals.dev.ptr)
            oracle.call("oracle_RayDecompressionWrapper", bases.host, sizes.host, sckeys.host, scvals.host,
                        skeys.host, svals.host, flags.host, scan.host, Cn, pkeys.host, pvals.host)
            eq(pkeys, N, "sorted keys"); eq(pvals, N, "sorted values")
            hk, hv = pkeys, pvals
        else:
            L.RayTrimmingWrapper(skeys.dev.ptr, svals.dev.ptr, W, H, LIGHT_MAX, flags.dev.ptr, scan.dev.ptr,
                                 pkeys.dev.ptr, pvals.dev.ptr, C.byref(tot))
            oracle.call("oracle_RayTrimmingWrapper", skeys.host, svals.host, W, H, LIGHT_MAX, flags.host, scan.host,
                        pkeys.host, pvals.host, C.byref(otot))
            assert tot.value == otot.value
            N = tot.value
            if N == 0:
                return 0
            eq(pkeys, N, "trimmed keys"); eq(pvals, N, "trimmed values")
            hk, hv = pkeys, pvals
        L.HierarchyCreationWrapper(rays.dev.ptr, hk.dev.ptr, hv.dev.ptr, N, 0.0, 0.0, hier.dev.ptr)
        oracle.call("oracle_HierarchyCreationWrapper", rays.host, hk.host, hv.host, N, 0.0, 0.0, hier.host)
        nn = (N + SUB - 1) // SUB
        nn += (nn + SUB - 1) // SUB
        got = hier.d()[:2 * nn]
        want = hier.host[:2 * nn]
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "hierarchy nodes"
        return N

    def traverse(N, shadow):
        """main.cpp:685-755 / 831-900: per triangle batch, warm-up + traversal + intersection."""
        batches = (T + TB - 1) // TB
        for i in range(batches):
            off = i * TB
            tb = TB - max(TB * (i + 1) - T, 0)
            ht, oht = C.c_uint32(0), C.c_uint32(0)
            mem = C.c_uint32(hit_mem)
            L.HierarchyTraversalWarmUpWrapper(hier.dev.ptr, usph.dev.ptr, tri.dev.ptr, K, N, tb, off, flags.dev.ptr,
                                              scan.dev.ptr, hits.dev.ptr, thits.dev.ptr, C.byref(ht), C.byref(mem))
            oracle.call("oracle_HierarchyTraversalWarmUpWrapper", hier.host, usph.host, tri.host, K, N, tb, off,
                        flags.host, scan.host, hits.host, thits.host, C.byref(oht))
            assert ht.value == oht.value, ("warm-up hit total", ht.value, oht.value)
            eq(thits, ht.value, "warm-up hits")
            L.HierarchyTraversalWrapper(hier.dev.ptr, tri.dev.ptr, N, tb, off, flags.dev.ptr, scan.dev.ptr,
                                        hits.dev.ptr, thits.dev.ptr, C.byref(ht), C.byref(mem))
            oracle.call("oracle_HierarchyTraversalWrapper", hier.host, tri.host, N, tb, off, flags.host, scan.host,
                        hits.host, thits.host, C.byref(oht))
            assert ht.value == oht.value, ("leaf hit total", ht.value, oht.value)
            eq(thits, ht.value, "leaf hits")
            if shadow:
                L.ShadowRayIntersectionWrapper(rays.dev.ptr, pkeys.dev.ptr, pvals.dev.ptr, thits.dev.ptr,
                                               tri.dev.ptr, ht.value, N, off, W, H, f3(eye), sflags.dev.ptr)
                oracle.call("oracle_ShadowRayIntersectionWrapper", rays.host, pkeys.host, pvals.host, thits.host,
                            tri.host, ht.value, N, off, W, H, sflags.host)
            else:
                L.ReflectionRayIntersectionWrapper(rays.dev.ptr, pkeys.dev.ptr, pvals.dev.ptr, thits.dev.ptr,
                                                   tri.dev.ptr, ht.value, N, off, W, H, 1, f3(eye), sflags.dev.ptr)
                oracle.call("oracle_ReflectionRayIntersectionWrapper", rays.host, pkeys.host, pvals.host,
                            thits.host, tri.host, ht.value, N, off, W, H, sflags.host)

    # ---- shadow iteration (main.cpp:1239-1259) ----
    (pk, pv) = (pkeys, pvals) if algorithm == 0 else (skeys, svals)
    L.ShadowRayCreationWrapper(W, H, LIGHT_MAX, rays.dev.ptr, flags.dev.ptr, pk.dev.ptr, pv.dev.ptr)
    oracle.call("oracle_ShadowRayCreationWrapper", W, H, LIGHT_MAX, rays.host, flags.host, pk.host, pv.host)
    live = flags.host[:S] == 0
    np.testing.assert_array_equal(flags.d()[:S], flags.host[:S])
    np.testing.assert_array_equal(pk.d()[live], pk.host[live])
    np.testing.assert_array_equal(rays.d().reshape(S, 2, 3)[live], rays.host.reshape(S, 2, 3)[live])
    N = cast_rays(pk, pv)
    assert N > 0
    L.ShadowRayPreparationWrapper(W, H, LIGHT_MAX, sflags.dev.ptr)
    oracle.call("oracle_ShadowRayPreparationWrapper", W, H, LIGHT_MAX, sflags.host)
    traverse(N, True)
    eq(sflags, None, "shadow flags")
    L.ShadowRayColoringWrapper(W, H, scene.light_total, f3(eye), sflags.dev.ptr, pixels.dev.ptr)
    oracle.call("oracle_ShadowRayColoringWrapper", W, H, scene.light_total, eye, sflags.host, pixels.host)
    eq(pixels, None, "shadow-ray shading")

    # ---- reflection iterations (main.cpp:1262-1302) ----
    L.ReflectionRayCreationWrapper(W, H, f3(eye), rays.dev.ptr, flags.dev.ptr, pk.dev.ptr, pv.dev.ptr)
    oracle.call("oracle_ReflectionRayCreationWrapper", W, H, eye, rays.host, flags.host, pk.host, pv.host)
    for it, create in ((1, True), (2, False)):
        N = cast_rays(pk, pv)
        if N == 0:
            break
        L.ReflectionRayPreparationWrapper(W, H, sflags.dev.ptr)
        oracle.call("oracle_ReflectionRayPreparationWrapper", W, H, sflags.host)
        traverse(N, False)
        eq(sflags, None, f"closest-hit records, bounce {it}")
        L.ReflectionRayColoringWrapper(tri.dev.ptr, unrm.dev.ptr, W, H, scene.light_total, f3(eye), bool(create),
                                       sflags.dev.ptr, rays.dev.ptr, flags.dev.ptr, pk.dev.ptr, pv.dev.ptr,
                                       pixels.dev.ptr)
        oracle.call("oracle_ReflectionRayColoringWrapper", tri.host, unrm.host, W, H, scene.light_total, eye,
                    create, sflags.host, rays.host, flags.host, pk.host, pv.host, pixels.host)
        eq(pixels, None, f"reflection shading, bounce {it}")

    # ---- refraction rays (RT.cu:793-842, never launched by the reference), traced for closest hits ----
    L.RefractionRayCreationWrapper(W, H, f3(eye), rays.dev.ptr, flags.dev.ptr, pk.dev.ptr, pv.dev.ptr)
    oracle.call("oracle_RefractionRayCreationWrapper", W, H, eye, rays.host, flags.host, pk.host, pv.host)
    live = flags.host[:S] == 0
    np.testing.assert_array_equal(flags.d()[:S], flags.host[:S])
    np.testing.assert_array_equal(pk.d()[live], pk.host[live])
    np.testing.assert_array_equal(rays.d().reshape(S, 2, 3)[live], rays.host.reshape(S, 2, 3)[live])
    N = cast_rays(pk, pv)
    assert N == int(live.sum()) > 0
    L.ReflectionRayPreparationWrapper(W, H, sflags.dev.ptr)
    oracle.call("oracle_ReflectionRayPreparationWrapper", W, H, sflags.host)
    traverse(N, False)
    eq(sflags, None, "refraction closest-hit records")
    assert L.rsh_compat_status() == 0, L.rsh_compat_error()
    print("compat frame ok, algorithm", algorithm, file=sys.stderr)


def test_sort_large_properties(rsh):
    """Full-size sort (2M chunks, the 1920x1080 ray count): output sorted,
    stable and a permutation of the input."""
    L = rsh.lib()
    n = 1920 * 1080
    rng = np.random.default_rng(99)
    keys = rng.integers(0, 1 << 29, n).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = rsh.DeviceArray.from_host(keys), rsh.DeviceArray.from_host(vals)
    ok_, ov = rsh.DeviceArray((n,), np.uint32), rsh.DeviceArray((n,), np.uint32)
    L.MemoryPreparationWrapper(1920, 1080, 0, None, None, None, None, None, None)
    L.RaySortingWrapper(dk.ptr, dv.ptr, n, ok_.ptr, ov.ptr)
    sk, sv = ok_.download(), ov.download()
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(sk, keys[order])
    np.testing.assert_array_equal(sv, vals[order])


@pytest.mark.parametrize("n,bits", [(1920 * 1080, 27), (1920 * 1080, 24), (300_000, 27), (5000, 17), (77, 9),
                                    (100_000, 32)])
def test_sort_pairs_bits(rsh, n, bits):
    """The sort on the low `bits` bits (9-bit digits for 17-18 and 25-27 bits):
    stable in those bits, the bits above carried along unchanged (a frame's
    pass keys share their type bits above the sorted ones)."""
    L = rsh.lib()
    rng = np.random.default_rng(bits * 7 + n % 97)
    low = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    top = (np.uint64(2) << np.uint64(bits)) if bits <= 29 else np.uint64(0)
    keys = ((low | top) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = rsh.DeviceArray.from_host(keys), rsh.DeviceArray.from_host(vals)
    ok_, ov = rsh.DeviceArray((n,), np.uint32), rsh.DeviceArray((n,), np.uint32)
    L.rsh_sort_pairs(dk.ptr, dv.ptr, n, bits, ok_.ptr, ov.ptr)
    assert L.rsh_compat_status() == 0, L.rsh_compat_error()
    order = np.argsort(keys & np.uint32((1 << bits) - 1 if bits < 32 else 0xFFFFFFFF), kind="stable")
    np.testing.assert_array_equal(ok_.download(), keys[order])
    np.testing.assert_array_equal(ov.download(), vals[order])


def test_capacity_guard(rsh, setup):
    """A traversal that would overrun hierarchyHitMemoryTotal reports RSH_CAPACITY."""
    L = rsh.lib()
    L.rsh_compat_set_exit_on_error(0)
    L.rsh_compat_reset_status()
    L.rsh_compat_configure(1, 2, 8, 0)
    dummy = rsh.DeviceArray((1024,), np.uint32)
    ht, mem = C.c_uint32(0), C.c_uint32(10)
    L.HierarchyTraversalWarmUpWrapper(dummy.ptr, dummy.ptr, dummy.ptr, 1, 4096, 5000, 0, dummy.ptr, dummy.ptr,
                                      dummy.ptr, dummy.ptr, C.byref(ht), C.byref(mem))
    assert L.rsh_compat_status() == 2
    L.rsh_compat_reset_status()


@pytest.mark.parametrize("bits", [9, 16, 24, 27, 32])
def test_sort_pairs_in_place(rsh, bits):
    """rsh_sort_pairs with the outputs aliasing the inputs, as the reference's
    buffer reuse allows: odd pass counts (9, 24 and 27 bits) start from a copy
    instead of scattering into the array other blocks still read."""
    L = rsh.lib()
    L.rsh_compat_set_exit_on_error(0)
    L.rsh_compat_reset_status()
    n = 300_001
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = rsh.DeviceArray.from_host(keys), rsh.DeviceArray.from_host(vals)
    L.rsh_sort_pairs(dk.ptr, dv.ptr, n, bits, dk.ptr, dv.ptr)
    assert L.rsh_compat_status() == 0, L.rsh_compat_error()
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(dk.download(), keys[order])
    np.testing.assert_array_equal(dv.download(), vals[order])
