// compat.hip -- the reference's extern "C" stage wrappers on gfx950.
//
// Same names, parameter lists and buffer layouts as
// projects/CUDA-Engine/source/RayTracer.cu:2042-3301.  State that the
// reference kept in texture references and globals (bound G-buffer, scene
// arrays, algorithmID, compile-time constants) lives in one library-level
// context.  Counts still return through host pointers after a blocking copy,
// as the reference's callers expect.  Differences:
//  * scans and the radix sort are the library's own (no CUB);
//  * decompression gathers per output ray instead of copying each chunk
//    serially (RayTracer.cu:1054);
//  * a traversal level that would write past the caller's hit buffers
//    (hierarchyHitMemoryTotal) fails with RSH_CAPACITY instead of
//    overrunning them (the reference has no guard, RayTracer.cu:2788).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/rsh.h"
#include "../../include/rsh_engine.h"
#include "rsh_kernels.hpp"

using namespace rsh;

namespace {

struct Compat {
    hipStream_t stream = nullptr;
    uint32_t light_max = 1, depth = 2, sub = 8;
    int algorithm = 0;
    int exit_on_error = 1;
    int status = 0;
    std::string error;
    const float4 *g_dif = nullptr, *g_spec = nullptr, *g_pos = nullptr, *g_nrm = nullptr;
    const float4 *tri_pos = nullptr, *tri_nrm = nullptr;
    const float2* tri_uv = nullptr;
    const int32_t *tri_obj = nullptr, *tri_mat = nullptr;
    const float4 *mat_dif = nullptr, *mat_spec = nullptr;
    const float4* spheres = nullptr;
    const float4 *light_pos = nullptr, *light_col = nullptr;
    const float2* light_int = nullptr;
    uint32_t T = 0, M = 0, K = 0, L = 0;
    ScanScratch sc;
    DevBuf dtotal;
    uint64_t scratch_items = 0;
};

Compat& C() {
    static Compat c;
    return c;
}

void raise(int code, const std::string& msg) {
    Compat& c = C();
    if (c.status == 0) { c.status = code; c.error = msg; }
    if (c.exit_on_error) {
        std::fprintf(stderr, "[rsh] %s\n", msg.c_str());
        std::exit(1);
    }
}

bool check(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    raise(RSH_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool ensure_scratch(uint64_t n) {
    Compat& c = C();
    if (n <= c.scratch_items && c.dtotal.p) return true;
    if (!check(c.sc.reserve(n), "scratch allocation")) return false;
    if (!check(c.dtotal.reserve(64), "scratch allocation")) return false;
    c.scratch_items = n;
    return true;
}

bool read_u32(const uint32_t* d, uint32_t* h) {
    Compat& c = C();
    if (!check(hipMemcpyAsync(h, d, 4, hipMemcpyDeviceToHost, c.stream), "count readback")) return false;
    return check(hipStreamSynchronize(c.stream), "count readback");
}

KeyCfg ref_keys() {
    KeyCfg k{};
    k.mode = 0;
    return k;
}

v3 V(rsh_float3 a) { return v3{a.x, a.y, a.z}; }

SceneView scene_view(uint32_t lights) {
    Compat& c = C();
    SceneView s{};
    s.tri_pos = c.tri_pos; s.tri_nrm = c.tri_nrm; s.tri_mat = c.tri_mat;
    s.mat_dif = c.mat_dif; s.mat_spec = c.mat_spec;
    s.lights = Lights{c.light_pos, c.light_col, c.light_int, lights};
    return s;
}

bool need(const void* p, const char* what) {
    if (p) return true;
    raise(RSH_INVALID_ARG, std::string(what) + " is not bound");
    return false;
}

void level_counts(uint32_t N, uint32_t* cnt, uint32_t* off) {
    Compat& c = C();
    level_sizes(N, c.sub, c.sub, c.depth, cnt, off);
}

}  // namespace

extern "C" {

void rsh_compat_configure(unsigned int lightMaximum, unsigned int hierarchyDepth, unsigned int hierarchySubdivision,
                          int algorithmID) {
    Compat& c = C();
    if (lightMaximum == 0 || hierarchyDepth == 0 || hierarchyDepth > 4 || hierarchySubdivision == 0 ||
        (algorithmID != 0 && algorithmID != 1)) {
        raise(RSH_INVALID_ARG, "rsh_compat_configure: invalid constants");
        return;
    }
    c.light_max = lightMaximum; c.depth = hierarchyDepth; c.sub = hierarchySubdivision; c.algorithm = algorithmID;
}
void rsh_compat_set_stream(void* s) { C().stream = static_cast<hipStream_t>(s); }
void rsh_compat_set_exit_on_error(int v) { C().exit_on_error = v; }
int rsh_compat_status(void) { return C().status; }
const char* rsh_compat_error(void) { return C().error.c_str(); }
void rsh_compat_reset_status(void) { C().status = 0; C().error.clear(); }

// ------------------------------------------------------------- binders ----
void bindDiffuseTextureArray(const rsh_float4* p) { C().g_dif = reinterpret_cast<const float4*>(p); }
void bindSpecularTextureArray(const rsh_float4* p) { C().g_spec = reinterpret_cast<const float4*>(p); }
void bindFragmentPositionArray(const rsh_float4* p) { C().g_pos = reinterpret_cast<const float4*>(p); }
void bindFragmentNormalArray(const rsh_float4* p) { C().g_nrm = reinterpret_cast<const float4*>(p); }
void bindTrianglePositions(float* p, unsigned int T) { C().tri_pos = reinterpret_cast<const float4*>(p); C().T = T; }
void bindTriangleNormals(float* p, unsigned int T) { C().tri_nrm = reinterpret_cast<const float4*>(p); (void)T; }
void bindTriangleTextureCoordinates(float* p, unsigned int T) { C().tri_uv = reinterpret_cast<const float2*>(p); (void)T; }
void bindTriangleObjectIDs(float* p, unsigned int T) { C().tri_obj = reinterpret_cast<const int32_t*>(p); (void)T; }
void bindTriangleMaterialIDs(float* p, unsigned int T) { C().tri_mat = reinterpret_cast<const int32_t*>(p); (void)T; }
void bindMaterialDiffuseProperties(float* p, unsigned int M) { C().mat_dif = reinterpret_cast<const float4*>(p); C().M = M; }
void bindMaterialSpecularProperties(float* p, unsigned int M) { C().mat_spec = reinterpret_cast<const float4*>(p); (void)M; }
void bindBoundingSpheres(float* p, unsigned int K) { C().spheres = reinterpret_cast<const float4*>(p); C().K = K; }
void bindLightPositions(float* p, unsigned int L) { C().light_pos = reinterpret_cast<const float4*>(p); C().L = L; }
void bindLightColors(float* p, unsigned int L) { C().light_col = reinterpret_cast<const float4*>(p); (void)L; }
void bindLightIntensities(float* p, unsigned int L) { C().light_int = reinterpret_cast<const float2*>(p); (void)L; }

// ------------------------------------------------------------- stages -----
void TriangleUpdateWrapper(float* model, float* normal, unsigned int T, rsh_float4* pos, rsh_float4* nrm) {
    Compat& c = C();
    if (!need(c.tri_pos, "bindTrianglePositions") || !need(c.tri_nrm, "bindTriangleNormals") ||
        !need(c.tri_obj, "bindTriangleObjectIDs"))
        return;
    launch_update_vertex(model, normal, 3 * T, c.tri_pos, c.tri_nrm, c.tri_obj, reinterpret_cast<float4*>(pos),
                         reinterpret_cast<float4*>(nrm), c.stream);
    check(hipGetLastError(), "TriangleUpdateWrapper");
}

void BoundingSphereUpdateWrapper(float* translation, float* scale, const unsigned int K, rsh_float3* out) {
    Compat& c = C();
    if (!need(c.spheres, "bindBoundingSpheres")) return;
    launch_update_sphere(translation, scale, K, c.spheres, reinterpret_cast<float*>(out), c.stream);
    check(hipGetLastError(), "BoundingSphereUpdateWrapper");
}

void MemoryPreparationWrapper(const unsigned int W, const unsigned int H, const unsigned int hitMemory,
                              unsigned int*, unsigned int*, unsigned int*, unsigned int*, unsigned int*,
                              unsigned int*) {
    Compat& c = C();
    uint64_t n = (uint64_t)W * H * c.light_max;
    if (hitMemory > n) n = hitMemory;
    ensure_scratch(n);
}

void ShadowRayCreationWrapper(const unsigned int W, const unsigned int H, const unsigned int L, rsh_float3* rays,
                              unsigned int* flags, unsigned int* keys, unsigned int* values) {
    Compat& c = C();
    if (!need(c.g_pos, "bindFragmentPositionArray") || !need(c.g_nrm, "bindFragmentNormalArray") ||
        !need(c.light_pos, "bindLightPositions"))
        return;
    if (L > c.L) {
        raise(RSH_INVALID_ARG, "ShadowRayCreationWrapper: lightTotal exceeds the bound light count");
        return;
    }
    launch_create_shadow(c.g_pos, c.g_nrm, W, H, L, c.light_pos, ref_keys(), single_view(W, H, v3{0, 0, 0}), 0u,
                         reinterpret_cast<float*>(rays), flags, keys, values, c.stream);
    check(hipGetLastError(), "ShadowRayCreationWrapper");
}

void ReflectionRayCreationWrapper(const unsigned int W, const unsigned int H, const rsh_float3 eye, rsh_float3* rays,
                                  unsigned int* flags, unsigned int* keys, unsigned int* values) {
    Compat& c = C();
    if (!need(c.g_pos, "bindFragmentPositionArray") || !need(c.g_nrm, "bindFragmentNormalArray")) return;
    launch_create_reflection(c.g_pos, c.g_nrm, c.g_dif ? c.g_dif : c.g_pos, W, H, c.light_max,
                             single_view(W, H, V(eye)), 0, 0, ref_keys(), reinterpret_cast<float*>(rays), flags, keys,
                             values, c.stream);
    check(hipGetLastError(), "ReflectionRayCreationWrapper");
}

// No wrapper in the reference: CreateRefractionRays (RayTracer.cu:793-842) is
// never launched there.  Same arguments as ReflectionRayCreationWrapper.
void RefractionRayCreationWrapper(const unsigned int W, const unsigned int H, const rsh_float3 eye, rsh_float3* rays,
                                  unsigned int* flags, unsigned int* keys, unsigned int* values) {
    Compat& c = C();
    if (!need(c.g_pos, "bindFragmentPositionArray") || !need(c.g_nrm, "bindFragmentNormalArray")) return;
    launch_create_refraction(c.g_pos, c.g_nrm, W, H, c.light_max, single_view(W, H, V(eye)), ref_keys(),
                             reinterpret_cast<float*>(rays), flags, keys, values, c.stream);
    check(hipGetLastError(), "RefractionRayCreationWrapper");
}

void RayTrimmingWrapper(unsigned int* keys, unsigned int* values, const unsigned int W, const unsigned int H,
                        const unsigned int L, unsigned int* flags, unsigned int*, unsigned int* tkeys,
                        unsigned int* tvalues, unsigned int* rayTotal) {
    Compat& c = C();
    uint64_t S = (uint64_t)W * H * L;
    if (!ensure_scratch(S)) return;
    if (!check(run_trim(flags, keys, values, S, tkeys, tvalues, c.dtotal.as<uint32_t>(), c.sc, c.stream),
               "RayTrimmingWrapper"))
        return;
    read_u32(c.dtotal.as<uint32_t>(), rayTotal);
}

void RayCompressionWrapper(unsigned int* tkeys, unsigned int*, const unsigned int N, unsigned int*, unsigned int*,
                           unsigned int* bases, unsigned int* sizes, unsigned int* ckeys, unsigned int* cvalues,
                           unsigned int* chunkTotal) {
    Compat& c = C();
    if (!ensure_scratch(N)) return;
    if (!check(run_compress(tkeys, N, bases, sizes, ckeys, cvalues, c.dtotal.as<uint32_t>(), c.sc, c.stream),
               "RayCompressionWrapper"))
        return;
    read_u32(c.dtotal.as<uint32_t>(), chunkTotal);
}

void RaySortingWrapper(unsigned int* ckeys, unsigned int* cvalues, const unsigned int Cn, unsigned int* skeys,
                       unsigned int* svalues) {
    Compat& c = C();
    if (!ensure_scratch(Cn)) return;
    check(run_radix_sort(ckeys, cvalues, Cn, 32, skeys, svalues, c.sc, c.stream), "RaySortingWrapper");
}

// The sorting stage on the low `bits` key bits (stable; 8- or 9-bit digits),
// as the engine runs it for a frame's shadow and reflection passes.
void rsh_sort_pairs(const unsigned int* keys, const unsigned int* values, const unsigned int n,
                    const unsigned int bits, unsigned int* out_keys, unsigned int* out_values) {
    Compat& c = C();
    if (bits == 0 || bits > 32) {
        raise(RSH_INVALID_ARG, "rsh_sort_pairs: bits must be 1..32");
        return;
    }
    if (!ensure_scratch(n)) return;
    if (check(run_radix_sort(keys, values, n, bits, out_keys, out_values, c.sc, c.stream), "rsh_sort_pairs"))
        check(hipStreamSynchronize(c.stream), "rsh_sort_pairs");
}

void RayDecompressionWrapper(unsigned int* bases, unsigned int* sizes, unsigned int*, unsigned int* scvalues,
                             unsigned int* tkeys, unsigned int* tvalues, unsigned int* skeleton, unsigned int* scan,
                             const unsigned int Cn, unsigned int* rkeys, unsigned int* rvalues) {
    Compat& c = C();
    if (!ensure_scratch(Cn)) return;
    check(run_decompress(bases, sizes, scvalues, tkeys, tvalues, skeleton, scan, Cn, 0, rkeys, rvalues, c.sc,
                         c.stream),
          "RayDecompressionWrapper");
}

void HierarchyCreationWrapper(rsh_float3* rays, unsigned int*, unsigned int* svalues, const unsigned int N,
                              const float r0, const float s0, rsh_float4* hierarchy) {
    Compat& c = C();
    if (N == 0) return;
    uint32_t cnt[4], off[4];
    level_counts(N, cnt, off);
    launch_hierarchy(nullptr, nullptr, reinterpret_cast<const float*>(rays), svalues, N, c.sub, c.sub, c.depth, r0, s0,
                     0, reinterpret_cast<float4*>(hierarchy), nullptr, cnt, off, c.stream);
    check(hipGetLastError(), "HierarchyCreationWrapper");
}

void HierarchyTraversalWarmUpWrapper(rsh_float4* hierarchy, rsh_float3* spheres, rsh_float4* tri_pos,
                                     const unsigned int K, const unsigned int N, const unsigned int Tb,
                                     const unsigned int offset, unsigned int* flags, unsigned int* scan,
                                     unsigned int* hits, unsigned int* thits, unsigned int* hitTotal,
                                     unsigned int* hitMemory) {
    Compat& c = C();
    const float4* h = reinterpret_cast<const float4*>(hierarchy);
    const float4* tp = reinterpret_cast<const float4*>(tri_pos);
    uint32_t cnt[4], off[4];
    level_counts(N, cnt, off);
    const uint32_t top = c.depth - 1, nodeOffset = off[top], nodeTotal = cnt[top];
    const uint64_t hitMax = (uint64_t)nodeTotal * Tb;
    if (hitMemory && hitMax > *hitMemory) {
        raise(RSH_CAPACITY, "HierarchyTraversalWarmUpWrapper: node x triangle pairs exceed hierarchyHitMemoryTotal");
        return;
    }
    if (!ensure_scratch(hitMax)) return;
    uint32_t* dt = c.dtotal.as<uint32_t>();
    uint32_t total = 0;
    if (c.algorithm == 0) {
        launch_fill_u32(flags, 0u, hitMax, c.stream);
        launch_bsphere_intersections(h, reinterpret_cast<const float*>(spheres), K, Tb, offset, nodeOffset, nodeTotal,
                                     flags, c.stream);
        if (!check(run_inclusive_scan(flags, scan, hitMax, c.sc, c.stream), "HierarchyTraversalWarmUpWrapper")) return;
        launch_bsphere_hits(Tb, nodeTotal, scan, flags, hits, c.stream);
        if (!check(run_trim(flags, hits, nullptr, hitMax, thits, nullptr, dt, c.sc, c.stream),
                   "HierarchyTraversalWarmUpWrapper"))
            return;
        if (!read_u32(dt, &total)) return;
    } else {
        total = (uint32_t)hitMax;
    }
    *hitTotal = total;
    if (total == 0) return;
    launch_level0_hits(h, tp, Tb, offset, total, nodeOffset, c.algorithm == 1 ? 1u : 0u, flags, hits, thits, c.stream);
    if (!check(run_trim(flags, hits, nullptr, total, thits, nullptr, dt, c.sc, c.stream),
               "HierarchyTraversalWarmUpWrapper"))
        return;
    read_u32(dt, hitTotal);
}

void HierarchyTraversalWrapper(rsh_float4* hierarchy, rsh_float4* tri_pos, const unsigned int N, const unsigned int Tb,
                               const unsigned int offset, unsigned int* flags, unsigned int*, unsigned int* hits,
                               unsigned int* thits, unsigned int* hitTotal, unsigned int* hitMemory) {
    Compat& c = C();
    (void)Tb;
    const float4* h = reinterpret_cast<const float4*>(hierarchy);
    const float4* tp = reinterpret_cast<const float4*>(tri_pos);
    uint32_t cnt[4], off[4];
    level_counts(N, cnt, off);
    uint32_t* dt = c.dtotal.as<uint32_t>();
    for (int level = (int)c.depth - 2; level >= 0; level--) {
        if (*hitTotal == 0) return;
        const uint64_t hitMax = (uint64_t)(*hitTotal) * c.sub;
        if (hitMemory && hitMax > *hitMemory) {
            raise(RSH_CAPACITY, "HierarchyTraversalWrapper: expanded pairs exceed hierarchyHitMemoryTotal");
            return;
        }
        if (!ensure_scratch(hitMax)) return;
        launch_levelN_hits(h, tp, offset, *hitTotal, off[level], cnt[level], c.sub, flags, hits, thits, c.stream);
        if (!check(run_trim(flags, hits, nullptr, hitMax, thits, nullptr, dt, c.sc, c.stream),
                   "HierarchyTraversalWrapper"))
            return;
        if (!read_u32(dt, hitTotal)) return;
    }
}

void ShadowRayPreparationWrapper(const unsigned int W, const unsigned int H, const unsigned int L, unsigned int* f) {
    launch_fill_u32(f, 0u, (uint64_t)W * H * L, C().stream);
    check(hipGetLastError(), "ShadowRayPreparationWrapper");
}

void ShadowRayIntersectionWrapper(rsh_float3* rays, unsigned int*, unsigned int* svalues, unsigned int* hits,
                                  rsh_float4* tri_pos, const unsigned int hitTotal, const unsigned int N,
                                  const unsigned int offset, const unsigned int W, const unsigned int H,
                                  const rsh_float3, unsigned int* sflags) {
    Compat& c = C();
    (void)H;
    if (!need(c.g_pos, "bindFragmentPositionArray")) return;
    launch_hits_intersect(reinterpret_cast<const float*>(rays), svalues, hits, reinterpret_cast<const float4*>(tri_pos),
                          hitTotal, N, offset, W, c.light_max, c.sub, c.g_pos, 1, sflags, c.stream);
    check(hipGetLastError(), "ShadowRayIntersectionWrapper");
}

void ShadowRayColoringWrapper(const unsigned int W, const unsigned int H, const unsigned int L, const rsh_float3 eye,
                              unsigned int* sflags, unsigned int* pixels) {
    Compat& c = C();
    if (!need(c.g_pos, "bindFragmentPositionArray") || !need(c.g_nrm, "bindFragmentNormalArray") ||
        !need(c.g_dif, "bindDiffuseTextureArray") || !need(c.g_spec, "bindSpecularTextureArray") ||
        !need(c.light_pos, "bindLightPositions") || !need(c.light_col, "bindLightColors") ||
        !need(c.light_int, "bindLightIntensities"))
        return;
    if (L > c.L || L > c.light_max) {
        raise(RSH_INVALID_ARG, "ShadowRayColoringWrapper: lightTotal exceeds the bound lights or LIGHT_SOURCE_MAXIMUM");
        return;
    }
    GBufView g{const_cast<float4*>(c.g_pos), const_cast<float4*>(c.g_nrm), const_cast<float4*>(c.g_dif),
               const_cast<float4*>(c.g_spec)};
    launch_shade_primary(g, W, H, c.light_max, Lights{c.light_pos, c.light_col, c.light_int, L},
                         single_view(W, H, V(eye)), sflags, 0u, pixels, c.stream);
    check(hipGetLastError(), "ShadowRayColoringWrapper");
}

void ReflectionRayPreparationWrapper(const unsigned int W, const unsigned int H, unsigned int* times) {
    launch_fill_u32(times, MISS32, (uint64_t)W * H * C().light_max, C().stream);
    check(hipGetLastError(), "ReflectionRayPreparationWrapper");
}

void ReflectionRayIntersectionWrapper(rsh_float3* rays, unsigned int*, unsigned int* svalues, unsigned int* hits,
                                      rsh_float4* tri_pos, const unsigned int hitTotal, const unsigned int N,
                                      const unsigned int offset, const unsigned int W, const unsigned int,
                                      const unsigned int, const rsh_float3, unsigned int* times) {
    Compat& c = C();
    launch_hits_intersect(reinterpret_cast<const float*>(rays), svalues, hits, reinterpret_cast<const float4*>(tri_pos),
                          hitTotal, N, offset, W, c.light_max, c.sub, nullptr, 0, times, c.stream);
    check(hipGetLastError(), "ReflectionRayIntersectionWrapper");
}

void ReflectionRayColoringWrapper(rsh_float4* tri_pos, rsh_float4* tri_nrm, const unsigned int W, const unsigned int H,
                                  const unsigned int L, const rsh_float3 eye, const bool create, unsigned int* times,
                                  rsh_float3* rays, unsigned int* flags, unsigned int* keys, unsigned int* values,
                                  unsigned int* pixels) {
    Compat& c = C();
    if (!need(c.tri_mat, "bindTriangleMaterialIDs") || !need(c.mat_dif, "bindMaterialDiffuseProperties") ||
        !need(c.mat_spec, "bindMaterialSpecularProperties") || !need(c.light_pos, "bindLightPositions") ||
        !need(c.light_col, "bindLightColors") || !need(c.light_int, "bindLightIntensities"))
        return;
    if (L > c.L) {
        raise(RSH_INVALID_ARG, "ReflectionRayColoringWrapper: lightTotal exceeds the bound light count");
        return;
    }
    launch_color_reflection_ref(reinterpret_cast<const float4*>(tri_pos), reinterpret_cast<const float4*>(tri_nrm), W,
                                H, c.light_max, scene_view(L), V(eye), create ? 1u : 0u, times,
                                reinterpret_cast<float*>(rays), flags, keys, values, pixels, c.stream);
    check(hipGetLastError(), "ReflectionRayColoringWrapper");
}

}  // extern "C"
