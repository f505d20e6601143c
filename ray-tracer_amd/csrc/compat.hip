// compat.hip -- the reference's extern "C" stage wrappers on gfx950, as a
// per-context, per-stream API (include/rsh_ctx.h) and as the drop-in
// wrappers (include/rsh_engine.h) over one process-wide context.
//
// Same names, parameter lists and buffer layouts as
// projects/CUDA-Engine/source/RayTracer.cu:2042-3301.  State that the
// reference kept in texture references and globals (bound G-buffer, scene
// arrays, algorithmID, compile-time constants, CUB temporary storage) lives in
// an rsh_ctx.  The drop-in wrappers return counts through host pointers after
// a blocking copy, as the reference's callers expect; the context API also
// writes them to device words and takes device counts, with no
// synchronisation.  Differences:
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
#include "../../include/rsh_ctx.h"
#include "../../include/rsh_engine.h"
#include "rsh_kernels.hpp"

using namespace rsh;

struct rsh_ctx {
    int device = 0;
    uint32_t light_max = 1, depth = 2, sub = 8;
    int algorithm = 0;
    std::string err;
    const float4 *g_dif = nullptr, *g_spec = nullptr, *g_pos = nullptr, *g_nrm = nullptr;
    const float4 *tri_pos = nullptr, *tri_nrm = nullptr;
    const int32_t *tri_obj = nullptr, *tri_mat = nullptr;
    const float4 *mat_dif = nullptr, *mat_spec = nullptr;
    const float4* spheres = nullptr;
    const float4 *light_pos = nullptr, *light_col = nullptr;
    const float2* light_int = nullptr;
    uint32_t T = 0, M = 0, K = 0, L = 0;
    ScanScratch sc;
    DevBuf dtotal;       // uint32[16]: counts of the stages given no device word
    DevBuf psz;          // PassSizes of a hierarchy over a device ray count
    uint64_t scratch_items = 0;
    void release() { sc.release(); dtotal.release(); psz.release(); scratch_items = 0; }
};

namespace {

rsh_status fail(rsh_ctx& c, rsh_status code, const std::string& msg) {
    c.err = msg;
    return code;
}

#define CTX_HIP(expr, what)                                                                       \
    do {                                                                                          \
        const hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(c, RSH_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define CTX_OK(expr)                                                                              \
    do {                                                                                          \
        const rsh_status r_ = (expr);                                                             \
        if (r_ != RSH_OK) return r_;                                                              \
    } while (0)

rsh_status need(rsh_ctx& c, const void* p, const char* what) {
    return p ? RSH_OK : fail(c, RSH_INVALID_ARG, std::string(what) + " is not bound");
}

rsh_status ensure_scratch(rsh_ctx& c, uint64_t n) {
    if (n <= c.scratch_items && c.dtotal.p && c.psz.p) return RSH_OK;
    CTX_HIP(c.sc.reserve(n), "scratch allocation");
    CTX_HIP(c.dtotal.reserve(64), "scratch allocation");
    CTX_HIP(c.psz.reserve(sizeof(PassSizes)), "scratch allocation");
    if (n > c.scratch_items) c.scratch_items = n;
    return RSH_OK;
}

// The count in device word d to *h (host) when h is given: the stream is
// waited for, as the reference's blocking cudaMemcpy.
rsh_status read_count(rsh_ctx& c, hipStream_t s, const uint32_t* d, uint32_t* h, const char* what) {
    if (!h) return RSH_OK;
    uint32_t v = 0;
    CTX_HIP(hipMemcpyAsync(&v, d, 4, hipMemcpyDeviceToHost, s), what);
    CTX_HIP(hipStreamSynchronize(s), what);
    *h = v;
    return RSH_OK;
}

KeyCfg ref_keys() {
    KeyCfg k{};
    k.mode = 0;
    return k;
}

v3 V(rsh_float3 a) { return v3{a.x, a.y, a.z}; }

SceneView scene_view(const rsh_ctx& c, uint32_t lights) {
    SceneView s{};
    s.tri_pos = c.tri_pos; s.tri_nrm = c.tri_nrm; s.tri_mat = c.tri_mat;
    s.mat_dif = c.mat_dif; s.mat_spec = c.mat_spec;
    s.lights = Lights{c.light_pos, c.light_col, c.light_int, lights};
    return s;
}

void level_counts(const rsh_ctx& c, uint32_t N, uint32_t* cnt, uint32_t* off) {
    level_sizes(N, c.sub, c.sub, c.depth, cnt, off);
}

hipStream_t S(void* stream) { return static_cast<hipStream_t>(stream); }

// ------------------------------------------------------------- stages -----
// Each stage on context c and stream s; the public functions below add the
// device switch, the drop-in wrappers the process-wide context.

rsh_status configure(rsh_ctx& c, unsigned lightMaximum, unsigned depth, unsigned sub, int algorithm) {
    if (lightMaximum == 0 || depth == 0 || depth > 4 || sub == 0 || (algorithm != 0 && algorithm != 1))
        return fail(c, RSH_INVALID_ARG, "configure: invalid constants");
    c.light_max = lightMaximum; c.depth = depth; c.sub = sub; c.algorithm = algorithm;
    return RSH_OK;
}

rsh_status triangle_update(rsh_ctx& c, hipStream_t s, float* model, float* normal, unsigned T, rsh_float4* pos,
                           rsh_float4* nrm) {
    CTX_OK(need(c, c.tri_pos, "bindTrianglePositions"));
    CTX_OK(need(c, c.tri_nrm, "bindTriangleNormals"));
    CTX_OK(need(c, c.tri_obj, "bindTriangleObjectIDs"));
    launch_update_vertex(model, normal, 3 * T, c.tri_pos, c.tri_nrm, c.tri_obj, reinterpret_cast<float4*>(pos),
                         reinterpret_cast<float4*>(nrm), s);
    CTX_HIP(hipGetLastError(), "TriangleUpdateWrapper");
    return RSH_OK;
}

rsh_status sphere_update(rsh_ctx& c, hipStream_t s, float* translation, float* scale, unsigned K, rsh_float3* out) {
    CTX_OK(need(c, c.spheres, "bindBoundingSpheres"));
    launch_update_sphere(translation, scale, K, c.spheres, reinterpret_cast<float*>(out), s);
    CTX_HIP(hipGetLastError(), "BoundingSphereUpdateWrapper");
    return RSH_OK;
}

rsh_status memory_preparation(rsh_ctx& c, unsigned W, unsigned H, unsigned hitMemory) {
    uint64_t n = (uint64_t)W * H * c.light_max;
    if (hitMemory > n) n = hitMemory;
    return ensure_scratch(c, n);
}

rsh_status shadow_creation(rsh_ctx& c, hipStream_t s, unsigned W, unsigned H, unsigned L, rsh_float3* rays,
                           unsigned* flags, unsigned* keys, unsigned* values) {
    CTX_OK(need(c, c.g_pos, "bindFragmentPositionArray"));
    CTX_OK(need(c, c.g_nrm, "bindFragmentNormalArray"));
    CTX_OK(need(c, c.light_pos, "bindLightPositions"));
    if (L > c.L) return fail(c, RSH_INVALID_ARG, "ShadowRayCreationWrapper: lightTotal exceeds the bound light count");
    launch_create_shadow(c.g_pos, c.g_nrm, W, H, L, c.light_pos, ref_keys(), single_view(W, H, v3{0, 0, 0}), 0u,
                         reinterpret_cast<float*>(rays), flags, keys, values, s);
    CTX_HIP(hipGetLastError(), "ShadowRayCreationWrapper");
    return RSH_OK;
}

rsh_status reflection_creation(rsh_ctx& c, hipStream_t s, unsigned W, unsigned H, rsh_float3 eye, rsh_float3* rays,
                               unsigned* flags, unsigned* keys, unsigned* values) {
    CTX_OK(need(c, c.g_pos, "bindFragmentPositionArray"));
    CTX_OK(need(c, c.g_nrm, "bindFragmentNormalArray"));
    launch_create_reflection(c.g_pos, c.g_nrm, c.g_dif ? c.g_dif : c.g_pos, W, H, c.light_max,
                             single_view(W, H, V(eye)), 0, 0, ref_keys(), reinterpret_cast<float*>(rays), flags, keys,
                             values, s);
    CTX_HIP(hipGetLastError(), "ReflectionRayCreationWrapper");
    return RSH_OK;
}

// No wrapper in the reference: CreateRefractionRays (RayTracer.cu:793-842) is
// never launched there.  Same arguments as the reflection ray creation.
rsh_status refraction_creation(rsh_ctx& c, hipStream_t s, unsigned W, unsigned H, rsh_float3 eye, rsh_float3* rays,
                               unsigned* flags, unsigned* keys, unsigned* values) {
    CTX_OK(need(c, c.g_pos, "bindFragmentPositionArray"));
    CTX_OK(need(c, c.g_nrm, "bindFragmentNormalArray"));
    launch_create_refraction(c.g_pos, c.g_nrm, W, H, c.light_max, single_view(W, H, V(eye)), ref_keys(),
                             reinterpret_cast<float*>(rays), flags, keys, values, s);
    CTX_HIP(hipGetLastError(), "RefractionRayCreationWrapper");
    return RSH_OK;
}

rsh_status trimming(rsh_ctx& c, hipStream_t s, unsigned* keys, unsigned* values, unsigned W, unsigned H, unsigned L,
                    unsigned* flags, unsigned* tkeys, unsigned* tvalues, unsigned* rayTotal, unsigned* rayTotalDev) {
    const uint64_t n = (uint64_t)W * H * L;
    CTX_OK(ensure_scratch(c, n));
    uint32_t* dt = rayTotalDev ? rayTotalDev : c.dtotal.as<uint32_t>();
    CTX_HIP(run_trim(flags, keys, values, n, tkeys, tvalues, dt, c.sc, s), "RayTrimmingWrapper");
    return read_count(c, s, dt, rayTotal, "RayTrimmingWrapper count");
}

rsh_status compression(rsh_ctx& c, hipStream_t s, unsigned* tkeys, unsigned N, const unsigned* N_dev,
                       unsigned* bases, unsigned* sizes, unsigned* ckeys, unsigned* cvalues, unsigned* chunkTotal,
                       unsigned* chunkTotalDev) {
    CTX_OK(ensure_scratch(c, N));
    uint32_t* dt = chunkTotalDev ? chunkTotalDev : c.dtotal.as<uint32_t>() + 1;
    CTX_HIP(run_compress(tkeys, N, bases, sizes, ckeys, cvalues, dt, c.sc, s, N_dev), "RayCompressionWrapper");
    return read_count(c, s, dt, chunkTotal, "RayCompressionWrapper count");
}

rsh_status sorting(rsh_ctx& c, hipStream_t s, const unsigned* keys, const unsigned* values, unsigned n,
                   const unsigned* n_dev, unsigned bits, unsigned* okeys, unsigned* ovalues, const char* what) {
    if (bits == 0 || bits > 32) return fail(c, RSH_INVALID_ARG, std::string(what) + ": bits must be 1..32");
    CTX_OK(ensure_scratch(c, n));
    if (n_dev) {
        if (keys == okeys || values == ovalues || keys == ovalues || values == okeys)
            return fail(c, RSH_INVALID_ARG, std::string(what) + ": outputs alias inputs with a device count");
        CTX_HIP(run_radix_sort_dev(keys, values, n, n_dev, bits, okeys, ovalues, c.sc, s), what);
    } else {
        CTX_HIP(run_radix_sort(keys, values, n, bits, okeys, ovalues, c.sc, s), what);
    }
    return RSH_OK;
}

rsh_status decompression(rsh_ctx& c, hipStream_t s, unsigned* bases, unsigned* sizes, unsigned* scvalues,
                         unsigned* tkeys, unsigned* tvalues, unsigned* skeleton, unsigned* scan, unsigned Cn,
                         const unsigned* Cn_dev, unsigned* rkeys, unsigned* rvalues) {
    CTX_OK(ensure_scratch(c, Cn));
    CTX_HIP(run_decompress(bases, sizes, scvalues, tkeys, tvalues, skeleton, scan, Cn, 0, rkeys, rvalues, c.sc, s,
                           Cn_dev),
            "RayDecompressionWrapper");
    return RSH_OK;
}

rsh_status hierarchy_creation(rsh_ctx& c, hipStream_t s, rsh_float3* rays, unsigned* svalues, unsigned N,
                              const unsigned* N_dev, float r0, float s0, rsh_float4* hierarchy) {
    if (N == 0) return RSH_OK;
    uint32_t cnt[4], off[4];
    level_counts(c, N, cnt, off);
    PassSizes* ps = nullptr;
    if (N_dev) {
        CTX_OK(ensure_scratch(c, 0));
        ps = c.psz.as<PassSizes>();
        launch_ray_sizes(N_dev, N, c.sub, c.sub, c.depth, ps, s);
    }
    launch_hierarchy(nullptr, nullptr, ps, reinterpret_cast<const float*>(rays), svalues, N, c.sub, c.sub, c.depth, r0,
                     s0, 0, reinterpret_cast<float4*>(hierarchy), nullptr, cnt, off, s);
    CTX_HIP(hipGetLastError(), "HierarchyCreationWrapper");
    return RSH_OK;
}

rsh_status traversal_warmup(rsh_ctx& c, hipStream_t s, rsh_float4* hierarchy, rsh_float3* spheres,
                            rsh_float4* tri_pos, unsigned K, unsigned N, unsigned Tb, unsigned offset, unsigned* flags,
                            unsigned* scan, unsigned* hits, unsigned* thits, unsigned* hitTotal, unsigned* hitMemory) {
    if (!hitTotal) return fail(c, RSH_INVALID_ARG, "HierarchyTraversalWarmUpWrapper: hierarchyHitTotal is NULL");
    const float4* h = reinterpret_cast<const float4*>(hierarchy);
    const float4* tp = reinterpret_cast<const float4*>(tri_pos);
    uint32_t cnt[4], off[4];
    level_counts(c, N, cnt, off);
    const uint32_t top = c.depth - 1, nodeOffset = off[top], nodeTotal = cnt[top];
    const uint64_t hitMax = (uint64_t)nodeTotal * Tb;
    if (hitMemory && hitMax > *hitMemory)
        return fail(c, RSH_CAPACITY,
                    "HierarchyTraversalWarmUpWrapper: node x triangle pairs exceed hierarchyHitMemoryTotal");
    CTX_OK(ensure_scratch(c, hitMax));
    uint32_t* dt = c.dtotal.as<uint32_t>() + 2;
    uint32_t total = 0;
    if (c.algorithm == 0) {
        launch_fill_u32(flags, 0u, hitMax, s);
        launch_bsphere_intersections(h, reinterpret_cast<const float*>(spheres), K, Tb, offset, nodeOffset, nodeTotal,
                                     flags, s);
        CTX_HIP(run_inclusive_scan(flags, scan, hitMax, c.sc, s), "HierarchyTraversalWarmUpWrapper");
        launch_bsphere_hits(Tb, nodeTotal, scan, flags, hits, s);
        CTX_HIP(run_trim(flags, hits, nullptr, hitMax, thits, nullptr, dt, c.sc, s), "HierarchyTraversalWarmUpWrapper");
        CTX_OK(read_count(c, s, dt, &total, "HierarchyTraversalWarmUpWrapper count"));
    } else {
        total = (uint32_t)hitMax;
    }
    *hitTotal = total;
    if (total == 0) return RSH_OK;
    launch_level0_hits(h, tp, Tb, offset, total, nodeOffset, c.algorithm == 1 ? 1u : 0u, flags, hits, thits, s);
    CTX_HIP(run_trim(flags, hits, nullptr, total, thits, nullptr, dt, c.sc, s), "HierarchyTraversalWarmUpWrapper");
    return read_count(c, s, dt, hitTotal, "HierarchyTraversalWarmUpWrapper count");
}

rsh_status traversal(rsh_ctx& c, hipStream_t s, rsh_float4* hierarchy, rsh_float4* tri_pos, unsigned N,
                     unsigned offset, unsigned* flags, unsigned* hits, unsigned* thits, unsigned* hitTotal,
                     unsigned* hitMemory) {
    if (!hitTotal) return fail(c, RSH_INVALID_ARG, "HierarchyTraversalWrapper: hierarchyHitTotal is NULL");
    const float4* h = reinterpret_cast<const float4*>(hierarchy);
    const float4* tp = reinterpret_cast<const float4*>(tri_pos);
    uint32_t cnt[4], off[4];
    level_counts(c, N, cnt, off);
    CTX_OK(ensure_scratch(c, 0));
    uint32_t* dt = c.dtotal.as<uint32_t>() + 2;
    for (int level = (int)c.depth - 2; level >= 0; level--) {
        if (*hitTotal == 0) return RSH_OK;
        const uint64_t hitMax = (uint64_t)(*hitTotal) * c.sub;
        if (hitMemory && hitMax > *hitMemory)
            return fail(c, RSH_CAPACITY, "HierarchyTraversalWrapper: expanded pairs exceed hierarchyHitMemoryTotal");
        CTX_OK(ensure_scratch(c, hitMax));
        launch_levelN_hits(h, tp, offset, *hitTotal, off[level], cnt[level], c.sub, flags, hits, thits, s);
        CTX_HIP(run_trim(flags, hits, nullptr, hitMax, thits, nullptr, dt, c.sc, s), "HierarchyTraversalWrapper");
        CTX_OK(read_count(c, s, dt, hitTotal, "HierarchyTraversalWrapper count"));
    }
    return RSH_OK;
}

rsh_status shadow_preparation(rsh_ctx& c, hipStream_t s, unsigned W, unsigned H, unsigned L, unsigned* f) {
    launch_fill_u32(f, 0u, (uint64_t)W * H * L, s);
    CTX_HIP(hipGetLastError(), "ShadowRayPreparationWrapper");
    return RSH_OK;
}

rsh_status shadow_intersection(rsh_ctx& c, hipStream_t s, rsh_float3* rays, unsigned* svalues, unsigned* hits,
                               rsh_float4* tri_pos, unsigned hitTotal, unsigned N, unsigned offset, unsigned W,
                               unsigned* sflags) {
    CTX_OK(need(c, c.g_pos, "bindFragmentPositionArray"));
    launch_hits_intersect(reinterpret_cast<const float*>(rays), svalues, hits, reinterpret_cast<const float4*>(tri_pos),
                          hitTotal, N, offset, W, c.light_max, c.sub, c.g_pos, 1, sflags, s);
    CTX_HIP(hipGetLastError(), "ShadowRayIntersectionWrapper");
    return RSH_OK;
}

rsh_status shadow_coloring(rsh_ctx& c, hipStream_t s, unsigned W, unsigned H, unsigned L, rsh_float3 eye,
                           unsigned* sflags, unsigned* pixels) {
    CTX_OK(need(c, c.g_pos, "bindFragmentPositionArray"));
    CTX_OK(need(c, c.g_nrm, "bindFragmentNormalArray"));
    CTX_OK(need(c, c.g_dif, "bindDiffuseTextureArray"));
    CTX_OK(need(c, c.g_spec, "bindSpecularTextureArray"));
    CTX_OK(need(c, c.light_pos, "bindLightPositions"));
    CTX_OK(need(c, c.light_col, "bindLightColors"));
    CTX_OK(need(c, c.light_int, "bindLightIntensities"));
    if (L > c.L || L > c.light_max)
        return fail(c, RSH_INVALID_ARG,
                    "ShadowRayColoringWrapper: lightTotal exceeds the bound lights or LIGHT_SOURCE_MAXIMUM");
    GBufView g{const_cast<float4*>(c.g_pos), const_cast<float4*>(c.g_nrm), const_cast<float4*>(c.g_dif),
               const_cast<float4*>(c.g_spec)};
    launch_shade_primary(g, W, H, c.light_max, Lights{c.light_pos, c.light_col, c.light_int, L},
                         single_view(W, H, V(eye)), sflags, 0u, pixels, s);
    CTX_HIP(hipGetLastError(), "ShadowRayColoringWrapper");
    return RSH_OK;
}

rsh_status reflection_preparation(rsh_ctx& c, hipStream_t s, unsigned W, unsigned H, unsigned* times) {
    launch_fill_u32(times, MISS32, (uint64_t)W * H * c.light_max, s);
    CTX_HIP(hipGetLastError(), "ReflectionRayPreparationWrapper");
    return RSH_OK;
}

rsh_status reflection_intersection(rsh_ctx& c, hipStream_t s, rsh_float3* rays, unsigned* svalues, unsigned* hits,
                                   rsh_float4* tri_pos, unsigned hitTotal, unsigned N, unsigned offset, unsigned W,
                                   unsigned* times) {
    launch_hits_intersect(reinterpret_cast<const float*>(rays), svalues, hits, reinterpret_cast<const float4*>(tri_pos),
                          hitTotal, N, offset, W, c.light_max, c.sub, nullptr, 0, times, s);
    CTX_HIP(hipGetLastError(), "ReflectionRayIntersectionWrapper");
    return RSH_OK;
}

rsh_status reflection_coloring(rsh_ctx& c, hipStream_t s, rsh_float4* tri_pos, rsh_float4* tri_nrm, unsigned W,
                               unsigned H, unsigned L, rsh_float3 eye, bool create, unsigned* times, rsh_float3* rays,
                               unsigned* flags, unsigned* keys, unsigned* values, unsigned* pixels) {
    CTX_OK(need(c, c.tri_mat, "bindTriangleMaterialIDs"));
    CTX_OK(need(c, c.mat_dif, "bindMaterialDiffuseProperties"));
    CTX_OK(need(c, c.mat_spec, "bindMaterialSpecularProperties"));
    CTX_OK(need(c, c.light_pos, "bindLightPositions"));
    CTX_OK(need(c, c.light_col, "bindLightColors"));
    CTX_OK(need(c, c.light_int, "bindLightIntensities"));
    if (L > c.L)
        return fail(c, RSH_INVALID_ARG, "ReflectionRayColoringWrapper: lightTotal exceeds the bound light count");
    launch_color_reflection_ref(reinterpret_cast<const float4*>(tri_pos), reinterpret_cast<const float4*>(tri_nrm), W,
                                H, c.light_max, scene_view(c, L), V(eye), create ? 1u : 0u, times,
                                reinterpret_cast<float*>(rays), flags, keys, values, pixels, s);
    CTX_HIP(hipGetLastError(), "ReflectionRayColoringWrapper");
    return RSH_OK;
}

// ------------------------------------------------- the drop-in wrappers ----
// One process-wide context (the reference's globals), the stream set with
// rsh_compat_set_stream, and the reference's exit-on-error.
struct Compat {
    rsh_ctx ctx;
    hipStream_t stream = nullptr;
    int exit_on_error = 1;
    int status = 0;
    std::string error;
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

void done(rsh_status st) {
    if (st != RSH_OK) raise(st, C().ctx.err);
}

rsh_ctx& X() { return C().ctx; }
hipStream_t XS() { return C().stream; }

// Public context calls: the context's device made current first.
rsh_status enter(rsh_ctx* c) {
    if (!c) return RSH_INVALID_ARG;
    const hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(*c, RSH_HIP_ERROR, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return RSH_OK;
}

#define CTX_ENTER(ctx)                                                                            \
    do {                                                                                          \
        const rsh_status r_ = enter(ctx);                                                         \
        if (r_ != RSH_OK) return r_;                                                              \
    } while (0)

}  // namespace

extern "C" {

// ------------------------------------------------------- context API -------
rsh_status rsh_ctx_create(int device, rsh_ctx** out) {
    if (!out) return RSH_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return RSH_INVALID_ARG;
    rsh_ctx* c = new rsh_ctx;
    c->device = device;
    *out = c;
    return RSH_OK;
}

void rsh_ctx_destroy(rsh_ctx* ctx) {
    if (!ctx) return;
    if (hipSetDevice(ctx->device) == hipSuccess) ctx->release();
    delete ctx;
}

rsh_status rsh_ctx_configure(rsh_ctx* ctx, unsigned int lightMaximum, unsigned int hierarchyDepth,
                             unsigned int hierarchySubdivision, int algorithmID) {
    if (!ctx) return RSH_INVALID_ARG;
    return configure(*ctx, lightMaximum, hierarchyDepth, hierarchySubdivision, algorithmID);
}

const char* rsh_ctx_last_error(const rsh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rsh_status rsh_set_gbuffer(rsh_ctx* ctx, const rsh_float4* diffuse, const rsh_float4* specular,
                           const rsh_float4* position, const rsh_float4* normal, unsigned int, unsigned int) {
    if (!ctx) return RSH_INVALID_ARG;
    ctx->g_dif = reinterpret_cast<const float4*>(diffuse);
    ctx->g_spec = reinterpret_cast<const float4*>(specular);
    ctx->g_pos = reinterpret_cast<const float4*>(position);
    ctx->g_nrm = reinterpret_cast<const float4*>(normal);
    return RSH_OK;
}

rsh_status rsh_set_scene(rsh_ctx* ctx, const float* pos, const float* nrm, const float* obj, const float* mat,
                         unsigned int T, const float* mdif, const float* mspec, unsigned int M, const float* spheres,
                         unsigned int K, const float* lpos, const float* lcol, const float* lint, unsigned int L) {
    if (!ctx) return RSH_INVALID_ARG;
    rsh_ctx& c = *ctx;
    c.tri_pos = reinterpret_cast<const float4*>(pos); c.tri_nrm = reinterpret_cast<const float4*>(nrm);
    c.tri_obj = reinterpret_cast<const int32_t*>(obj); c.tri_mat = reinterpret_cast<const int32_t*>(mat);
    c.T = T;
    c.mat_dif = reinterpret_cast<const float4*>(mdif); c.mat_spec = reinterpret_cast<const float4*>(mspec); c.M = M;
    c.spheres = reinterpret_cast<const float4*>(spheres); c.K = K;
    c.light_pos = reinterpret_cast<const float4*>(lpos); c.light_col = reinterpret_cast<const float4*>(lcol);
    c.light_int = reinterpret_cast<const float2*>(lint); c.L = L;
    return RSH_OK;
}

rsh_status rsh_triangle_update(rsh_ctx* ctx, void* stream, float* model, float* normal, unsigned int T,
                               rsh_float4* pos, rsh_float4* nrm) {
    CTX_ENTER(ctx);
    return triangle_update(*ctx, S(stream), model, normal, T, pos, nrm);
}
rsh_status rsh_bounding_sphere_update(rsh_ctx* ctx, void* stream, float* translation, float* scale, unsigned int K,
                                      rsh_float3* out) {
    CTX_ENTER(ctx);
    return sphere_update(*ctx, S(stream), translation, scale, K, out);
}
rsh_status rsh_memory_preparation(rsh_ctx* ctx, void*, unsigned int W, unsigned int H, unsigned int hitMemory) {
    CTX_ENTER(ctx);
    return memory_preparation(*ctx, W, H, hitMemory);
}
rsh_status rsh_shadow_ray_creation(rsh_ctx* ctx, void* stream, unsigned int W, unsigned int H, unsigned int L,
                                   rsh_float3* rays, unsigned int* flags, unsigned int* keys, unsigned int* values) {
    CTX_ENTER(ctx);
    return shadow_creation(*ctx, S(stream), W, H, L, rays, flags, keys, values);
}
rsh_status rsh_reflection_ray_creation(rsh_ctx* ctx, void* stream, unsigned int W, unsigned int H, rsh_float3 eye,
                                       rsh_float3* rays, unsigned int* flags, unsigned int* keys,
                                       unsigned int* values) {
    CTX_ENTER(ctx);
    return reflection_creation(*ctx, S(stream), W, H, eye, rays, flags, keys, values);
}
rsh_status rsh_refraction_ray_creation(rsh_ctx* ctx, void* stream, unsigned int W, unsigned int H, rsh_float3 eye,
                                       rsh_float3* rays, unsigned int* flags, unsigned int* keys,
                                       unsigned int* values) {
    CTX_ENTER(ctx);
    return refraction_creation(*ctx, S(stream), W, H, eye, rays, flags, keys, values);
}
rsh_status rsh_ray_trimming(rsh_ctx* ctx, void* stream, unsigned int* keys, unsigned int* values, unsigned int W,
                            unsigned int H, unsigned int L, unsigned int* flags, unsigned int*, unsigned int* tkeys,
                            unsigned int* tvalues, unsigned int* rayTotal, unsigned int* rayTotalDevice) {
    CTX_ENTER(ctx);
    return trimming(*ctx, S(stream), keys, values, W, H, L, flags, tkeys, tvalues, rayTotal, rayTotalDevice);
}
rsh_status rsh_ray_compression(rsh_ctx* ctx, void* stream, unsigned int* tkeys, unsigned int*, unsigned int N,
                               const unsigned int* rayTotalDevice, unsigned int*, unsigned int*, unsigned int* bases,
                               unsigned int* sizes, unsigned int* ckeys, unsigned int* cvalues,
                               unsigned int* chunkTotal, unsigned int* chunkTotalDevice) {
    CTX_ENTER(ctx);
    return compression(*ctx, S(stream), tkeys, N, rayTotalDevice, bases, sizes, ckeys, cvalues, chunkTotal,
                       chunkTotalDevice);
}
rsh_status rsh_ray_sorting(rsh_ctx* ctx, void* stream, unsigned int* ckeys, unsigned int* cvalues, unsigned int Cn,
                           const unsigned int* chunkTotalDevice, unsigned int* skeys, unsigned int* svalues) {
    CTX_ENTER(ctx);
    return sorting(*ctx, S(stream), ckeys, cvalues, Cn, chunkTotalDevice, 32, skeys, svalues, "RaySortingWrapper");
}
rsh_status rsh_ctx_sort_pairs(rsh_ctx* ctx, void* stream, const unsigned int* keys, const unsigned int* values,
                              unsigned int n, const unsigned int* nDevice, unsigned int bits, unsigned int* out_keys,
                              unsigned int* out_values) {
    CTX_ENTER(ctx);
    return sorting(*ctx, S(stream), keys, values, n, nDevice, bits, out_keys, out_values, "rsh_ctx_sort_pairs");
}
rsh_status rsh_ray_decompression(rsh_ctx* ctx, void* stream, unsigned int* bases, unsigned int* sizes, unsigned int*,
                                 unsigned int* scvalues, unsigned int* tkeys, unsigned int* tvalues,
                                 unsigned int* skeleton, unsigned int* scan, unsigned int Cn,
                                 const unsigned int* chunkTotalDevice, unsigned int* rkeys, unsigned int* rvalues) {
    CTX_ENTER(ctx);
    return decompression(*ctx, S(stream), bases, sizes, scvalues, tkeys, tvalues, skeleton, scan, Cn,
                         chunkTotalDevice, rkeys, rvalues);
}
rsh_status rsh_hierarchy_creation(rsh_ctx* ctx, void* stream, rsh_float3* rays, unsigned int*, unsigned int* svalues,
                                  unsigned int N, const unsigned int* rayTotalDevice, float r0, float s0,
                                  rsh_float4* hierarchy) {
    CTX_ENTER(ctx);
    return hierarchy_creation(*ctx, S(stream), rays, svalues, N, rayTotalDevice, r0, s0, hierarchy);
}
rsh_status rsh_hierarchy_traversal_warmup(rsh_ctx* ctx, void* stream, rsh_float4* hierarchy, rsh_float3* spheres,
                                          rsh_float4* tri_pos, unsigned int K, unsigned int N, unsigned int Tb,
                                          unsigned int offset, unsigned int* flags, unsigned int* scan,
                                          unsigned int* hits, unsigned int* thits, unsigned int* hitTotal,
                                          unsigned int* hitMemory) {
    CTX_ENTER(ctx);
    return traversal_warmup(*ctx, S(stream), hierarchy, spheres, tri_pos, K, N, Tb, offset, flags, scan, hits, thits,
                            hitTotal, hitMemory);
}
rsh_status rsh_hierarchy_traversal(rsh_ctx* ctx, void* stream, rsh_float4* hierarchy, rsh_float4* tri_pos,
                                   unsigned int N, unsigned int, unsigned int offset, unsigned int* flags,
                                   unsigned int*, unsigned int* hits, unsigned int* thits, unsigned int* hitTotal,
                                   unsigned int* hitMemory) {
    CTX_ENTER(ctx);
    return traversal(*ctx, S(stream), hierarchy, tri_pos, N, offset, flags, hits, thits, hitTotal, hitMemory);
}
rsh_status rsh_shadow_ray_preparation(rsh_ctx* ctx, void* stream, unsigned int W, unsigned int H, unsigned int L,
                                      unsigned int* f) {
    CTX_ENTER(ctx);
    return shadow_preparation(*ctx, S(stream), W, H, L, f);
}
rsh_status rsh_shadow_ray_intersection(rsh_ctx* ctx, void* stream, rsh_float3* rays, unsigned int*,
                                       unsigned int* svalues, unsigned int* hits, rsh_float4* tri_pos,
                                       unsigned int hitTotal, unsigned int N, unsigned int offset, unsigned int W,
                                       unsigned int, rsh_float3, unsigned int* sflags) {
    CTX_ENTER(ctx);
    return shadow_intersection(*ctx, S(stream), rays, svalues, hits, tri_pos, hitTotal, N, offset, W, sflags);
}
rsh_status rsh_shadow_ray_coloring(rsh_ctx* ctx, void* stream, unsigned int W, unsigned int H, unsigned int L,
                                   rsh_float3 eye, unsigned int* sflags, unsigned int* pixels) {
    CTX_ENTER(ctx);
    return shadow_coloring(*ctx, S(stream), W, H, L, eye, sflags, pixels);
}
rsh_status rsh_reflection_ray_preparation(rsh_ctx* ctx, void* stream, unsigned int W, unsigned int H,
                                          unsigned int* times) {
    CTX_ENTER(ctx);
    return reflection_preparation(*ctx, S(stream), W, H, times);
}
rsh_status rsh_reflection_ray_intersection(rsh_ctx* ctx, void* stream, rsh_float3* rays, unsigned int*,
                                           unsigned int* svalues, unsigned int* hits, rsh_float4* tri_pos,
                                           unsigned int hitTotal, unsigned int N, unsigned int offset, unsigned int W,
                                           unsigned int, unsigned int, rsh_float3, unsigned int* times) {
    CTX_ENTER(ctx);
    return reflection_intersection(*ctx, S(stream), rays, svalues, hits, tri_pos, hitTotal, N, offset, W, times);
}
rsh_status rsh_reflection_ray_coloring(rsh_ctx* ctx, void* stream, rsh_float4* tri_pos, rsh_float4* tri_nrm,
                                       unsigned int W, unsigned int H, unsigned int L, rsh_float3 eye, bool create,
                                       unsigned int* times, rsh_float3* rays, unsigned int* flags, unsigned int* keys,
                                       unsigned int* values, unsigned int* pixels) {
    CTX_ENTER(ctx);
    return reflection_coloring(*ctx, S(stream), tri_pos, tri_nrm, W, H, L, eye, create, times, rays, flags, keys,
                               values, pixels);
}

// ------------------------------------------------------- drop-in setup -----
void rsh_compat_configure(unsigned int lightMaximum, unsigned int hierarchyDepth, unsigned int hierarchySubdivision,
                          int algorithmID) {
    if (configure(X(), lightMaximum, hierarchyDepth, hierarchySubdivision, algorithmID) != RSH_OK)
        raise(RSH_INVALID_ARG, "rsh_compat_configure: invalid constants");
}
void rsh_compat_set_stream(void* s) { C().stream = static_cast<hipStream_t>(s); }
void rsh_compat_set_exit_on_error(int v) { C().exit_on_error = v; }
int rsh_compat_status(void) { return C().status; }
const char* rsh_compat_error(void) { return C().error.c_str(); }
void rsh_compat_reset_status(void) { C().status = 0; C().error.clear(); }

// ------------------------------------------------------------- binders ----
void bindDiffuseTextureArray(const rsh_float4* p) { X().g_dif = reinterpret_cast<const float4*>(p); }
void bindSpecularTextureArray(const rsh_float4* p) { X().g_spec = reinterpret_cast<const float4*>(p); }
void bindFragmentPositionArray(const rsh_float4* p) { X().g_pos = reinterpret_cast<const float4*>(p); }
void bindFragmentNormalArray(const rsh_float4* p) { X().g_nrm = reinterpret_cast<const float4*>(p); }
void bindTrianglePositions(float* p, unsigned int T) { X().tri_pos = reinterpret_cast<const float4*>(p); X().T = T; }
void bindTriangleNormals(float* p, unsigned int T) { X().tri_nrm = reinterpret_cast<const float4*>(p); (void)T; }
void bindTriangleTextureCoordinates(float* p, unsigned int T) { (void)p; (void)T; }   // no stage reads them
void bindTriangleObjectIDs(float* p, unsigned int T) { X().tri_obj = reinterpret_cast<const int32_t*>(p); (void)T; }
void bindTriangleMaterialIDs(float* p, unsigned int T) { X().tri_mat = reinterpret_cast<const int32_t*>(p); (void)T; }
void bindMaterialDiffuseProperties(float* p, unsigned int M) { X().mat_dif = reinterpret_cast<const float4*>(p); X().M = M; }
void bindMaterialSpecularProperties(float* p, unsigned int M) { X().mat_spec = reinterpret_cast<const float4*>(p); (void)M; }
void bindBoundingSpheres(float* p, unsigned int K) { X().spheres = reinterpret_cast<const float4*>(p); X().K = K; }
void bindLightPositions(float* p, unsigned int L) { X().light_pos = reinterpret_cast<const float4*>(p); X().L = L; }
void bindLightColors(float* p, unsigned int L) { X().light_col = reinterpret_cast<const float4*>(p); (void)L; }
void bindLightIntensities(float* p, unsigned int L) { X().light_int = reinterpret_cast<const float2*>(p); (void)L; }

// ------------------------------------------------------------- stages -----
void TriangleUpdateWrapper(float* model, float* normal, unsigned int T, rsh_float4* pos, rsh_float4* nrm) {
    done(triangle_update(X(), XS(), model, normal, T, pos, nrm));
}

void BoundingSphereUpdateWrapper(float* translation, float* scale, const unsigned int K, rsh_float3* out) {
    done(sphere_update(X(), XS(), translation, scale, K, out));
}

void MemoryPreparationWrapper(const unsigned int W, const unsigned int H, const unsigned int hitMemory,
                              unsigned int*, unsigned int*, unsigned int*, unsigned int*, unsigned int*,
                              unsigned int*) {
    done(memory_preparation(X(), W, H, hitMemory));
}

void ShadowRayCreationWrapper(const unsigned int W, const unsigned int H, const unsigned int L, rsh_float3* rays,
                              unsigned int* flags, unsigned int* keys, unsigned int* values) {
    done(shadow_creation(X(), XS(), W, H, L, rays, flags, keys, values));
}

void ReflectionRayCreationWrapper(const unsigned int W, const unsigned int H, const rsh_float3 eye, rsh_float3* rays,
                                  unsigned int* flags, unsigned int* keys, unsigned int* values) {
    done(reflection_creation(X(), XS(), W, H, eye, rays, flags, keys, values));
}

void RefractionRayCreationWrapper(const unsigned int W, const unsigned int H, const rsh_float3 eye, rsh_float3* rays,
                                  unsigned int* flags, unsigned int* keys, unsigned int* values) {
    done(refraction_creation(X(), XS(), W, H, eye, rays, flags, keys, values));
}

void RayTrimmingWrapper(unsigned int* keys, unsigned int* values, const unsigned int W, const unsigned int H,
                        const unsigned int L, unsigned int* flags, unsigned int*, unsigned int* tkeys,
                        unsigned int* tvalues, unsigned int* rayTotal) {
    done(trimming(X(), XS(), keys, values, W, H, L, flags, tkeys, tvalues, rayTotal, nullptr));
}

void RayCompressionWrapper(unsigned int* tkeys, unsigned int*, const unsigned int N, unsigned int*, unsigned int*,
                           unsigned int* bases, unsigned int* sizes, unsigned int* ckeys, unsigned int* cvalues,
                           unsigned int* chunkTotal) {
    done(compression(X(), XS(), tkeys, N, nullptr, bases, sizes, ckeys, cvalues, chunkTotal, nullptr));
}

void RaySortingWrapper(unsigned int* ckeys, unsigned int* cvalues, const unsigned int Cn, unsigned int* skeys,
                       unsigned int* svalues) {
    done(sorting(X(), XS(), ckeys, cvalues, Cn, nullptr, 32, skeys, svalues, "RaySortingWrapper"));
}

// The sorting stage on the low `bits` key bits (stable; 8- or 9-bit digits),
// as the engine runs it for a frame's shadow and reflection passes.
void rsh_sort_pairs(const unsigned int* keys, const unsigned int* values, const unsigned int n,
                    const unsigned int bits, unsigned int* out_keys, unsigned int* out_values) {
    rsh_status st = sorting(X(), XS(), keys, values, n, nullptr, bits, out_keys, out_values, "rsh_sort_pairs");
    if (st == RSH_OK) {
        const hipError_t e = hipStreamSynchronize(XS());
        if (e != hipSuccess) st = fail(X(), RSH_HIP_ERROR, std::string("rsh_sort_pairs: ") + hipGetErrorString(e));
    }
    done(st);
}

void RayDecompressionWrapper(unsigned int* bases, unsigned int* sizes, unsigned int*, unsigned int* scvalues,
                             unsigned int* tkeys, unsigned int* tvalues, unsigned int* skeleton, unsigned int* scan,
                             const unsigned int Cn, unsigned int* rkeys, unsigned int* rvalues) {
    done(decompression(X(), XS(), bases, sizes, scvalues, tkeys, tvalues, skeleton, scan, Cn, nullptr, rkeys,
                       rvalues));
}

void HierarchyCreationWrapper(rsh_float3* rays, unsigned int*, unsigned int* svalues, const unsigned int N,
                              const float r0, const float s0, rsh_float4* hierarchy) {
    done(hierarchy_creation(X(), XS(), rays, svalues, N, nullptr, r0, s0, hierarchy));
}

void HierarchyTraversalWarmUpWrapper(rsh_float4* hierarchy, rsh_float3* spheres, rsh_float4* tri_pos,
                                     const unsigned int K, const unsigned int N, const unsigned int Tb,
                                     const unsigned int offset, unsigned int* flags, unsigned int* scan,
                                     unsigned int* hits, unsigned int* thits, unsigned int* hitTotal,
                                     unsigned int* hitMemory) {
    done(traversal_warmup(X(), XS(), hierarchy, spheres, tri_pos, K, N, Tb, offset, flags, scan, hits, thits,
                          hitTotal, hitMemory));
}

void HierarchyTraversalWrapper(rsh_float4* hierarchy, rsh_float4* tri_pos, const unsigned int N, const unsigned int Tb,
                               const unsigned int offset, unsigned int* flags, unsigned int*, unsigned int* hits,
                               unsigned int* thits, unsigned int* hitTotal, unsigned int* hitMemory) {
    (void)Tb;
    done(traversal(X(), XS(), hierarchy, tri_pos, N, offset, flags, hits, thits, hitTotal, hitMemory));
}

void ShadowRayPreparationWrapper(const unsigned int W, const unsigned int H, const unsigned int L, unsigned int* f) {
    done(shadow_preparation(X(), XS(), W, H, L, f));
}

void ShadowRayIntersectionWrapper(rsh_float3* rays, unsigned int*, unsigned int* svalues, unsigned int* hits,
                                  rsh_float4* tri_pos, const unsigned int hitTotal, const unsigned int N,
                                  const unsigned int offset, const unsigned int W, const unsigned int H,
                                  const rsh_float3, unsigned int* sflags) {
    (void)H;
    done(shadow_intersection(X(), XS(), rays, svalues, hits, tri_pos, hitTotal, N, offset, W, sflags));
}

void ShadowRayColoringWrapper(const unsigned int W, const unsigned int H, const unsigned int L, const rsh_float3 eye,
                              unsigned int* sflags, unsigned int* pixels) {
    done(shadow_coloring(X(), XS(), W, H, L, eye, sflags, pixels));
}

void ReflectionRayPreparationWrapper(const unsigned int W, const unsigned int H, unsigned int* times) {
    done(reflection_preparation(X(), XS(), W, H, times));
}

void ReflectionRayIntersectionWrapper(rsh_float3* rays, unsigned int*, unsigned int* svalues, unsigned int* hits,
                                      rsh_float4* tri_pos, const unsigned int hitTotal, const unsigned int N,
                                      const unsigned int offset, const unsigned int W, const unsigned int,
                                      const unsigned int, const rsh_float3, unsigned int* times) {
    done(reflection_intersection(X(), XS(), rays, svalues, hits, tri_pos, hitTotal, N, offset, W, times));
}

void ReflectionRayColoringWrapper(rsh_float4* tri_pos, rsh_float4* tri_nrm, const unsigned int W, const unsigned int H,
                                  const unsigned int L, const rsh_float3 eye, const bool create, unsigned int* times,
                                  rsh_float3* rays, unsigned int* flags, unsigned int* keys, unsigned int* values,
                                  unsigned int* pixels) {
    done(reflection_coloring(X(), XS(), tri_pos, tri_nrm, W, H, L, eye, create, times, rays, flags, keys, values,
                             pixels));
}

}  // extern "C"
