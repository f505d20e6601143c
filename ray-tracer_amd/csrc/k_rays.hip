// k_rays.hip -- ray creation and hashing, surface records and shading (gfx950).
//
// Reference: projects/CUDA-Engine/source/RayTracer.cu
//   UpdateVertex :448-511, UpdateBoundingSphere :513-570
//   CreateShadowRays :682-740, CreateReflectionRays :742-791
//   ColorPrimaryShadowRay :1616-1709, ColorReflectionRay :1776-1882,
//   ColorAndCreateReflectionRay :1884-2017
// plus the primary-ray stage that replaces the OpenGL G-buffer
// (main.cpp:1082-1108; formula of RayTracer.cu:667).
//
// All kernels are one thread per slot or pixel with coalesced row-major
// access; the ray buffer keeps the reference's interleaved layout
// (origin at 2*slot, direction at 2*slot+1, float3), written as float2 pairs.
#include "rsh_kernels.hpp"

namespace rsh {

namespace {

inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 1u << 20) {
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return (uint32_t)(g < cap ? g : cap);
}

__device__ __forceinline__ void store_ray(float* rays, size_t slot, v3 o, v3 d) {
    float2* r = reinterpret_cast<float2*>(rays + 6 * slot);
    r[0] = make_float2(o.x, o.y);
    r[1] = make_float2(o.z, d.x);
    r[2] = make_float2(d.y, d.z);
}
__device__ __forceinline__ void load_ray(const float* rays, size_t slot, v3& o, v3& d) {
    const float2* r = reinterpret_cast<const float2*>(rays + 6 * slot);
    float2 a = r[0], b = r[1], c = r[2];
    o = mk(a.x, a.y, b.x);
    d = mk(b.y, c.x, c.y);
}

// type: 0 primary, 1 shadow (light l), 2 mirror, 3 refraction
// seg: the segment (view) id OR-ed in at kc.h.segment_shift when that is set.
__device__ __forceinline__ uint32_t make_key(const KeyCfg& kc, v3 o, v3 d, uint32_t type, uint32_t l, uint32_t seg) {
    if (kc.mode == 0)
        return type == 1 ? shadow_key_ref(d, l) : type == 3 ? refraction_key_ref(o, d) : reflection_key_ref(o, d);
    uint32_t k = key_oct(kc.h, o, d, type);
    if (kc.h.segment_shift) k |= seg << kc.h.segment_shift;
    return k;
}
__device__ __forceinline__ uint32_t seg_of(const KeyCfg& kc, const Views& V, const uint32_t* s_off, bool lds,
                                           uint32_t p) {
    return kc.h.segment_shift ? view_index(V, s_off, lds, p) : 0u;
}

// Pixel p of view v: (x, y) within the view, dir = D + R*(x/w - 0.5) + U*(y/h - 0.5).
// Primary rays are keyed by screen position, not by the hash (the primary
// stage replaces the reference's rasterised G-buffer, so it has no reference
// order): 8x8 pixel tiles in row-major order over the view, Morton order inside
// a tile, the view id above segment_shift.  Sorted, a leaf of 8 rays is a 4x2
// pixel block and a 64-ray node one tile (measured on the 1080p atrium: the
// primary traversal 1.82 -> 1.00 ms against the direction-hash order).
__device__ __forceinline__ uint32_t primary_key(uint32_t x, uint32_t y, uint32_t w, uint32_t view,
                                                uint32_t segment_shift) {
    const uint32_t lx = x & 7u, ly = y & 7u;
    const uint32_t m = (lx & 1u) | ((ly & 1u) << 1) | ((lx & 2u) << 1) | ((ly & 2u) << 2) | ((lx & 4u) << 2) |
                       ((ly & 4u) << 3);
    uint32_t k = (((y >> 3) * ((w + 7u) >> 3) + (x >> 3)) << 6) | m;
    if (segment_shift) k |= view << segment_shift;
    return k;
}

__device__ __forceinline__ v3 primary_dir(const ViewDesc& vd, uint32_t x, uint32_t y) {
    const float a = (float)x / (float)vd.w - 0.5f;
    const float c = (float)y / (float)vd.h - 0.5f;
    const v3 D = mk(vd.dir[0], vd.dir[1], vd.dir[2]), R = mk(vd.right[0], vd.right[1], vd.right[2]);
    const v3 U = mk(vd.up[0], vd.up[1], vd.up[2]);
    return normalize((D + R * a) + U * c);
}

// Flags: F = uint8_t for the engine's own slot flags (one byte per slot),
// uint32_t for the stage API's buffers; values == nullptr: not written (the
// engine's trim takes each live slot's index as its value).
template <typename F>
__global__ void __launch_bounds__(256) k_create_primary(Views V, uint32_t P, KeyCfg kc, float* rays, F* flags,
                                                        uint32_t* keys, uint32_t* values) {
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const uint32_t v = view_index(V, s_off, lds, p);
        const ViewDesc vd = V.n ? V.d[v] : V.one;
        const uint32_t q = p - vd.off;
        const uint32_t x = q % vd.w, y = q / vd.w;
        store_ray(rays, p, mk(vd.eye[0], vd.eye[1], vd.eye[2]), primary_dir(vd, x, y));
        keys[p] = primary_key(x, y, vd.w, v, kc.h.segment_shift);
        if (values) values[p] = p;
        flags[p] = 0;
    }
}

// The camera rays already in key order.  The keys are unique, so sorted
// position q holds, within its view's block (views in offset order, which is
// view-id order), the q-th pixel in tile-major, Morton-inside order.  Tile
// rows hold 8 w pixels (the last one h_t w); in a tile row every tile left of
// tile tx is 8 wide, so it holds 8 h_t; inside a tile of w_t x h_t valid
// pixels the position is the rank among the Morton indices m with
// (mx, my) < (w_t, h_t).  Writes the slot-order rays (the surface stage reads
// them) and the sorted hierarchy records of k_gather: rec0 = (o, slot bits),
// rec1 = (d, 0).
__global__ void __launch_bounds__(256) k_create_primary_sorted(Views V, uint32_t P, float4* rec0, float4* rec1) {
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < P; q += gridDim.x * blockDim.x) {
        const ViewDesc vd = view_at(V, s_off, lds, q);
        const uint32_t w = vd.w, h = vd.h, r = q - vd.off;
        const uint32_t ty = r / (8u * w), r1 = r - ty * 8u * w;
        const uint32_t ht = min(8u, h - 8u * ty);
        const uint32_t tx = r1 / (8u * ht), r2 = r1 - tx * 8u * ht;
        const uint32_t wt = min(8u, w - 8u * tx);
        uint32_t m = r2;
        if (wt != 8u || ht != 8u) {
            uint32_t seen = 0;
            for (m = 0; m < 64u; m++) {
                const uint32_t mx = (m & 1u) | ((m >> 1) & 2u) | ((m >> 2) & 4u);
                const uint32_t my = ((m >> 1) & 1u) | ((m >> 2) & 2u) | ((m >> 3) & 4u);
                if (mx < wt && my < ht && seen++ == r2) break;
            }
        }
        const uint32_t x = 8u * tx + ((m & 1u) | ((m >> 1) & 2u) | ((m >> 2) & 4u));
        const uint32_t y = 8u * ty + (((m >> 1) & 1u) | ((m >> 2) & 2u) | ((m >> 3) & 4u));
        const uint32_t p = vd.off + x + y * w;
        const v3 o = mk(vd.eye[0], vd.eye[1], vd.eye[2]), d = primary_dir(vd, x, y);
        rec0[q] = f4(o, __uint_as_float(p));
        rec1[q] = f4(d, 0.0f);
    }
}

// Ray queries: the caller's rays, keyed like frame rays of the given type.
template <typename F>
__global__ void __launch_bounds__(256) k_query_keys(const float* rays, const uint32_t* segs, uint32_t n, KeyCfg kc,
                                                    uint32_t type, F* flags, uint32_t* keys, uint32_t* values) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v3 o, d;
        load_ray(rays, i, o, d);
        keys[i] = make_key(kc, o, d, type, 0, segs ? segs[i] : 0u);
        if (values) values[i] = i;
        flags[i] = 0;
    }
}

// RayTracer.cu:682-740; slot i = x + W*(l + Lmax*y).
template <typename F>
__global__ void __launch_bounds__(256) k_create_shadow(const float4* gpos, const float4* gnrm, uint32_t W, uint32_t H,
                                                       uint32_t Lmax, const float4* light_pos, KeyCfg kc, Views V,
                                                       uint32_t soft, float* rays, F* flags, uint32_t* keys,
                                                       uint32_t* values, FastDiv fw, FastDiv fl) {
    // soft: 16 rays per slot, ray 16 i + c from grid point c about the light
    // (RayTracer.cu:1550-1603) towards the slot's surface point
    const uint32_t sh = soft ? 4u : 0u;
    const uint32_t n = (W * H * Lmax) << sh;
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const uint32_t i = r >> sh;
        const uint32_t qw = fdiv(i, fw);          // l + Lmax y
        const uint32_t x = i - qw * W, y = fdiv(qw, fl), l = qw - y * Lmax;
        size_t px = (size_t)x + (size_t)y * W;
        v3 P = xyz(gpos[px]);
        const v3 Nr = xyz(gnrm[px]);
        uint32_t flag = 1;
        if (dot(P, P) != 0.0f) {                   // len(P) != 0 without the square root
            const v3 N = normalize(Nr);
            const v3 Lp = xyz(light_pos[l]);
            const float diffuse = fmaxf(dot(normalize(P - Lp), -N), 0.0f);
            if (diffuse > 0.0f) {
                v3 o, d;
                shadow_ray(P, Lp, soft != 0u, r & 15u, o, d);
                if (rays) store_ray(rays, r, o, d);     // null: the gather recomputes them
                keys[r] = make_key(kc, o, d, 1, l, seg_of(kc, V, s_off, lds, (uint32_t)px));
                if (values) values[r] = r;
                flag = 0;
            }
        }
        flags[r] = (F)flag;
    }
}

// RayTracer.cu:742-791; slot x + y*W*Lmax.  gate: only diffuse.w > 0 reflects.
// mark_gaps: also flag the unused slots between rows when Lmax > 1.
template <typename F>
__global__ void __launch_bounds__(256) k_create_reflection(const float4* gpos, const float4* gnrm, const float4* gdif,
                                                           uint32_t W, uint32_t H, uint32_t Lmax, Views V,
                                                           uint32_t gate, uint32_t mark_gaps, KeyCfg kc, float* rays,
                                                           F* flags, uint32_t* keys, uint32_t* values, FastDiv fw) {
    const uint32_t n = W * H;
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t y = fdiv(p, fw), x = p - y * W;
        uint32_t slot = x + y * W * Lmax;
        // the gate first: the surface point and normal of a pixel that does
        // not reflect are never read
        bool live = !gate || gdif[p].w > 0.0f;
        v3 P = mk(0.0f, 0.0f, 0.0f);
        if (live) {
            P = xyz(gpos[p]);
            live = dot(P, P) != 0.0f;              // len(P) != 0 without the square root
        }
        if (live) {
            const v3 N = normalize(xyz(gnrm[p]));
            const uint32_t v = view_index(V, s_off, lds, p);
            v3 d = reflect(normalize(P - view_eye(V, v)), normalize(N));
            store_ray(rays, slot, P, d);
            keys[slot] = make_key(kc, P, d, 2, 0, v);
            if (values) values[slot] = slot;
            flags[slot] = 0;
        } else {
            flags[slot] = 1;
        }
        if (mark_gaps)
            for (uint32_t j = 1; j < Lmax; j++) flags[slot + j * W] = 1;
    }
}

// RayTracer.cu:793-842 CreateRefractionRays (no caller in the reference):
// slot x + y*W*Lmax, origin = the surface point, direction = refract(view
// direction, normal, 1 / 1.52) for every covered pixel.
__global__ void __launch_bounds__(256) k_create_refraction(const float4* gpos, const float4* gnrm, uint32_t W,
                                                           uint32_t H, uint32_t Lmax, Views V, KeyCfg kc, float* rays,
                                                           uint32_t* flags, uint32_t* keys, uint32_t* values) {
    const uint32_t n = W * H;
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        uint32_t x = p % W, y = p / W;
        uint32_t slot = x + y * W * Lmax;
        v3 P = xyz(gpos[p]);
        v3 N = normalize(xyz(gnrm[p]));
        if (len(P) != 0.0f) {
            const uint32_t v = view_index(V, s_off, lds, p);
            v3 d = refract(normalize(P - view_eye(V, v)), normalize(N), 1.0f / 1.52f);
            store_ray(rays, slot, P, d);
            keys[slot] = make_key(kc, P, d, 3, 0, v);
            values[slot] = slot;
            flags[slot] = 0;
        } else {
            flags[slot] = 1;
        }
    }
}

// Next bounce (RayTracer.cu:1995-2010): origin = hit point, direction =
// reflect(previous direction, barycentric normal), in place.
template <typename F>
__global__ void __launch_bounds__(256) k_create_bounce(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax,
                                                       const unsigned long long* rec, uint32_t hit_mode,
                                                       uint32_t gate, KeyCfg kc, Views V, float* rays, F* flags,
                                                       uint32_t* keys, uint32_t* values, FastDiv fw) {
    const uint32_t n = W * H;
    const unsigned long long miss = hit_mode == 0 ? MISS64 : (unsigned long long)MISS32;
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t y = fdiv(p, fw), x = p - y * W;
        uint32_t slot = x + y * W * Lmax;
        bool live = rec[slot] != miss && (!gate || g.dif[p].w > 0.0f);
        if (live) {
            v3 o, d;
            load_ray(rays, slot, o, d);
            v3 P = xyz(g.pos[p]), N = xyz(g.nrm[p]);
            v3 nd = reflect(d, N);
            store_ray(rays, slot, P, nd);
            keys[slot] = make_key(kc, P, nd, 2, 0, seg_of(kc, V, s_off, lds, p));
            if (values) values[slot] = slot;
            flags[slot] = 0;
        } else {
            flags[slot] = 1;
        }
        for (uint32_t j = 1; j < Lmax; j++) flags[slot + j * W] = 1;
    }
}

// Surface record at a hit (RayTracer.cu:1807-1843): point o + d*(eps + t),
// area-ratio barycentric normal, material of the triangle's first vertex.
// rays == nullptr: the camera pass of views V; each hit pixel's ray is
// recomputed as k_create_primary_sorted made it (same view, same operations).
__global__ void __launch_bounds__(256) k_surface(const float* rays, const unsigned long long* rec, uint32_t hit_mode,
                                                 uint32_t W, uint32_t H, uint32_t slot_stride, uint32_t normalize_n,
                                                 SceneView sc, GBufView g, Views V, FastDiv fw) {
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = rays == nullptr && stage_views(V, s_off);
    const uint32_t n = W * H;
    const unsigned long long miss = hit_mode == 0 ? MISS64 : (unsigned long long)MISS32;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t y = fdiv(p, fw), x = p - y * W;
        uint32_t slot = x + y * W * slot_stride;
        unsigned long long r = rec[slot];
        if (r == miss) {
            float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            g.pos[p] = z; g.nrm[p] = z; g.dif[p] = z; g.spec[p] = z;
            continue;
        }
        uint32_t t = hit_mode == 0 ? (uint32_t)(r & 0xFFFFFFFFull) : (uint32_t)(r & 0x000FFFFFull);
        v3 v0 = xyz(sc.tri_pos[3 * (size_t)t]), v1 = xyz(sc.tri_pos[3 * (size_t)t + 1]);
        v3 v2 = xyz(sc.tri_pos[3 * (size_t)t + 2]);
        v3 o, d;
        if (rays) {
            load_ray(rays, slot, o, d);
        } else {
            const ViewDesc vd = view_at(V, s_off, lds, p);
            const uint32_t q = p - vd.off;
            o = mk(vd.eye[0], vd.eye[1], vd.eye[2]);
            d = primary_dir(vd, q % vd.w, q / vd.w);
        }
        float th = ray_triangle(o + d * EPS, d, v0, v1 - v0, v2 - v0);
        v3 P = o + d * (EPS + th);
        float aABC = len(cross(v1 - v0, v2 - v0));
        float aPBC = len(cross(v1 - P, v2 - P));
        float aPCA = len(cross(v0 - P, v2 - P));
        v3 N = ((aPBC / aABC) * xyz(sc.tri_nrm[3 * (size_t)t]) + (aPCA / aABC) * xyz(sc.tri_nrm[3 * (size_t)t + 1])) +
               ((1.0f - (aPBC / aABC)) - (aPCA / aABC)) * xyz(sc.tri_nrm[3 * (size_t)t + 2]);
        if (normalize_n) N = normalize(N);
        int32_t m = sc.tri_mat[3 * (size_t)t];
        g.pos[p] = f4(P, 1.0f);
        g.nrm[p] = f4(N, 0.0f);
        g.dif[p] = sc.mat_dif[m];
        g.spec[p] = sc.mat_spec[m];
    }
}

// RayTracer.cu:1616-1709 ColorPrimaryShadowRay
// soft: 16 flags per slot (soft shadows), shaded by their grid weights.
__device__ inline uint32_t soft_mask(const uint32_t* sflags, size_t slot) {
    const uint4* q = reinterpret_cast<const uint4*>(sflags + 16 * slot);
    uint32_t m = 0;
    for (uint32_t k = 0; k < 4; k++) {
        const uint4 v = q[k];
        m |= (v.x == MISS32 ? 1u : 0u) << (4 * k);
        m |= (v.y == MISS32 ? 2u : 0u) << (4 * k);
        m |= (v.z == MISS32 ? 4u : 0u) << (4 * k);
        m |= (v.w == MISS32 ? 8u : 0u) << (4 * k);
    }
    return m;
}

__global__ void __launch_bounds__(256) k_shade_primary(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, Lights L,
                                                       Views V, const uint32_t* sflags, uint32_t soft,
                                                       uint32_t* pixels, const uint32_t* live, FastDiv fw) {
    // a shadow pass with no live ray leaves the pixels as they are (the frame
    // driver shades only after a non-empty shadow pass); live: its device count
    if (live && *live == 0u) return;
    const uint32_t n = W * H;
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t y = fdiv(p, fw), x = p - y * W;
        v3 color = mk(0.0f, 0.0f, 0.0f);
        v3 P = xyz(g.pos[p]), N = xyz(g.nrm[p]);
        if (dot(P, P) != 0.0f) {                   // len(P) != 0 without the square root
            const v3 eye = view_eye(V, view_index(V, s_off, lds, p));
            float4 Kd = g.dif[p], Ks = g.spec[p];
            for (uint32_t l = 0; l < L.total; l++) {
                const size_t slot = (size_t)(x + y * W * Lmax) + (size_t)W * l;
                if (!soft) {
                    if (sflags[slot] != MISS32) color = color + light_term(L, l, P, N, eye, Kd, Ks, false);
                } else {
                    const uint32_t m = soft_mask(sflags, slot);
                    if (m != 0xFFFFu) color = color + light_term(L, l, P, N, eye, Kd, Ks, false, true, soft_factor(m));
                }
            }
        }
        pixels[p] = rgb_to_int(color.x * 255.0f, color.y * 255.0f, color.z * 255.0f);
    }
}

// RayTracer.cu:1776-1882 shading from the surface record; with bounce shadow
// flags a light whose shadow ray is occluded is skipped.
__global__ void __launch_bounds__(256) k_shade_reflection(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, Lights L,
                                                          Views V, const unsigned long long* rec, uint32_t hit_mode,
                                                          const uint32_t* sflags, uint32_t soft, uint32_t* pixels,
                                                          FastDiv fw) {
    const uint32_t n = W * H;
    const unsigned long long miss = hit_mode == 0 ? MISS64 : (unsigned long long)MISS32;
    __shared__ uint32_t s_off[VIEW_LDS];
    const bool lds = stage_views(V, s_off);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t y = fdiv(p, fw), x = p - y * W;
        uint32_t slot = x + y * W * Lmax;
        if (rec[slot] == miss) continue;
        const v3 eye = view_eye(V, view_index(V, s_off, lds, p));
        v3 P = xyz(g.pos[p]), N = xyz(g.nrm[p]);
        float4 Kd = g.dif[p], Ks = g.spec[p];
        v3 color = mk(0.0f, 0.0f, 0.0f);
        for (uint32_t l = 0; l < L.total; l++) {
            const size_t sl = (size_t)slot + (size_t)W * l;
            if (sflags && soft) {
                const uint32_t m = soft_mask(sflags, sl);
                if (m != 0xFFFFu) color = color + light_term(L, l, P, N, eye, Kd, Ks, true, true, soft_factor(m));
                continue;
            }
            if (sflags && sflags[sl] == MISS32) continue;
            color = color + light_term(L, l, P, N, eye, Kd, Ks, true);
        }
        pixels[p] += rgb_to_int(color.x * Ks.w, color.y * Ks.w, color.z * Ks.w);
    }
}

// Per-triangle data of the fused traversal: bounding sphere (RayTracer.cu:301-376
// evaluated once per triangle instead of once per candidate pair) and the
// vertex / edge vectors of the Moller-Trumbore test.
__global__ void __launch_bounds__(256) k_triangle_prep(const float4* tp, uint32_t T, float4* tri_sphere, float4* tri_mt) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        v3 a = xyz(tp[3 * (size_t)t]), b = xyz(tp[3 * (size_t)t + 1]), c = xyz(tp[3 * (size_t)t + 2]);
        tri_sphere[t] = triangle_sphere(a, b, c);
        tri_mt[3 * (size_t)t] = f4(a, 0.0f);
        tri_mt[3 * (size_t)t + 1] = f4(b - a, 0.0f);
        tri_mt[3 * (size_t)t + 2] = f4(c - a, 0.0f);
    }
}

// Cluster spheres: one sphere over the triangle spheres of up to 64
// consecutive triangles of one object (range[i] = first, last).  Center =
// middle of the bounding box of the triangle spheres; radius = the largest
// |c_i - q| + r_i, inflated by 1e-5 relative so that FP32 rounding cannot
// leave a triangle sphere outside.
__global__ void __launch_bounds__(256) k_cluster_build(const float4* tri_sphere, const uint2* range, uint32_t n,
                                                       float4* cl_sphere) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint2 r = range[i];
        v3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t t = r.x; t <= r.y; t++) {
            const float4 s = tri_sphere[t];
            lo = mk(fminf(lo.x, s.x - s.w), fminf(lo.y, s.y - s.w), fminf(lo.z, s.z - s.w));
            hi = mk(fmaxf(hi.x, s.x + s.w), fmaxf(hi.y, s.y + s.w), fmaxf(hi.z, s.z + s.w));
        }
        const v3 q = (lo + hi) * 0.5f;
        float rho = 0.0f;
        bool finite = true;
        for (uint32_t t = r.x; t <= r.y; t++) {
            const float4 s = tri_sphere[t];
            const float d = len(xyz(s) - q) + s.w;
            finite = finite && isfinite(d);
            rho = fmaxf(rho, d);
        }
        // A non-finite member makes the cluster sphere NaN: the pre-test then passes.
        cl_sphere[i] = finite ? f4(q, rho * 1.00001f + 1e-6f) : make_float4(NAN, NAN, NAN, NAN);
    }
}

// Sub-clusters worth a ray mask: bit k of sub_use[c] is set when sub-cluster k
// of cluster c has a radius below SUB_RATIO of the cluster's (a smaller sphere
// meets fewer of the rays that meet the cluster's).
__global__ void __launch_bounds__(256) k_sub_use(const float4* cl_sphere, const float4* sub_sphere, uint32_t nc,
                                                 uint8_t* sub_use) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const float rc = cl_sphere[c].w;
        uint32_t bits = 0;
        for (uint32_t k = 0; k < SUB_PER_CLUSTER; k++)
            bits |= sub_sphere[(size_t)c * SUB_PER_CLUSTER + k].w < SUB_RATIO * rc ? 1u << k : 0u;
        sub_use[c] = (uint8_t)bits;
    }
}

// RayTracer.cu:448-511
__global__ void __launch_bounds__(256) k_update_vertex(const float* model, const float* normal, uint32_t n,
                                                       const float4* pos_in, const float4* nrm_in,
                                                       const int32_t* obj, float4* pos_out, float4* nrm_out) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x) {
        uint32_t m = (uint32_t)obj[x];
        const float* M = model + 16 * (size_t)m;
        float4 v = pos_in[x];
        float u[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            u[i] = 0.0f;
            u[i] += M[i * 4 + 0] * v.x; u[i] += M[i * 4 + 1] * v.y;
            u[i] += M[i * 4 + 2] * v.z; u[i] += M[i * 4 + 3] * v.w;
        }
        pos_out[x] = make_float4(u[0], u[1], u[2], (float)m);
        const float* Nm = normal + 16 * (size_t)m;
        float4 q = nrm_in[x];
        float w[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            w[i] = 0.0f;
            w[i] += Nm[i * 4 + 0] * q.x; w[i] += Nm[i * 4 + 1] * q.y;
            w[i] += Nm[i * 4 + 2] * q.z; w[i] += Nm[i * 4 + 3] * q.w;
        }
        nrm_out[x] = f4(normalize(mk(w[0], w[1], w[2])), 0.0f);
    }
}

// RayTracer.cu:513-570
__global__ void __launch_bounds__(256) k_update_sphere(const float* translation, const float* scale, uint32_t K,
                                                       const float4* sin4, float* sout3) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < K; x += gridDim.x * blockDim.x) {
        const float* Tm = translation + 16 * (size_t)x;
        const float* Sm = scale + 16 * (size_t)x;
        float4 c = sin4[2 * (size_t)x], rb = sin4[2 * (size_t)x + 1];
        float u[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            u[i] = 0.0f;
            u[i] += Tm[i * 4 + 0] * c.x; u[i] += Tm[i * 4 + 1] * c.y;
            u[i] += Tm[i * 4 + 2] * c.z; u[i] += Tm[i * 4 + 3] * 1.0f;
        }
        float ms = 1.17549435e-38f;
        for (int i = 0; i < 3; i++)
            if (fabsf(Sm[i * 4 + i]) > ms) ms = fabsf(Sm[i * 4 + i]);
        float* o = sout3 + 6 * (size_t)x;
        o[0] = u[0]; o[1] = u[1]; o[2] = u[2];
        o[3] = rb.x * ms; o[4] = rb.y; o[5] = rb.z;
    }
}

// RayTracer.cu:1776-2017 (reference 32-bit records, recomputed hit, optional
// in-place creation of the next bounce).
__global__ void __launch_bounds__(256) k_color_reflection_ref(const float4* tp, const float4* tn, uint32_t W, uint32_t H,
                                                              uint32_t Lmax, SceneView sc, v3 eye, uint32_t create,
                                                              const uint32_t* times, float* rays, uint32_t* flags,
                                                              uint32_t* keys, uint32_t* values, uint32_t* pixels) {
    const uint32_t n = W * H;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        uint32_t x = p % W, y = p / W;
        uint32_t r = x + y * W * Lmax;
        uint32_t rec = times[r], t = rec & 0x000FFFFFu;
        if (rec == MISS32) { if (create) flags[r] = 1; continue; }
        int32_t m = sc.tri_mat[3 * (size_t)t];
        v3 v0 = xyz(tp[3 * (size_t)t]), v1 = xyz(tp[3 * (size_t)t + 1]), v2 = xyz(tp[3 * (size_t)t + 2]);
        v3 o, d;
        load_ray(rays, r, o, d);
        float th = ray_triangle(o + d * EPS, d, v0, v1 - v0, v2 - v0);
        v3 P = o + d * (EPS + th);
        float aABC = len(cross(v1 - v0, v2 - v0));
        float aPBC = len(cross(v1 - P, v2 - P));
        float aPCA = len(cross(v0 - P, v2 - P));
        v3 N = ((aPBC / aABC) * xyz(tn[3 * (size_t)t]) + (aPCA / aABC) * xyz(tn[3 * (size_t)t + 1])) +
               ((1.0f - (aPBC / aABC)) - (aPCA / aABC)) * xyz(tn[3 * (size_t)t + 2]);
        float4 Kd = sc.mat_dif[m], Ks = sc.mat_spec[m];
        v3 color = mk(0.0f, 0.0f, 0.0f);
        for (uint32_t l = 0; l < sc.lights.total; l++) color = color + light_term(sc.lights, l, P, N, eye, Kd, Ks, true);
        pixels[p] += rgb_to_int(color.x * Ks.w, color.y * Ks.w, color.z * Ks.w);
        if (create) {
            v3 nd = reflect(d, N);
            store_ray(rays, r, P, nd);
            keys[r] = reflection_key_ref(P, nd);
            values[r] = r;
            flags[r] = 0;
        }
    }
}

}  // namespace

template <typename F>
void launch_query_keys(const float* rays, const uint32_t* segs, uint32_t n, KeyCfg kc, uint32_t type, F* flags,
                       uint32_t* keys, uint32_t* values, hipStream_t s) {
    if (n) k_query_keys<<<grid_for(n, 256, 65536), 256, 0, s>>>(rays, segs, n, kc, type, flags, keys, values);
}
template <typename F>
void launch_create_primary(const Views& V, uint32_t P, KeyCfg kc, float* rays, F* flags, uint32_t* keys,
                           uint32_t* values, hipStream_t s) {
    k_create_primary<<<grid_for(P, 256, 65536), 256, 0, s>>>(V, P, kc, rays, flags, keys, values);
}
void launch_create_primary_sorted(const Views& V, uint32_t P, float4* rec0, float4* rec1, hipStream_t s) {
    k_create_primary_sorted<<<grid_for(P, 256, 65536), 256, 0, s>>>(V, P, rec0, rec1);
}
template <typename F>
void launch_create_shadow(const float4* gpos, const float4* gnrm, uint32_t W, uint32_t H, uint32_t Lmax,
                          const float4* light_pos, KeyCfg kc, const Views& V, uint32_t soft, float* rays, F* flags,
                          uint32_t* keys, uint32_t* values, hipStream_t s) {
    k_create_shadow<<<grid_for((uint64_t)W * H * Lmax << (soft ? 4 : 0), 256, 65536), 256, 0, s>>>(
        gpos, gnrm, W, H, Lmax, light_pos, kc, V, soft, rays, flags, keys, values, fast_div(W), fast_div(Lmax));
}
template <typename F>
void launch_create_reflection(const float4* gpos, const float4* gnrm, const float4* gdif, uint32_t W, uint32_t H,
                              uint32_t Lmax, const Views& V, uint32_t gate, uint32_t mark_gaps, KeyCfg kc,
                              float* rays, F* flags, uint32_t* keys, uint32_t* values, hipStream_t s) {
    k_create_reflection<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(gpos, gnrm, gdif, W, H, Lmax, V, gate,
                                                                               mark_gaps, kc, rays, flags, keys, values,
                                                                               fast_div(W));
}
void launch_create_refraction(const float4* gpos, const float4* gnrm, uint32_t W, uint32_t H, uint32_t Lmax,
                              const Views& V, KeyCfg kc, float* rays, uint32_t* flags, uint32_t* keys,
                              uint32_t* values, hipStream_t s) {
    k_create_refraction<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(gpos, gnrm, W, H, Lmax, V, kc, rays,
                                                                               flags, keys, values);
}
template <typename F>
void launch_create_bounce(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, const unsigned long long* rec,
                          uint32_t hit_mode, uint32_t gate, KeyCfg kc, const Views& V, float* rays, F* flags,
                          uint32_t* keys, uint32_t* values, hipStream_t s) {
    k_create_bounce<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(g, W, H, Lmax, rec, hit_mode, gate, kc, V,
                                                                           rays, flags, keys, values, fast_div(W));
}
// the engine's byte flags and the stage API's 32-bit ones
#define RSH_FLAG_LAUNCHERS(F)                                                                                        \
    template void launch_query_keys<F>(const float*, const uint32_t*, uint32_t, KeyCfg, uint32_t, F*, uint32_t*,      \
                                       uint32_t*, hipStream_t);                                                      \
    template void launch_create_primary<F>(const Views&, uint32_t, KeyCfg, float*, F*, uint32_t*, uint32_t*,         \
                                           hipStream_t);                                                             \
    template void launch_create_shadow<F>(const float4*, const float4*, uint32_t, uint32_t, uint32_t, const float4*, \
                                          KeyCfg, const Views&, uint32_t, float*, F*, uint32_t*, uint32_t*,          \
                                          hipStream_t);                                                              \
    template void launch_create_reflection<F>(const float4*, const float4*, const float4*, uint32_t, uint32_t,       \
                                              uint32_t, const Views&, uint32_t, uint32_t, KeyCfg, float*, F*,        \
                                              uint32_t*, uint32_t*, hipStream_t);                                    \
    template void launch_create_bounce<F>(GBufView, uint32_t, uint32_t, uint32_t, const unsigned long long*,         \
                                          uint32_t, uint32_t, KeyCfg, const Views&, float*, F*, uint32_t*,           \
                                          uint32_t*, hipStream_t);
RSH_FLAG_LAUNCHERS(uint8_t)
RSH_FLAG_LAUNCHERS(uint32_t)
#undef RSH_FLAG_LAUNCHERS
void launch_surface(const float* rays, const unsigned long long* rec, uint32_t hit_mode, uint32_t W, uint32_t H,
                    uint32_t slot_stride, uint32_t normalize_n, SceneView sc, GBufView g, hipStream_t s,
                    const Views* V) {
    k_surface<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(rays, rec, hit_mode, W, H, slot_stride,
                                                                     normalize_n, sc, g, V ? *V : Views{},
                                                                     fast_div(W));
}
void launch_shade_primary(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, Lights L, const Views& V,
                          const uint32_t* sflags, uint32_t soft, uint32_t* pixels, hipStream_t s,
                          const uint32_t* live) {
    k_shade_primary<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(g, W, H, Lmax, L, V, sflags, soft, pixels,
                                                                           live, fast_div(W));
}
void launch_shade_reflection(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, Lights L, const Views& V,
                             const unsigned long long* rec, uint32_t hit_mode, const uint32_t* sflags,
                             uint32_t soft, uint32_t* pixels, hipStream_t s) {
    k_shade_reflection<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(g, W, H, Lmax, L, V, rec, hit_mode,
                                                                              sflags, soft, pixels, fast_div(W));
}
void launch_cluster_build(const float4* tri_sphere, const uint2* range, uint32_t n, float4* cl_sphere,
                          hipStream_t s) {
    if (n) k_cluster_build<<<grid_for(n, 256), 256, 0, s>>>(tri_sphere, range, n, cl_sphere);
}

void launch_sub_use(const float4* cl_sphere, const float4* sub_sphere, uint32_t nc, uint8_t* sub_use, hipStream_t s) {
    if (nc) k_sub_use<<<grid_for(nc, 256), 256, 0, s>>>(cl_sphere, sub_sphere, nc, sub_use);
}

void launch_triangle_prep(const float4* tri_pos, uint32_t T, float4* tri_sphere, float4* tri_mt, hipStream_t s) {
    if (T) k_triangle_prep<<<grid_for(T, 256), 256, 0, s>>>(tri_pos, T, tri_sphere, tri_mt);
}
void launch_update_vertex(const float* model, const float* normal, uint32_t n, const float4* pos_in,
                          const float4* nrm_in, const int32_t* obj_ids, float4* pos_out, float4* nrm_out,
                          hipStream_t s) {
    if (n) k_update_vertex<<<grid_for(n, 256), 256, 0, s>>>(model, normal, n, pos_in, nrm_in, obj_ids, pos_out, nrm_out);
}
void launch_update_sphere(const float* translation, const float* scale, uint32_t K, const float4* spheres_in,
                          float* spheres_out, hipStream_t s) {
    if (K) k_update_sphere<<<grid_for(K, 256), 256, 0, s>>>(translation, scale, K, spheres_in, spheres_out);
}
void launch_color_reflection_ref(const float4* tp, const float4* tn, uint32_t W, uint32_t H, uint32_t Lmax,
                                 SceneView sc, v3 eye, uint32_t create, const uint32_t* times, float* rays,
                                 uint32_t* flags, uint32_t* keys, uint32_t* values, uint32_t* pixels, hipStream_t s) {
    k_color_reflection_ref<<<grid_for((uint64_t)W * H, 256, 65536), 256, 0, s>>>(tp, tn, W, H, Lmax, sc, eye, create,
                                                                                  times, rays, flags, keys, values,
                                                                                  pixels);
}

}  // namespace rsh
