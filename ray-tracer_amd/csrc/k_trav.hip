// k_trav.hip -- hierarchy construction and traversal + intersection (gfx950).
//
// Reference: projects/CUDA-Engine/source/RayTracer.cu
//   CreateHierarchyLevel0 :1061-1108, CreateHierarchyLevelN :1110-1150
//   CalculateBoundingSpheresIntersections :1152-1280, CreateHierarchyBoundingSphereHits :1282-1313
//   CreateCRSH/RAHHierarchyLevel0Hits :1315-1405, CreateHierarchyLevelNHits :1407-1462
//   CreateTrimmedHierarchyHits :1464-1496
//   CalculateShadowRayIntersections :1498-1614, CalculateReflectionRayIntersections :1711-1774
//
// The native path fuses traversal and intersection: one wave64 owns the 64
// sorted rays under one top-level node (8 leaves of 8 rays for the reference
// fan-out), keeps each lane's ray and best hit in registers, and streams the
// node's candidate triangles -- the triangles of the objects whose spheres
// pass the node -- 64 at a time against the top node, then every survivor
// against the lane's own chain of lower nodes before the Moller-Trumbore
// test.  No (node, triangle) pair is materialised.  A top node's candidate
// stream is cut into equal chunks (work items, one wave each) so that nodes
// with a wide cone over large objects do not serialise the pass; waves of
// the same node merge closest hits with a 64-bit atomicMin.  The candidate
// set per (leaf, triangle) is the reference's; only the order of tests
// differs, and both any-hit and closest-hit are order independent.
#include <type_traits>

#include "rsh_internal.hpp"
#include "rsh_kernels.hpp"

namespace rsh {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 1u << 20) {
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return (uint32_t)(g < cap ? g : cap);
}

// -------------------------------------------------------------- gather ----
// Sorted ray records: rec0 = (origin, slot bits), rec1 = (direction, tmax).
// tmax = |o - P| - 2 eps for shadow rays (RayTracer.cu:1610), P the surface
// point of the slot's pixel, or the caller's per-slot tmax for ray queries;
// unused for closest-hit passes.  slot_shift: 4 for soft shadows (16 rays per
// pixel-light slot), else 0.
// rays == nullptr (a frame's shadow pass): each ray is recomputed from its
// slot's surface point and light as k_create_shadow made it (shadow_ray).
__global__ void __launch_bounds__(256) k_gather(const float* rays, const uint32_t* sorted_values, uint32_t N,
                                                const PassSizes* ps, const float4* shadow_pos, const float* tmax_in,
                                                const float4* light_pos, uint32_t W, uint32_t Lmax,
                                                uint32_t slot_shift, float4* rec0, float4* rec1, FastDiv fw,
                                                FastDiv fl) {
    if (ps) N = min(N, ps->N);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        uint32_t slot = sorted_values[p];
        v3 o, d;
        v3 P = mk(0.0f, 0.0f, 0.0f);
        if (shadow_pos) {
            const uint32_t i = slot >> slot_shift;
            const uint32_t qw = fdiv(i, fw);              // l + Lmax y
            const uint32_t x = i - qw * W, y = fdiv(qw, fl);
            P = xyz(shadow_pos[(size_t)x + (size_t)y * W]);
            if (!rays) shadow_ray(P, xyz(light_pos[qw - y * Lmax]), slot_shift != 0u, slot & 15u, o, d);
        }
        if (rays) {
            const float2* r = reinterpret_cast<const float2*>(rays + 6 * (size_t)slot);
            float2 a = r[0], b = r[1], c = r[2];
            o = mk(a.x, a.y, b.x); d = mk(b.y, c.x, c.y);
        }
        float tmax = 0.0f;
        if (tmax_in) {
            tmax = tmax_in[slot];
        } else if (shadow_pos) {
            tmax = len(o - P) - EPS * 2.0f;
        }
        rec0[p] = f4(o, __uint_as_float(slot));
        rec1[p] = f4(d, tmax);
    }
}

// ----------------------------------------------------------- hierarchy ----
// Level 0: node j over sorted rays [j*leaf, min((j+1)*leaf, N)).  The
// reference's loop bound reads one slot past the last ray (RayTracer.cu:1090);
// here the node stops at the last live ray.
template <bool FROM_RECORDS>
__global__ void __launch_bounds__(256) k_level0(const float4* rec0, const float4* rec1, const float* rays,
                                                const uint32_t* sorted_values, uint32_t N, const PassSizes* ps,
                                                uint32_t leaf, float r0, float s0, uint32_t mode, float4* nodes,
                                                float2* aux) {
    if (ps) N = min(N, ps->N);
    const uint32_t n0 = (N + leaf - 1) / leaf;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n0; j += gridDim.x * blockDim.x) {
        auto fetch = [&](uint32_t p, v3& o, v3& d) {
            if (FROM_RECORDS) {
                float4 a = rec0[p], b = rec1[p];
                o = xyz(a); d = xyz(b);
            } else {
                const float* r = rays + 6 * (size_t)sorted_values[p];
                o = mk(r[0], r[1], r[2]); d = mk(r[3], r[4], r[5]);
            }
        };
        v3 o, d;
        float4 sphere, cone;
        if (FROM_RECORDS && leaf == 8) {
            // the eight records' loads issued together, then the same fold
            float4 a[8], b[8];
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                const uint32_t p = j * 8 + i;
                a[i] = p < N ? rec0[p] : make_float4(0, 0, 0, 0);
                b[i] = p < N ? rec1[p] : make_float4(0, 0, 0, 0);
            }
            sphere = f4(xyz(a[0]), r0);
            cone = f4(mode == 0 ? xyz(b[0]) : normalize(xyz(b[0])), s0);
#pragma unroll
            for (uint32_t i = 1; i < 8; i++) {
                if (j * 8 + i >= N) break;
                sphere = sphere_union(sphere, f4(xyz(a[i]), r0));
                cone = cone0(cone, xyz(b[i]), s0, mode);
            }
        } else {
            fetch(j * leaf, o, d);
            sphere = f4(o, r0);
            cone = f4(mode == 0 ? d : normalize(d), s0);
            for (uint32_t i = 1; i < leaf; i++) {
                uint32_t p = j * leaf + i;
                if (p >= N) break;
                fetch(p, o, d);
                sphere = sphere_union(sphere, f4(o, r0));
                cone = cone0(cone, d, s0, mode);
            }
        }
        nodes[2 * (size_t)j] = sphere;
        nodes[2 * (size_t)j + 1] = cone;
        if (aux) aux[j] = make_float2(cos_s(cone.w), tan_s(cone.w));
    }
}

__global__ void __launch_bounds__(256) k_levelN(float4* nodes, float2* aux, uint32_t read_off, uint32_t read_count,
                                                uint32_t write_off, uint32_t write_count, uint32_t branch,
                                                uint32_t mode, const PassSizes* ps, uint32_t level) {
    if (ps) {       // device-sized: level `level` over level - 1
        read_off = ps->offset[level - 1]; read_count = ps->count[level - 1];
        write_off = ps->offset[level]; write_count = ps->count[level];
    }
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < write_count; j += gridDim.x * blockDim.x) {
        size_t c0 = read_off + (size_t)j * branch;
        float4 sphere = nodes[2 * c0], cone = nodes[2 * c0 + 1];
        for (uint32_t i = 1; i < branch; i++) {
            if (j * branch + i >= read_count) break;
            size_t c = c0 + i;
            sphere = sphere_union(sphere, nodes[2 * c]);
            cone = coneN(cone, nodes[2 * c + 1], mode);
        }
        nodes[2 * (size_t)(write_off + j)] = sphere;
        nodes[2 * (size_t)(write_off + j) + 1] = cone;
        if (aux) aux[write_off + j] = make_float2(cos_s(cone.w), tan_s(cone.w));
    }
}

// ---------------------------------------------------- fused traversal ----
// Device-sized passes: the pass's N and level sizes from device memory into
// the kernel's copy of the parameters.
__device__ __forceinline__ void apply_sizes(TravParams& P) {
    if (!P.ps) return;
    const PassSizes q = *P.ps;
    P.N = q.N;
#pragma unroll
    for (int k = 0; k < 4; k++) { P.level_count[k] = q.count[k]; P.level_offset[k] = q.offset[k]; }
}

// Object whose triangle range contains t (objects tile [0, T) in order);
// wave-uniform binary search over the first-triangle column.
__device__ __forceinline__ uint32_t object_of(const uint4* ranges, uint32_t K, uint32_t t) {
    uint32_t a = 0, b = K;      // ranges[a].x <= t < ranges[b].x (b = K: end)
    while (b - a > 1) {
        uint32_t m = (a + b) >> 1;
        if (ranges[m].x <= t) a = m; else b = m;
    }
    return a;
}

template <int DEPTH>
__global__ void __launch_bounds__(256) k_traverse(TravParams P) {
    const uint32_t lane = lane_id();
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t per_top = P.waves_per_top * P.splits;
    const uint32_t top = wave / per_top;
    const uint32_t rem = wave - top * per_top;
    const uint32_t sub = rem / P.splits;
    const uint32_t split = rem - sub * P.splits;
    const uint32_t topLevel = DEPTH - 1;
    if (top >= P.level_count[topLevel]) return;
    const uint32_t local = sub * 64 + lane;
    const uint32_t pos = top * P.rays_per_top + local;
    const bool valid = local < P.rays_per_top && pos < P.N;
    if (__ballot(valid) == 0ull) return;

    v3 o = mk(0, 0, 0), d = mk(0, 0, 0);
    uint32_t slot = 0;
    float tmax = 0.0f;
    if (valid) {
        float4 a = P.rec0[pos], b = P.rec1[pos];
        o = xyz(a); d = xyz(b);
        slot = __float_as_uint(a.w);
        tmax = b.w;
    }
    const v3 oe = o + d * EPS;

    // top node (wave-uniform)
    const uint32_t gt = P.level_offset[topLevel] + top;
    const float4 tS = P.nodes[2 * (size_t)gt], tC = P.nodes[2 * (size_t)gt + 1];
    const float2 tA = P.aux[gt];

    // per-lane chain of lower nodes, level DEPTH-2 down to 0
    float4 cS[DEPTH > 1 ? DEPTH - 1 : 1], cC[DEPTH > 1 ? DEPTH - 1 : 1];
    float2 cA[DEPTH > 1 ? DEPTH - 1 : 1];
    bool leaf_head = false;
    if (DEPTH > 1) {
        uint32_t span = P.leaf;
        uint32_t idx0 = valid ? pos / P.leaf : 0;
        leaf_head = valid && (pos % P.leaf) == 0;
#pragma unroll
        for (int k = 0; k < DEPTH - 1; k++) {
            uint32_t idx = valid ? pos / span : 0;
            (void)idx0;
            uint32_t g = P.level_offset[k] + idx;
            cS[k] = P.nodes[2 * (size_t)g];
            cC[k] = P.nodes[2 * (size_t)g + 1];
            cA[k] = P.aux[g];
            span *= P.branch;
        }
    } else {
        leaf_head = valid && (pos % P.leaf) == 0;
    }

    unsigned long long best = P.hit_mode == 0 ? MISS64 : (unsigned long long)MISS32;
    bool occluded = false;
    uint64_t alive = __ballot(valid);
    const bool count = P.counters != nullptr;
    uint64_t n_top = 0, n_leaf = 0, n_tests = 0, n_surv = 0;
    bool done = false;

    auto process_range = [&](uint32_t first, uint32_t last, uint32_t exlo, uint32_t exhi) {
        for (uint32_t tb = first; tb <= last && !done; tb += 64) {
            const uint32_t t = tb + lane;
            const bool inr = t <= last && !(t >= exlo && t < exhi);
            bool p = false;
            if (inr) p = sphere_node(tS, tC, P.tri_sphere[t], tA.x, tA.y);
            uint64_t m = __ballot(p);
            if (count && sub == 0) {
                n_top += (uint64_t)__popcll(__ballot(inr));
                n_surv += (uint64_t)__popcll(m);
            }
            while (m) {
                const uint32_t tt = __builtin_amdgcn_readfirstlane(tb + (uint32_t)__builtin_ctzll(m));
                m &= m - 1;
                const float4 ts = P.tri_sphere[tt];
                bool lp = valid && !occluded;
                if (DEPTH > 1) {
#pragma unroll
                    for (int k = DEPTH - 2; k >= 0; k--) lp = lp && sphere_node(cS[k], cC[k], ts, cA[k].x, cA[k].y);
                }
                const uint64_t lm = __ballot(lp);
                if (count) {
                    n_leaf += (uint64_t)__popcll(__ballot(lp && leaf_head));
                    n_tests += (uint64_t)__popcll(lm);
                }
                if (lm == 0ull) continue;
                const float4 A = P.tri_mt[3 * (size_t)tt], B = P.tri_mt[3 * (size_t)tt + 1], Cc = P.tri_mt[3 * (size_t)tt + 2];
                if (lp) {
                    float th = ray_triangle(oe, d, xyz(A), xyz(B), xyz(Cc));
                    if (P.kind == 1) {
                        if (th > EPS && th < tmax) occluded = true;
                    } else if (th > EPS) {
                        unsigned long long rec = pack_hit(th, tt, P.hit_mode);
                        best = rec < best ? rec : best;
                    }
                }
                if (P.kind == 1 && (__ballot(occluded) & alive) == alive) { done = true; break; }
            }
        }
    };

    const uint32_t lo = (uint32_t)(((uint64_t)split * P.T) / P.splits);
    const uint32_t hi = (uint32_t)(((uint64_t)(split + 1) * P.T) / P.splits);
    if (hi > lo) {
        if (P.algorithm == 1) {
            process_range(lo, hi - 1, 0, 0);
        } else {
            const uint32_t k0 = object_of(P.obj_range, P.K, lo);
            const uint32_t k1 = object_of(P.obj_range, P.K, hi - 1);
            for (uint32_t ob = k0; ob <= k1 && !done; ob += 64) {
                const uint32_t oi = ob + lane;
                bool po = false;
                if (oi <= k1) po = sphere_node(tS, tC, P.obj_sphere[oi], tA.x, tA.y);
                uint64_t m = __ballot(po);
                while (m && !done) {
                    const uint32_t ok = __builtin_amdgcn_readfirstlane(ob + (uint32_t)__builtin_ctzll(m));
                    m &= m - 1;
                    const uint4 r = P.obj_range[ok];
                    process_range(r.x > lo ? r.x : lo, r.y < hi - 1 ? r.y : hi - 1, r.z, r.w);
                }
            }
        }
    }

    if (valid) {
        if (P.kind == 1) { if (occluded) P.shadow[slot] = MISS32; }
        else if (P.splits == 1) P.closest[slot] = best;
        else if (best != (P.hit_mode == 0 ? MISS64 : (unsigned long long)MISS32)) atomicMin(&P.closest[slot], best);
    }
    if (count) {
        if (lane == 0) {
            unsigned long long* cs = P.counters + (size_t)(wave % COUNTER_SLOTS) * COUNTERS_PER_SLOT;
            atomicAdd(&cs[CNT_TOP_PAIRS], (unsigned long long)n_top);
            atomicAdd(&cs[CNT_TRI_TESTS], (unsigned long long)n_top);
            atomicAdd(&cs[CNT_LEAF_PAIRS], (unsigned long long)n_leaf);
            atomicAdd(&cs[CNT_RAY_TESTS], (unsigned long long)n_tests);
            atomicAdd(&cs[CNT_TOP_SURV], (unsigned long long)n_surv);
        }
    }
}

// ------------------------------------------- top-level object culling ------
// One wave per top node (depth 2, 64 rays): the object-sphere tests of
// RayTracer.cu:1152-1280, once per node.  Passing objects are listed in
// triangle order with the number of candidate clusters (CLUSTER consecutive
// triangles of one object) before each, so the node's candidates form one
// stream of clusters that is cut into equal chunks (work items) for the
// traversal kernel.  The count pass sizes the compact lists (npass, cand per
// node; entry n_top is written as 0 so an exclusive scan over n_top + 1
// entries ends with the total), the list pass fills them.  RAH (no object
// culling) streams every cluster through one shared list of all objects.
// With counters on, the count pass adds the reference's candidate pair count
// (triangles of passing objects outside the excluded batch) per node.
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x = fmaxf(x, __shfl_xor(x, d, 64));
    return x;
}

// Object j of the walk order (obj_order, or index order), j < K.
__device__ __forceinline__ uint32_t object_at(const TravParams& P, uint32_t j) {
    return P.obj_order ? P.obj_order[j] : j;
}

// The counting pass stores each node's passing objects as 64-bit masks
// (pmask, ceil(K / 64) words per node); the list pass reads them back instead
// of repeating the sphere tests.
template <bool LIST>
__global__ void __launch_bounds__(64) k_cull(TravParams P, uint32_t* npass, uint32_t* cand, uint4* obj_ent,
                                             uint32_t* obj_cum, unsigned long long* pmask) {
    apply_sizes(P);
    const uint32_t lane = lane_id();
    const uint32_t top = blockIdx.x;
    const uint32_t ntop = P.level_count[1];
    if (top > ntop || (LIST && top == ntop)) return;
    if (top == ntop) {
        if (lane == 0) { npass[top] = 0; cand[top] = 0; }
        return;
    }
    unsigned long long* cs =
        P.counters ? P.counters + (size_t)(top % COUNTER_SLOTS) * COUNTERS_PER_SLOT : nullptr;
    if (P.algorithm == 1) {
        if (!LIST && lane == 0) {
            npass[top] = P.K; cand[top] = P.n_clusters;
            if (cs) atomicAdd(&cs[CNT_TOP_PAIRS], (unsigned long long)P.T);
        }
        return;
    }
    const uint32_t gt = P.level_offset[1] + top;
    const float4 tS = P.nodes[2 * (size_t)gt], tC = P.nodes[2 * (size_t)gt + 1];
    const float2 tA = P.aux[gt];
    const size_t base = LIST ? P.obj_off[top] : 0;
    const uint32_t words = (P.K + 63) / 64;
    unsigned long long* pm = pmask + (size_t)top * words;
    const float rcos = node_rcos(tA.x);
    uint32_t n = 0, cum = 0;
    if (LIST) {
        // only the words of the count pass's mask with a passing object
        for (uint32_t ob = 0; ob < P.K; ob += 64) {
            const uint64_t m = pm[ob / 64];
            if (m == 0ull) continue;
            const bool po = (m >> lane) & 1ull;
            const uint32_t oi = po ? object_at(P, ob + lane) : 0u;
            uint4 r = make_uint4(0u, 0u, 0u, 0u);
            uint32_t ncl = 0;
            if (po) {
                r = P.obj_range[oi];
                ncl = (r.y - r.x + CLUSTER) / CLUSTER;
            }
            const uint32_t incl = wave_incl_add(ncl);
            if (po) {
                const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                obj_ent[base + n + rank] = make_uint4(r.x, r.y, P.obj_cl_base[oi], r.z);
                obj_cum[base + n + rank] = cum + incl - ncl;
            }
            n += (uint32_t)__popcll(m);
            cum += __shfl(incl, 63, 64);
        }
        return;
    }
    // count pass: per-lane sums, one reduction at the end
    unsigned long long pairs = 0;
    for (uint32_t ob = 0; ob < P.K; ob += 64) {
        const uint32_t oi = ob + lane < P.K ? object_at(P, ob + lane) : 0u;
        bool po = false;
        if (ob + lane < P.K) po = sphere_node_fast(tS, tC, P.obj_sphere[oi], tA.x, tA.y, rcos);
        if (po) {
            const uint4 r = P.obj_range[oi];
            cum += (r.y - r.x + CLUSTER) / CLUSTER;
            pairs += (r.y - r.x + 1) - (r.w - r.z);
        }
        const uint64_t m = __ballot(po);
        if (lane == 0) pm[ob / 64] = m;
        n += (uint32_t)__popcll(m);
    }
    for (int dlt = 32; dlt > 0; dlt >>= 1) cum += __shfl_xor(cum, dlt, 64);
    if (cs) {
        for (int dlt = 32; dlt > 0; dlt >>= 1) pairs += __shfl_xor(pairs, dlt, 64);
        if (lane == 0) atomicAdd(&cs[CNT_TOP_PAIRS], pairs);
    }
    if (lane == 0) { npass[top] = n; cand[top] = cum; }
}

// Work items per top node for a chunk size; entry n_top is 0.
__global__ void __launch_bounds__(256) k_item_count(const uint32_t* cand, uint32_t chunk, uint32_t ntop,
                                                    uint32_t* nitems) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > ntop) return;
    nitems[i] = i == ntop ? 0u : (uint32_t)(((uint64_t)cand[i] + chunk - 1) / chunk);
}

// Item totals per chunk size: table[c - 1] = sum over top nodes of ceil(cand /
// c) for c = 1..ITEM_TABLE, table[ITEM_TABLE + c - 1] the same for chunk 2c;
// table[2 ITEM_TABLE] = top nodes whose cone is wide (tangent of the spread
// outside [0, WIDE_TAN]: a spread clamped to pi/2 has a large negative float
// tangent; or not finite); table[2 ITEM_TABLE + 1] = the sum of npass (the
// pass's list entries; npass may be null), so one read-back carries
// everything.  The table must start zeroed (the item scan clears it for the
// next pass).  One wave per slice of TABLE_SLICE top nodes,
// lane c - 1 sums chunk sizes c and 2c in registers; one read-back of the
// table sizes a pass and picks its traversal mode.
constexpr uint32_t TABLE_SLICE = 256;
// host (if set, device-visible coherent host memory): the block that finishes
// last (counted in *done, which it resets) copies the table to host[1..] and
// then writes `seq` to host[0], so the host reads the table without a copy.
__global__ void __launch_bounds__(64) k_item_table(const uint32_t* cand, const float2* top_aux, uint32_t ntop,
                                                   unsigned long long* table, bool fast, const uint32_t* npass,
                                                   uint32_t* done, unsigned long long* host, uint32_t seq) {
    __shared__ uint4 s_cand4[TABLE_SLICE / 4];
    uint32_t* s_cand = reinterpret_cast<uint32_t*>(s_cand4);
    const uint32_t lane = threadIdx.x & 63u, c = lane + 1u;
    const uint32_t t0 = blockIdx.x * TABLE_SLICE;
    const uint32_t t1 = t0 + TABLE_SLICE < ntop ? t0 + TABLE_SLICE : ntop;
    for (uint32_t t = lane; t < TABLE_SLICE; t += 64u) s_cand[t] = t0 + t < t1 ? cand[t0 + t] : 0u;   // coalesced
    __builtin_amdgcn_wave_barrier();
    // four candidates per LDS read (the zero padding adds nothing)
    const uint32_t q1 = (t1 - t0 + 3u) / 4u;
    uint32_t lo = 0, lo2 = 0, lo4 = 0;         // per-read partial sums stay below 2^32
    unsigned long long sum = 0ull, sum2 = 0ull, sum4 = 0ull;
    if (fast) {       // every n + 4c - 1 within div_hi's bound for 4c <= 256
        const uint32_t m1 = div_hi_magic(c), m2 = div_hi_magic(2u * c), m4 = div_hi_magic(4u * c);
        auto f1 = [&](uint32_t n) { return div_hi(n + c - 1u, c, m1); };
        auto f2 = [&](uint32_t n) { return div_hi(n + 2u * c - 1u, 2u * c, m2); };
        auto f4 = [&](uint32_t n) { return div_hi(n + 4u * c - 1u, 4u * c, m4); };
#pragma unroll 4
        for (uint32_t q = 0; q < q1; q++) {
            const uint4 v = s_cand4[q];
            lo = f1(v.x) + f1(v.y);
            lo2 = f2(v.x) + f2(v.y);
            lo4 = f4(v.x) + f4(v.y);
            sum += lo; sum2 += lo2; sum4 += lo4;
            lo = f1(v.z) + f1(v.w);
            lo2 = f2(v.z) + f2(v.w);
            lo4 = f4(v.z) + f4(v.w);
            sum += lo; sum2 += lo2; sum4 += lo4;
        }
    } else {
        for (uint32_t t = 0; t < t1 - t0; t++) {
            const uint32_t n = s_cand[t];
            sum += (n + c - 1u) / c;
            sum2 += (n + 2u * c - 1u) / (2u * c);
            sum4 += (n + 4u * c - 1u) / (4u * c);
        }
    }
    if (sum) atomicAdd(&table[c - 1u], sum);
    if (sum2) atomicAdd(&table[ITEM_TABLE + c - 1u], sum2);
    if (sum4) atomicAdd(&table[2 * ITEM_TABLE + c - 1u], sum4);
    uint32_t wide = 0;
    for (uint32_t t = t0 + lane; t < t1; t += 64u) wide += !(top_aux[t].y >= 0.0f && top_aux[t].y <= WIDE_TAN) ? 1u : 0u;
    for (int d = 32; d > 0; d >>= 1) wide += __shfl_xor(wide, d, 64);
    if (lane == 0 && wide) atomicAdd(&table[TABLE_WIDE], (unsigned long long)wide);
    if (npass) {
        unsigned long long np = 0ull;
        for (uint32_t t = t0 + lane; t < t1; t += 64u) np += npass[t];
        for (int d = 32; d > 0; d >>= 1) np += __shfl_xor(np, d, 64);
        if (lane == 0 && np) atomicAdd(&table[TABLE_LIST], np);
    }
    if (host) {
        __threadfence();
        uint32_t prev = 0;
        if (lane == 0) prev = atomicAdd(done, 1u);
        prev = __shfl(prev, 0, 64);
        if (prev == gridDim.x - 1u) {
            __threadfence();
            for (uint32_t i = lane; i < ITEM_TABLE_WORDS; i += 64u)
                host[1 + i] = __hip_atomic_load(&table[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence_system();
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                *done = 0u;
                __hip_atomic_store(host, (unsigned long long)seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// Work-item descriptors, one wave per top node, one lane per item:
//   x top node, y absolute list entry holding the item's first cluster,
//   z first cluster within that entry, w clusters of the item | (1 << 31)
//   when the item is the node's only one (no merge needed) | (1 << 30) when
//   it is the node's first.
// The entry is found by binary search of the node's candidate offsets (RAH:
// the shared list of all objects with the cluster bases as offsets).
__global__ void __launch_bounds__(64) k_item_desc(TravParams P, uint4* desc) {
    const uint32_t top = blockIdx.x;
    if (top >= P.level_count[1]) return;
    const bool rah = P.algorithm == 1;
    const uint32_t f = P.item_first[top], k = P.item_first[top + 1] - f, ncand = P.cand[top];
    const uint32_t lbase = rah ? 0u : P.obj_off[top];
    const uint32_t np = rah ? P.K : P.obj_off[top + 1] - lbase;
    for (uint32_t c = lane_id(); c < k; c += 64) {
        const uint32_t s0 = c * P.chunk;
        const uint32_t rem = ncand - s0 < P.chunk ? ncand - s0 : P.chunk;
        uint32_t a = 0, b = np;       // last entry with offset <= s0: cum[a] <= s0 < cum[b]
        while (b - a > 1) {
            const uint32_t m = (a + b) >> 1;
            if (P.obj_cum[lbase + m] <= s0) a = m; else b = m;
        }
        desc[f + c] = make_uint4(top, lbase + a, s0 - P.obj_cum[lbase + a],
                                 rem | (k == 1 ? 0x80000000u : 0u) | (c == 0 ? 0x40000000u : 0u));
    }
}

// List pass and work-item descriptors in one (CRSH): one wave per top node
// walks the passing objects of its count-pass mask in order, 64 objects (one
// mask word) at a time, writes each as a list entry with the node's running
// cluster count before it (as k_cull<true>), and then the descriptors of the
// items whose first cluster s0 = i * chunk falls inside those objects'
// clusters, one item per lane: (top, entry, s0 - c0, clusters | single << 31
// | first << 30)
// for the object with the largest first cluster c0 <= s0 (a binary search
// over the word's objects in LDS) -- the entry k_item_desc finds by binary
// search.  Every listed object has at least one cluster, so each item's start
// lies in exactly one of them.  (Writing the descriptors per object instead
// left one lane writing all of a large object's items while the others
// idled.)
__global__ void __launch_bounds__(64) k_cull_items(TravParams P, uint4* obj_ent, uint32_t* obj_cum,
                                                   const unsigned long long* pmask, uint4* desc) {
    __shared__ uint32_t s_c0[1][64];   // first cluster of each passing object of the word
    __shared__ uint32_t s_e[1][64];    // and its list entry
    apply_sizes(P);
    const uint32_t lane = lane_id(), w = 0;
    const uint32_t top = blockIdx.x;
    if (top >= P.level_count[1]) return;
    // device-sized passes: a list or descriptor table over its capacity is
    // not written (the frame reports an overflow and is rendered again)
    if (P.ps && (P.obj_off[top + 1] > P.list_cap || P.item_first[top + 1] > P.items_cap)) {
        if (lane == 0) atomicOr(P.overflow, 1u);
        return;
    }
    const uint32_t words = (P.K + 63) / 64;
    const unsigned long long* pm = pmask + (size_t)top * words;
    const uint32_t base = P.obj_off[top];
    const uint32_t f = P.item_first[top], k = P.item_first[top + 1] - f, ncand = P.cand[top];
    const uint32_t chunk = P.chunk;
    const uint32_t single = k == 1 ? 0x80000000u : 0u;
    uint32_t n = 0, cum = 0;
    for (uint32_t ob = 0; ob < P.K; ob += 64) {
        const uint64_t m = pm[ob / 64];
        if (m == 0ull) continue;
        const bool po = (m >> lane) & 1ull;
        const uint32_t oi = po ? object_at(P, ob + lane) : 0u;
        uint4 r = make_uint4(0u, 0u, 0u, 0u);
        uint32_t ncl = 0;
        if (po) {
            r = P.obj_range[oi];
            ncl = (r.y - r.x + CLUSTER) / CLUSTER;
        }
        const uint32_t incl = wave_incl_add(ncl);
        if (po) {
            const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            const uint32_t e = base + n + rank;
            const uint32_t c0 = cum + incl - ncl;
            obj_ent[e] = make_uint4(r.x, r.y, P.obj_cl_base[oi], r.z);
            obj_cum[e] = c0;
            s_c0[w][rank] = c0;
            s_e[w][rank] = e;
        }
        const uint32_t cnt = (uint32_t)__popcll(m);
        const uint32_t cum_end = cum + __shfl(incl, 63, 64);
        __builtin_amdgcn_wave_barrier();
        const uint32_t i1 = (cum_end + chunk - 1u) / chunk;
        for (uint32_t i = (cum + chunk - 1u) / chunk + lane; i < i1; i += 64u) {
            const uint32_t s0 = i * chunk;
            uint32_t lo = 0, hi = cnt;          // s_c0[w][lo] <= s0 < s_c0[w][hi] (ascending)
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_c0[w][mid] <= s0) lo = mid; else hi = mid;
            }
            const uint32_t rem = ncand - s0 < chunk ? ncand - s0 : chunk;
            desc[f + i] = make_uint4(top, s_e[w][lo], s0 - s_c0[w][lo], rem | single | (i == 0 ? 0x40000000u : 0u));
        }
        __builtin_amdgcn_wave_barrier();      // the word's LDS entries are read before the next word's
        n += cnt;
        cum = cum_end;
    }
}

// ------------------------------------- fused traversal, 64-ray top nodes ----
// Specialisation for depth 2 with leaf * branch == 64: one workgroup = one
// wave = one work item = one chunk of one top node's candidate cluster stream
// (k_cull, k_item_desc); the top node covers BRANCH leaves of LEAF rays (lane
// i holds sorted ray i of the node).
//  A  The wave walks its chunk: 64 clusters (CLUSTER consecutive triangles of
//     a listed object) at a time are pre-tested, one per lane, with the
//     conservative cluster_may_pass; the triangles of each passing cluster
//     are tested against the top node 64 at a time, NB blocks' bounding
//     spheres loaded back to back.  Survivors (index, sphere and the mask of
//     leaves the cluster may pass) are appended to an LDS queue.  The
//     pre-tests only skip pairs whose own test would fail, so the survivors
//     and the passing (leaf, triangle) pairs are exactly the reference's.
//  B  Per round of up to SCH queued survivors, the (survivor, leaf)
//     candidates allowed by the masks are listed and tested 64 per step;
//     passing pairs are compacted in place.
//  C  The pair list is walked BRANCH pairs per step, lane = pair p0 + lane /
//     LEAF, ray lane % LEAF of the pair's leaf (MODE 2: one pair per lane,
//     its candidate rays listed one per round).  With the pre-test (MODE >= 1)
//     each ray's line is first tested against the survivor's bounding sphere
//     (ray_may_hit_sphere); the (survivor, ray) tests that may hit go to an LDS
//     ring and run the Moller-Trumbore test 64 at a time, the vectors loaded
//     from memory per test.  Without it, the vectors of the survivors in a pair
//     are staged in LDS and every (pair, ray) runs the exact test.  Closest
//     hits merge per ray in LDS (atomicMin), shadow rays set a flag.
// The set of (leaf, triangle) pairs and ray tests is the reference's; only
// the order differs, and any-hit / closest-hit are order independent.
// One wave per workgroup: items differ widely in length, and a workgroup's LDS
// is held until its last wave ends, so a finished item must not wait for
// others (measured: 2x faster than four-wave workgroups on the 1080p atrium).
// MODE 0 runs the exact test on every (pair, ray); MODE 1 first tests each
// ray's line against the triangle's bounding sphere (rejects 99% of the tests
// of the atrium's second bounce and about 60% of primary ones); MODE 2 also
// computes, per passing cluster, which of the 64 rays may meet the cluster's
// sphere and lists only those (pair, ray) candidates.  MODE 2 pays for
// spread-out secondary rays (2x on the atrium's bounces) and costs on coherent
// ones; the engine picks the mode per pass.  Modes 1 and 2 are built for six
// waves per SIMD (80 VGPRs, no spills, in the instances without counters;
// 5.6 / 6.5 KB of LDS with 64-entry survivor queues, one block of triangle
// spheres per stage-A step: 24 waves per CU of 160 KB; mode 2's 8.9 KB with a
// 128-entry queue allowed 17 per CU: 1080p atrium frame 5.63 -> 5.40 ms,
// 64-view batch 59.6 -> 57.4 ms when it went to 64 entries); mode 0, four.
// LDS is allocated in 512-byte granules.  The counting instances (CNT) carry
// the reference's pair and test counters; without them the others need 16
// fewer scalar and 4 fewer vector registers, which removed every VGPR spill
// (frame 5.22 -> 5.10 ms, batch 52.7 -> 50.9 ms) and let mode 2 reach six
// waves per SIMD too (its passes 2 % faster; with counters it spilled 10
// VGPRs there and lost).
// Occupancy targets (waves per SIMD) of the pre-test modes and of mode 0,
// the blocks of triangle spheres loaded back to back in stage A, and the two
// distance prunings (RSH_TRAV_TPRUNE: each ray's pre-test also bounds the
// triangle's parameter range by EPS and the ray's length or closest hit so
// far; RSH_TRAV_NPRUNE: mode 1 shadow items skip clusters and triangles
// beyond the reach of all their rays).
// Measured on the 1080p atrium frame and its 64-view batch (median frame /
// graph-replayed batch): 2 blocks 6.02 ms / 66.3 ms; 1 block 6.11 / 67.9;
// 3 blocks 6.09 / 66.9; 4 blocks 6.37 / 70.3; 6 blocks 6.84 / 76.6.  Four or
// six waves per SIMD for the pre-test modes measured slower than five while
// mode 1's 7.0 KB of LDS left room for only 22 waves per CU.  With its queue
// at 64 entries (5.6 KB) and one block per step, six waves per SIMD fit
// with 2 VGPRs spilled: camera pass 0.81 -> 0.72 ms, frame 5.43 -> 5.31 ms,
// batch 57.8 -> 55.1 ms (the 64-entry queue alone at five waves 0.78 ms; six
// waves with two blocks per step spill 25 VGPRs and lose the first shadow
// pass, 0.85 -> 0.95 ms; seven waves with one block spill 10, 5.33 ms).
// Pruning, same measures: neither 6.04 / 66.3; TPRUNE 5.90 / 65.1; with
// NPRUNE in every pass 6.05 / 65.2 (closest-hit passes lose: their reach is
// infinite until every ray of the item has a hit), in mode 1 shadow passes
// only 5.78 / 62.3.  With mode 2 at 20 waves per CU (64-entry survivor
// queue): one block per step in mode 2 5.35 / 56.8 against two 5.44 / 57.4
// (mode 1 keeps two); six waves per SIMD for mode 2 5.79 / 60.4; a
// sub-cluster threshold of 16 or 32 instead of 64 within 0.5 % (at six waves
// per SIMD with the block skips, 32 measured 1 % faster on the frame,
// traversal 3.88 -> 3.81 ms, and is the default).  Seeding a split node's
// items with the records its other items had merged (5.99 / 65.5) and the exact mode for camera rays (6.14 /
// 70.1) measured slower.
#ifndef RSH_TRAV_WAVES_PRE
#define RSH_TRAV_WAVES_PRE 6
#endif
#ifndef RSH_TRAV_WAVES_MASKS
#define RSH_TRAV_WAVES_MASKS 6
#endif
#ifndef RSH_TRAV_WAVES_EXACT
#define RSH_TRAV_WAVES_EXACT 4
#endif
#ifndef RSH_TRAV_NB
#define RSH_TRAV_NB 1
#endif
#ifndef RSH_TRAV_NB_MASKS
#define RSH_TRAV_NB_MASKS 1
#endif
#ifndef RSH_TRAV_TPRUNE
#define RSH_TRAV_TPRUNE 1
#endif
#ifndef RSH_TRAV_SUB_MIN
#define RSH_TRAV_SUB_MIN 32
#endif
#ifndef RSH_TRAV_NPRUNE
#define RSH_TRAV_NPRUNE 1
#endif
#ifndef RSH_TRAV_QCAP_PRE
#define RSH_TRAV_QCAP_PRE 64
#endif
// Mode 2: a block's cluster ray mask before its triangle spheres are loaded
// and tested (a block no ray may reach is skipped whole).
#ifndef RSH_TRAV_RM_FIRST
#define RSH_TRAV_RM_FIRST 1
#endif
// Mode 1: a block whose cluster sphere no live ray's line may meet is
// skipped before its triangle spheres are loaded, and its leaf mask keeps
// only leaves holding such a ray (1: shadow passes; 2: every mode-1 pass;
// 3: shadow passes and passes without a front-to-back order).
#ifndef RSH_TRAV_SKIP_PRE
#define RSH_TRAV_SKIP_PRE 3
#endif
// CNT: the counting instance (P.counters set); the other carries no counters.
template <int LEAF, int MODE, bool CNT>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(MODE == 2 ? RSH_TRAV_WAVES_MASKS : MODE == 1 ? RSH_TRAV_WAVES_PRE
                                                                   : RSH_TRAV_WAVES_EXACT))) k_traverse64(TravParams P) {
    constexpr bool PRE = MODE >= 1;               // ray / triangle-sphere pre-test
    constexpr bool RM = MODE == 2;                // and ray / cluster-sphere masks
    constexpr uint32_t BRANCH = 64 / LEAF;
    constexpr uint32_t TQ = 128;                  // rings of (survivor, ray) tests
    constexpr unsigned long long LRAYS = LEAF >= 64 ? ~0ull : (1ull << LEAF) - 1ull;   // one leaf's rays
    // Survivor queue.  A block adds up to 64 survivors; modes 1 and 2 keep 64
    // entries (flushed when full, the block's remaining survivors queued
    // after the flush) so that their LDS stays small enough for 24 and 20
    // waves per CU; mode 0 keeps 128 and flushes once 64 are queued.
    constexpr uint32_t QCAP = RM ? 64 : RSH_TRAV_QCAP_PRE;
    // leaf masks, one bit per leaf of the top node
    using LMask = typename std::conditional<BRANCH <= 8, uint8_t,
                  typename std::conditional<BRANCH <= 16, uint16_t, uint32_t>::type>::type;
    constexpr uint32_t PCAP = 512;                // (survivor, leaf) candidate / pair list
    constexpr uint32_t SCH = PCAP / BRANCH < 64 ? PCAP / BRANCH : 64;   // survivors per B+C round
    constexpr uint32_t SUB_MIN = RSH_TRAV_SUB_MIN;   // RM: survivors x rays that earn a sub-cluster mask
    __shared__ float4 s_qsph[QCAP];
    __shared__ uint32_t s_qidx[QCAP];
    __shared__ LMask s_qmask[QCAP];            // leaves the survivor's cluster may pass
    __shared__ float4 s_mt[PRE ? 1 : 64][3];      // Moller-Trumbore vectors of a round's survivors
    __shared__ float4 s_ray0[64];              // (o + d*eps, tmax)
    __shared__ float4 s_ray1[64];              // (d, dot(d, d)), -1 for no ray
    __shared__ unsigned long long s_best[64];  // closest record / shadow flag per ray
    __shared__ uint16_t s_pair[PCAP];          // survivor << 5 | leaf
    __shared__ float4 s_lsph[BRANCH];          // leaf nodes of the top node
    __shared__ float4 s_lcone[BRANCH];
    __shared__ float4 s_laux[BRANCH];          // (cos, tan, lmax, rcos): node_bound
    __shared__ uint8_t s_used[PRE ? 1 : 64];   // survivor of the round appears in a pair
    __shared__ uint16_t s_tq[PRE ? TQ : 1];       // (survivor, ray) tests
    __shared__ uint16_t s_tq2[RM ? TQ : 1];       // RM: candidates that passed the ray pre-test
    __shared__ unsigned long long s_qrm[RM ? QCAP : 1];   // RM: rays that may meet the survivor's cluster
    __shared__ uint8_t s_grp[64];                 // passing clusters of a group, in order
    __shared__ LMask s_gm[64];                    // leaf mask of each cluster of a group
    apply_sizes(P);
    const uint32_t lane = lane_id();
    uint32_t item = blockIdx.x;
    if (P.ps) {     // device-sized: the item total, bounded by the capacity the grid covers
        if (*P.overflow) return;       // a list or descriptor table was cut: this frame is rendered again
        const uint32_t it = P.item_first[P.level_count[1]];
        P.items = it < P.items_cap ? it : P.items_cap;
    }
    if (P.phase == 1) {           // block = top node: its first item
        if (item >= P.level_count[1]) return;
        const uint32_t f = P.item_first[item];
        if (f >= P.item_first[item + 1] || f >= P.items) return;
        item = f;
    }
    if (item >= P.items) return;
    const uint4 desc = P.item_desc[item];
    const uint32_t top = desc.x;
    if (top >= P.level_count[1]) return;
    if (P.phase == 2 && (desc.w & 0x40000000u)) return;     // the node's first item ran in phase 1
    const uint32_t pos = top * 64 + lane;
    const bool valid = pos < P.N;

    uint32_t slot = 0;
    {
        v3 o = mk(0, 0, 0), d = mk(0, 0, 0);
        float tmax = 0.0f;
        if (valid) {
            float4 a = P.rec0[pos], b = P.rec1[pos];
            o = xyz(a); d = xyz(b);
            slot = __float_as_uint(a.w);
            tmax = b.w;
        }
        s_ray0[lane] = f4(o + d * EPS, tmax);
        s_ray1[lane] = f4(d, valid ? dot(d, d) : -1.0f);     // NaN directions never hit: no ray either
    }
    const unsigned long long miss = P.kind == 1 ? 0ull : (P.hit_mode == 0 ? MISS64 : (unsigned long long)MISS32);
    // phase 2: a split node's item starts from what phase 1 merged (the
    // closest record, or the shadow ray already occluded)
    unsigned long long seed = miss;
    if (P.phase == 2 && valid && (desc.w >> 31) == 0u)
        seed = P.kind == 1 ? (P.shadow[slot] != 0u ? 1ull : 0ull) : P.closest[slot];
    s_best[lane] = seed;
    if (lane < BRANCH) {
        const uint32_t bl = top * BRANCH + lane;
        float4 ls = make_float4(0, 0, 0, 0), lc = make_float4(0, 0, 1, 0), la = make_float4(0, 0, 0, 0);
        if (bl < P.level_count[0]) {
            ls = P.nodes[2 * (size_t)bl];
            lc = P.nodes[2 * (size_t)bl + 1];
            const float2 a = P.aux[bl];
            const float2 nb = node_bound(ls, lc, a.x, a.y);
            la = make_float4(a.x, a.y, nb.x, nb.y);
        }
        s_lsph[lane] = ls; s_lcone[lane] = lc; s_laux[lane] = la;
    }
    const uint32_t gt = P.level_offset[1] + top;
    const float4 tS = P.nodes[2 * (size_t)gt], tC = P.nodes[2 * (size_t)gt + 1];
    const float2 tA = P.aux[gt];
    const uint64_t alive = __ballot(valid);
    const bool count = CNT && P.counters != nullptr;
    uint64_t n_top = 0, n_surv = 0, n_leaf = 0, n_tests = 0, n_bsteps = 0, n_csteps = 0, n_mt = 0, n_psteps = 0;
    bool done = false;
    uint32_t qlen = 0;
    __builtin_amdgcn_wave_barrier();

    // Pre-test of ray `ray` (records r0, r1) against a triangle's or a
    // cluster's bounding sphere: its line, and with RSH_TRAV_TPRUNE the part
    // of it a hit can still use -- beyond EPS and, for a shadow ray, short of
    // its length; for a closest-hit ray (exact 64-bit records), short of the
    // closest hit found so far.
    auto may_hit = [&](uint32_t ray, float4 r0, float4 r1, float4 sph) -> bool {
        if (RSH_TRAV_TPRUNE != 0 && !P.line_only) {
            float hi = r0.w;
            if (P.kind != 1)
                hi = P.hit_mode == 0 ? __uint_as_float((uint32_t)(s_best[ray] >> 32)) : __int_as_float(0x7FC00000);
            return ray_may_hit_sphere_range(xyz(r0), xyz(r1), r1.w, sph, EPS, hi);
        } else {
            return ray_may_hit_sphere(xyz(r0), xyz(r1), r1.w, sph);
        }
    };

    // Reach of the item (RSH_TRAV_NPRUNE, not in counting runs, whose pair
    // and test counts stay the reference's): every live ray i starts within
    // r_org of the top node's centre s and can still change its result only
    // at a distance below hi_i |d_i| from its start (hi as in may_hit), so a
    // sphere (q, rho) with |q - s| > (reach + r_org + rho) 1.001 + 1e-5,
    // reach = max_i hi_i |d_i|, holds no triangle that changes a result (the
    // slack covers the FP32 rounding of this test and of ray_triangle's
    // parameter, ~1e-6 of these magnitudes).  reach is infinite while some
    // closest-hit ray has no hit, and is recomputed after each flush.
    const bool prune = RSH_TRAV_NPRUNE != 0 && !count && ((MODE == 1 && P.kind == 1) || P.phase != 0);
    float r_org = 0.0f, dlen = 0.0f, reach = __builtin_inff();
    if (prune) {
        const v3 o = xyz(s_ray0[lane]), d = xyz(s_ray1[lane]);
        dlen = valid ? len(d) : 0.0f;
        r_org = wave_max(valid ? len(o - xyz(tS)) : 0.0f);
    }
    auto update_reach = [&]() {
        if (!prune) return;
        float v = 0.0f;        // no live ray: nothing left to reach
        const unsigned long long b = s_best[lane];
        if (valid && (P.kind != 1 || b == 0ull)) {
            const float hi = P.kind == 1 ? s_ray0[lane].w
                                         : (P.hit_mode == 0 ? __uint_as_float((uint32_t)(b >> 32)) : __builtin_inff());
            v = hi * dlen;
            if (!(v <= 3.0e38f)) v = __builtin_inff();    // NaN (no hit yet) or overflow
        }
        reach = wave_max(v);
    };
    auto beyond = [&](float4 sph) -> bool {
        const v3 w = xyz(sph) - xyz(tS);
        const float lim = (reach + r_org + sph.w) * 1.001f + 1e-5f;
        return dot(w, w) > lim * lim;
    };
    update_reach();

    // ---- stages B + C over the queued survivors, SCH at a time:
    //  B  the (survivor, leaf) candidates allowed by each survivor's cluster
    //     leaf mask are listed in s_pair and tested 64 per step; passing pairs
    //     are compacted in place and their survivors marked used;
    //  load the Moller-Trumbore vectors of the used survivors only;
    //  C  the pair list is walked BRANCH pairs per step, lane = pair p0 + lane / LEAF,
    //     ray lane % LEAF of the pair's leaf; hits merge per ray in LDS.
    auto flush = [&]() {
        for (uint32_t c0 = 0; c0 < qlen && !done; c0 += SCH) {
            const uint32_t n = qlen - c0 < SCH ? qlen - c0 : SCH;
            uint32_t lm = lane < n ? s_qmask[c0 + lane] : 0u;
            const uint32_t kc = (uint32_t)__popc(lm);
            const uint32_t incl = wave_incl_add(kc);
            const uint32_t ncand_ = __shfl(incl, 63, 64);
            uint32_t o = incl - kc;
            while (lm) {
                const uint32_t b = (uint32_t)__builtin_ctz(lm);
                lm &= lm - 1u;
                s_pair[o++] = (uint16_t)((lane << 5) | b);
            }
            if constexpr (!PRE) s_used[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            uint32_t np_ = 0;
            for (uint32_t c = 0; c < ncand_; c += 64) {
                const uint32_t ci = c + lane;
                bool lp = false;
                uint32_t e = 0;
                if (ci < ncand_) {
                    e = s_pair[ci];
                    const uint32_t b = e & 31u;
                    const float4 la = s_laux[b];
                    lp = sphere_node_fast(s_lsph[b], s_lcone[b], s_qsph[c0 + (e >> 5)], la.x, la.y, la.w);
                }
                const uint64_t pm = __ballot(lp);
                if (lp) {     // in place: entries c.. were read above, writes land below c + 64
                    s_pair[np_ + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull))] = (uint16_t)e;
                    if constexpr (!PRE) s_used[e >> 5] = 1;
                }
                np_ += (uint32_t)__popcll(pm);
                if (count) { n_leaf += (uint64_t)__popcll(pm); n_bsteps++; }
            }
            __builtin_amdgcn_wave_barrier();
            if (np_ == 0) continue;
            if constexpr (!PRE) {
                if (lane < n && s_used[lane]) {
                    size_t t = s_qidx[c0 + lane];
                    if (t >= P.T) t = 0;     // queue entries are < T by construction
                    s_mt[lane][0] = P.tri_mt[3 * t];
                    s_mt[lane][1] = P.tri_mt[3 * t + 1];
                    s_mt[lane][2] = P.tri_mt[3 * t + 2];
                }
                __builtin_amdgcn_wave_barrier();
            }
            // exact test of survivor sv against ray `ray` (a valid ray), merged per ray
            auto exact = [&](uint32_t sv, uint32_t ray, float4 r0, float4 r1) {
                float4 A, B, Cc;
                if constexpr (PRE) {     // few tests pass the pre-test: vectors straight from memory
                    size_t t = s_qidx[c0 + sv];
                    if (t >= P.T) t = 0;     // queue entries are < T by construction
                    A = P.tri_mt[3 * t]; B = P.tri_mt[3 * t + 1]; Cc = P.tri_mt[3 * t + 2];
                } else {
                    A = s_mt[sv][0]; B = s_mt[sv][1]; Cc = s_mt[sv][2];
                }
                const float th = ray_triangle(xyz(r0), xyz(r1), xyz(A), xyz(B), xyz(Cc));
                if (count) n_mt++;
                if (P.kind == 1) {
                    if (th > EPS && th < r0.w) s_best[ray] = 1ull;
                } else if (th > EPS) {
                    const unsigned long long rec = pack_hit(th, s_qidx[c0 + sv], P.hit_mode);
                    if (rec < s_best[ray]) atomicMin(&s_best[ray], rec);
                }
            };
            if constexpr (RM) {
                // C0: the (pair, ray) candidates -- rays of the pair's leaf that are
                // live and may meet the survivor's cluster -- are listed in the ring
                // s_tq, one pair per lane and one ray per lane and round;
                // C1: each candidate ray's line against the survivor's bounding
                // sphere, 64 per step, passing ones to the ring s_tq2;
                // C2: the exact test, 64 per step.
                const uint64_t lt = (1ull << lane) - 1ull;
                uint32_t qh = 0, qt = 0, rh = 0, rt = 0;
                auto c2 = [&](uint32_t cnt) {
                    if (count) n_csteps++;
                    if (lane < cnt) {
                        const uint32_t code = s_tq2[(rh + lane) & (TQ - 1)];
                        const uint32_t sv = code >> 6, ray = code & 63u;
                        if (P.kind != 1 || s_best[ray] == 0ull) exact(sv, ray, s_ray0[ray], s_ray1[ray]);
                    }
                    rh += cnt;
                    __builtin_amdgcn_wave_barrier();
                };
                auto c1 = [&](uint32_t cnt) {
                    if (count) n_psteps++;
                    bool keep = false;
                    uint32_t code = 0;
                    if (lane < cnt) {
                        code = s_tq[(qh + lane) & (TQ - 1)];
                        const uint32_t sv = code >> 6, ray = code & 63u;
                        if (P.kind != 1 || s_best[ray] == 0ull)
                            keep = may_hit(ray, s_ray0[ray], s_ray1[ray], s_qsph[c0 + sv]);
                    }
                    const uint64_t km = __ballot(keep);
                    if (keep) s_tq2[(rt + (uint32_t)__popcll(km & lt)) & (TQ - 1)] = (uint16_t)code;
                    rt += (uint32_t)__popcll(km);
                    qh += cnt;
                    __builtin_amdgcn_wave_barrier();
                    if (rt - rh >= 64) c2(64);
                };
                uint64_t live = alive;
                if (P.kind == 1) live &= ~__ballot(s_best[lane] != 0ull);
                for (uint32_t p0 = 0; p0 < np_; p0 += 64) {
                    const uint32_t pi = p0 + lane;
                    uint32_t rb = 0, sv = 0, b = 0;
                    if (pi < np_) {
                        const uint32_t e = s_pair[pi];
                        sv = e >> 5;
                        b = e & 31u;
                        const uint32_t lv = (uint32_t)((live >> (b * LEAF)) & LRAYS);
                        if (count) n_tests += (uint64_t)__popc(lv);
                        rb = (uint32_t)(s_qrm[c0 + sv] >> (b * LEAF)) & lv;
                    }
                    for (;;) {
                        const uint64_t hm = __ballot(rb != 0u);
                        if (hm == 0ull) break;
                        if (rb) {
                            const uint32_t r = (uint32_t)__builtin_ctz(rb);
                            rb &= rb - 1u;
                            s_tq[(qt + (uint32_t)__popcll(hm & lt)) & (TQ - 1)] = (uint16_t)((sv << 6) | (b * LEAF + r));
                        }
                        qt += (uint32_t)__popcll(hm);
                        __builtin_amdgcn_wave_barrier();
                        if (qt - qh >= 64) c1(64);
                    }
                }
                while (qt != qh) c1(qt - qh < 64 ? qt - qh : 64);
                while (rt != rh) c2(rt - rh < 64 ? rt - rh : 64);
            } else if constexpr (PRE) {
                // C1: each ray's line against the survivor's bounding sphere, BRANCH
                // pairs per step; tests that may hit queue in the ring s_tq
                // (survivor << 6 | ray), and C2 runs the exact test 64 at a time.
                uint32_t qh = 0, qt = 0;
                auto c2 = [&](uint32_t cnt) {
                    if (count) n_csteps++;
                    if (lane < cnt) {
                        const uint32_t code = s_tq[(qh + lane) & (TQ - 1)];
                        const uint32_t sv = code >> 6, ray = code & 63u;
                        if (P.kind != 1 || s_best[ray] == 0ull) exact(sv, ray, s_ray0[ray], s_ray1[ray]);
                    }
                    qh += cnt;
                    __builtin_amdgcn_wave_barrier();
                };
                for (uint32_t p0 = 0; p0 < np_; p0 += BRANCH) {
                    const uint32_t pi = p0 + lane / LEAF;
                    bool keep = false;
                    uint32_t code = 0;
                    if (count) n_psteps++;
                    if (pi < np_) {
                        const uint32_t e = s_pair[pi];
                        const uint32_t sv = e >> 5;
                        const uint32_t ray = (e & 31u) * LEAF + lane % LEAF;
                        const float4 r0 = s_ray0[ray], r1 = s_ray1[ray];
                        if (r1.w >= 0.0f && (P.kind != 1 || s_best[ray] == 0ull)) {
                            if (count) n_tests++;
                            keep = may_hit(ray, r0, r1, s_qsph[c0 + sv]);
                            code = (sv << 6) | ray;
                        }
                    }
                    const uint64_t km = __ballot(keep);
                    if (keep) s_tq[(qt + (uint32_t)__popcll(km & ((1ull << lane) - 1ull))) & (TQ - 1)] = (uint16_t)code;
                    qt += (uint32_t)__popcll(km);
                    __builtin_amdgcn_wave_barrier();
                    if (qt - qh >= 64) c2(64);
                }
                while (qt != qh) c2(qt - qh < 64 ? qt - qh : 64);
            } else {
                for (uint32_t p0 = 0; p0 < np_; p0 += BRANCH) {
                    const uint32_t pi = p0 + lane / LEAF;
                    if (count) n_csteps++;
                    if (pi < np_) {
                        const uint32_t e = s_pair[pi];
                        const uint32_t sv = e >> 5;
                        const uint32_t ray = (e & 31u) * LEAF + lane % LEAF;
                        const float4 r0 = s_ray0[ray], r1 = s_ray1[ray];
                        const bool live = r1.w >= 0.0f && (P.kind != 1 || s_best[ray] == 0ull);
                        if (live) {
                            if (count) n_tests++;
                            exact(sv, ray, r0, r1);
                        }
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (P.kind == 1 && (__ballot(s_best[lane] != 0ull) & alive) == alive) done = true;
        }
        qlen = 0;
        update_reach();
    };

    // ---- cursor over this item's clusters of the node's cluster stream, from the
    // descriptor's list entry (RAH: the list of all objects, offsets obj_cl_base)
    const bool rah = P.algorithm == 1;
    const bool single = (desc.w >> 31) != 0u;
    const float2 t_bound = node_bound(tS, tC, tA.x, tA.y);
    uint64_t n_cl = 0;
    uint32_t rem = desc.w & 0x3FFFFFFFu;      // clusters of the item not yet taken from an entry
    uint32_t j = desc.y;                       // current entry (absolute index) of the object list
    uint32_t e_first = 0, e_last = 0, e_exlo = 0, e_exhi = 0, e_cbase = 0, e_ncl = 0;
    uint32_t lc_cur = 0, lc_end = 0;           // local clusters of the entry still to pre-test
    uint32_t g0 = 0;                           // first local cluster of the tested group
    uint64_t gmask = 0;                        // clusters of the group that passed the pre-test
    // Take the `rem`-bounded part of list entry jj, starting `off` clusters in.
    auto load_entry = [&](uint32_t jj, uint32_t off) {
        const uint4 en = P.obj_ent[jj];          // (first, last, first cluster, excluded from)
        e_first = en.x; e_last = en.y; e_cbase = en.z;
        e_exlo = en.w; e_exhi = en.w ? en.y : 0u;  // an excluded batch ends at the object's last triangle
        e_ncl = (en.y - en.x + CLUSTER) / CLUSTER;
        const uint32_t avail = e_ncl - off;
        const uint32_t take = avail < rem ? avail : rem;
        lc_cur = off; lc_end = off + take;
        rem -= take;
    };
    if (rem) load_entry(j, desc.z);
    // Leaf masks of the clusters in gm (bits = local cluster index in the group
    // at g0): each passing cluster is tested against the BRANCH leaves
    // (conservative, as the cluster pre-test), CPS clusters per wave step.
    // Returns the clusters that may pass some leaf; their masks are in s_gm.
    auto leaf_masks = [&](uint64_t gm) -> uint64_t {
        constexpr uint32_t CPS = 64 / BRANCH;
        constexpr uint32_t LMASK = BRANCH >= 32 ? 0xFFFFFFFFu : (1u << BRANCH) - 1u;
        const uint32_t g = (uint32_t)__popcll(gm);
        if ((gm >> lane) & 1ull) s_grp[__popcll(gm & ((1ull << lane) - 1ull))] = (uint8_t)lane;
        __builtin_amdgcn_wave_barrier();
        const uint32_t leaf = lane % BRANCH;
        const bool leaf_ok = top * BRANCH + leaf < P.level_count[0];
        const float4 ls = s_lsph[leaf], lc = s_lcone[leaf], la = s_laux[leaf];
        for (uint32_t j0 = 0; j0 < g; j0 += CPS) {
            const uint32_t ci = j0 + lane / BRANCH;
            bool lp = false;
            if (ci < g && leaf_ok)
                lp = cluster_may_pass(ls, lc, la.y, make_float2(la.z, la.w), P.cl_sphere[e_cbase + g0 + s_grp[ci]]);
            const uint64_t bm = __ballot(lp);
            if (lane < CPS && j0 + lane < g) s_gm[s_grp[j0 + lane]] = (LMask)((bm >> (lane * BRANCH)) & LMASK);
        }
        __builtin_amdgcn_wave_barrier();
        return __ballot(((gm >> lane) & 1ull) && s_gm[lane] != 0u);
    };
    // Next block of one cluster's triangles [tb, tl] that passed the cluster
    // pre-test and may pass some leaf (cm: those leaves); clusters are
    // pre-tested 64 at a time, one per lane.
    auto next_block = [&](uint32_t& tb, uint32_t& tl, uint32_t& cm, uint32_t& cid) -> bool {
        while (gmask == 0ull) {
            if (lc_cur >= lc_end) {
                if (rem == 0) return false;     // the item's cluster count is exact
                j++;
                load_entry(j, 0);
                continue;
            }
            g0 = lc_cur;
            const uint32_t n = lc_end - lc_cur < 64 ? lc_end - lc_cur : 64;
            lc_cur += n;
            uint64_t gm = 1ull;                 // one cluster: the object test already stands for it
            if (e_ncl != 1) {
                bool cp = false;
                if (lane < n) {
                    const float4 cs = P.cl_sphere[e_cbase + g0 + lane];
                    cp = cluster_may_pass(tS, tC, tA.y, t_bound, cs) && !(prune && beyond(cs));
                }
                gm = __ballot(cp);
                if (count) n_cl += n;
            }
            if (gm) gmask = leaf_masks(gm);
        }
        const uint32_t i = (uint32_t)__builtin_ctzll(gmask);
        gmask &= gmask - 1ull;
        cm = s_gm[i];
        cid = e_cbase + g0 + i;
        tb = e_first + (g0 + i) * CLUSTER;
        tl = tb + (CLUSTER - 1) < e_last ? tb + (CLUSTER - 1) : e_last;
        return true;
    };

    // the lane's ray may still change its result (a shadow ray drops out once occluded)
    auto live_lane = [&]() -> bool { return valid && (P.kind != 1 || s_best[lane] == 0ull); };
    // leaves holding a ray of mask r
    auto leaves_of = [&](uint64_t r) {
        uint32_t lr = 0;
#pragma unroll
        for (uint32_t b = 0; b < BRANCH; b++) lr |= (((r >> (b * LEAF)) & LRAYS) != 0ull ? 1u : 0u) << b;
        return lr;
    };
    // NB blocks per iteration: their sphere loads are issued back to back and
    // consumed in order, so only the first load's latency is exposed.
    constexpr int NB = RM ? RSH_TRAV_NB_MASKS : RSH_TRAV_NB;
    bool have = rem != 0 || lc_cur < lc_end;
    while (have && !done) {
        uint32_t tbk[NB], cmk[NB], cidk[NB];
        bool ink[NB];
        float4 tsk[NB];
        uint64_t rmk[RM ? NB : 1];
#pragma unroll
        for (int k = 0; k < NB; k++) {
            uint32_t tb = 0, tl = 0, cm = 0, cid = 0;
            have = have && next_block(tb, tl, cm, cid);
            if constexpr (RM) {
                // rays whose line (and usable range) may meet the cluster's sphere;
                // counting runs keep every ray, so the counts stay the reference's
                rmk[k] = ~0ull;
                if (RSH_TRAV_RM_FIRST && have && !count) {
                    rmk[k] = __ballot(live_lane() && may_hit(lane, s_ray0[lane], s_ray1[lane], P.cl_sphere[cid]));
                    cm &= leaves_of(rmk[k]);
                }
            } else if constexpr (MODE == 1) {
                if (RSH_TRAV_SKIP_PRE != 0 && have && !count &&
                    (RSH_TRAV_SKIP_PRE == 2 || P.kind == 1 || (RSH_TRAV_SKIP_PRE == 3 && P.obj_order == nullptr)))
                    cm &= leaves_of(__ballot(live_lane() && may_hit(lane, s_ray0[lane], s_ray1[lane], P.cl_sphere[cid])));
            }
            const uint32_t t = tb + lane;
            ink[k] = have && cm != 0u && t <= tl && !(t >= e_exlo && t < e_exhi);
            tbk[k] = tb;
            cmk[k] = cm;
            cidk[k] = cid;
            tsk[k] = make_float4(0, 0, 0, 0);
            if (ink[k]) tsk[k] = P.tri_sphere[t];
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
            if (done) break;             // every ray of the wave is occluded: stop queueing
            // ---- stage A
            bool p = false;
            if (ink[k]) p = sphere_node_fast(tS, tC, tsk[k], tA.x, tA.y, t_bound.y) && !(prune && beyond(tsk[k]));
            uint64_t m = __ballot(p);
            if (count) { n_top += (uint64_t)__popcll(__ballot(ink[k])); n_surv += (uint64_t)__popcll(m); }
            if (m == 0ull) continue;
            uint32_t cm = cmk[k];            // leaves the block's cluster (RM: the lane's survivor) may pass
            uint64_t rm = ~0ull;
            if constexpr (RM) {
                // Rays whose line may meet the cluster's sphere, then, per
                // sub-cluster holding survivors, its sphere (the argument behind
                // ray_may_hit_sphere holds for any sphere that holds the
                // triangle); a survivor keeps the rays of its sub-cluster and the
                // leaves holding one of them.  Counting runs keep every ray and
                // leaf, so the counts stay the reference's.
                if (!count) {
                    const float4 ro = s_ray0[lane], rd = s_ray1[lane];
                    const uint32_t use = P.sub_use[cidk[k]];
                    if (RSH_TRAV_RM_FIRST) {
                        rm = rmk[k];         // cm already holds only leaves with such rays
                    } else {
                        rm = __ballot(((alive >> lane) & 1ull) && may_hit(lane, ro, rd, P.cl_sphere[cidk[k]]));
                        cm &= leaves_of(rm);
                        if (cm == 0u) continue;
                    }
                    // sub-clusters worth a mask: small enough (sub_use) and survivors
                    // x rays at least SUB_MIN
                    const uint32_t nrc = (uint32_t)__popcll(rm);
                    uint32_t subs = 0;
#pragma unroll
                    for (uint32_t k8 = 0; k8 < SUB_PER_CLUSTER; k8++)
                        subs |= (uint32_t)__popcll((m >> (k8 * SUBCLUSTER)) & ((1ull << SUBCLUSTER) - 1ull)) * nrc >= SUB_MIN
                                    ? 1u << k8 : 0u;
                    subs &= use;
                    if (subs) {
                        const bool rin = (rm >> lane) & 1ull;
                        const float4* ssp = P.sub_sphere + (size_t)cidk[k] * SUB_PER_CLUSTER;
                        do {
                            const uint32_t k8 = (uint32_t)__builtin_ctz(subs);
                            subs &= subs - 1u;
                            const uint64_t rk = __ballot(rin && may_hit(lane, ro, rd, ssp[k8]));
                            if (lane / SUBCLUSTER == k8) rm = rk;
                        } while (subs);
                        cm &= leaves_of(rm);
                        p = p && cm != 0u;
                        m = __ballot(p);
                        if (m == 0ull) continue;
                    }
                }
            }
            const uint32_t tot = qlen + (uint32_t)__popcll(m);
            const uint32_t r = qlen + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            auto enqueue = [&](uint32_t at) {
                s_qsph[at] = tsk[k];
                s_qidx[at] = tbk[k] + lane;
                s_qmask[at] = (LMask)cm;
                if constexpr (RM) s_qrm[at] = rm;
            };
            if constexpr (QCAP >= 128) {     // qlen < 64 here: the block fits
                if (p) enqueue(r);
                qlen = tot;
                __builtin_amdgcn_wave_barrier();
                if (qlen >= 64) flush();
            } else {
                if (p && r < QCAP) enqueue(r);
                __builtin_amdgcn_wave_barrier();
                if (tot < QCAP) {
                    qlen = tot;
                    continue;
                }
                qlen = QCAP;
                flush();
                if (done) break;
                if (p && r >= QCAP) enqueue(r - QCAP);
                qlen = tot - QCAP;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (qlen && !done) flush();

    // Results start as misses; a node cut into several items merges by atomicMin.
    __builtin_amdgcn_wave_barrier();
    const unsigned long long best = s_best[lane];
    if (valid) {
        if (P.kind == 1) { if (best != 0ull) P.shadow[slot] = MISS32; }
        else if (single) P.closest[slot] = best;
        else if (best != seed) atomicMin(&P.closest[slot], best);
    }
    if (count) {
        unsigned long long tsum = n_tests, msum = n_mt;
        for (int dlt = 32; dlt > 0; dlt >>= 1) {
            tsum += __shfl_xor(tsum, dlt, 64);
            msum += __shfl_xor(msum, dlt, 64);
        }
        if (lane == 0) {
            unsigned long long* cs = P.counters + (size_t)(item % COUNTER_SLOTS) * COUNTERS_PER_SLOT;
            atomicAdd(&cs[CNT_LEAF_PAIRS], (unsigned long long)n_leaf);
            atomicAdd(&cs[CNT_RAY_TESTS], tsum);
            atomicAdd(&cs[CNT_TOP_SURV], (unsigned long long)n_surv);
            atomicAdd(&cs[CNT_TRI_TESTS], (unsigned long long)n_top);
            atomicAdd(&cs[CNT_CLUSTER_TESTS], (unsigned long long)n_cl);
            atomicAdd(&cs[CNT_B_STEPS], (unsigned long long)n_bsteps);
            atomicAdd(&cs[CNT_C_STEPS], (unsigned long long)(n_csteps + n_psteps));
            atomicAdd(&cs[CNT_MT_TESTS], msum);
        }
    }
}

// ------------------------------------------- materialised traversal ------
__device__ __forceinline__ uint32_t create_hit(uint32_t node, uint32_t tri) {
    return ((node << 14) & 0xFFFFC000u) + (tri & 0x00003FFFu);
}
__device__ __forceinline__ uint32_t hit_node(uint32_t h) { return (h & 0xFFFFC000u) >> 14; }
__device__ __forceinline__ uint32_t hit_tri(uint32_t h) { return h & 0x00003FFFu; }

__device__ __forceinline__ float4 tri_sphere_at(const float4* tp, uint32_t t) {
    return triangle_sphere(xyz(tp[3 * (size_t)t]), xyz(tp[3 * (size_t)t + 1]), xyz(tp[3 * (size_t)t + 2]));
}
__device__ __forceinline__ bool node_test(const float4* h, uint32_t g, float4 ts) {
    float4 s = h[2 * (size_t)g], c = h[2 * (size_t)g + 1];
    return sphere_node(s, c, ts, cos_s(c.w), tan_s(c.w));
}

// RayTracer.cu:1152-1280 (run-length culling flags by atomicAdd)
__global__ void __launch_bounds__(256) k_bsphere_intersections(const float4* h, const float* spheres3, uint32_t K,
                                                               uint32_t Tb, uint32_t offset, uint32_t nodeOffset,
                                                               uint32_t nodeTotal, uint32_t* flags) {
    const uint64_t n = (uint64_t)nodeTotal * K;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t node = (uint32_t)(x / K), b = (uint32_t)(x % K);
        float4 s = h[2 * (size_t)(nodeOffset + node)], c = h[2 * (size_t)(nodeOffset + node) + 1];
        const float* p = spheres3 + 6 * (size_t)b;
        float ry = p[4], rz = p[5];
        bool res = sphere_node(s, c, make_float4(p[0], p[1], p[2], p[3]), cos_s(c.w), tan_s(c.w));
        const float lower = (float)offset, upper = (float)(offset + Tb);
        uint32_t* f = flags + (size_t)node * Tb;
        if (res) {
            if (ry >= lower && ry < upper && rz >= lower && rz < upper)
                atomicAdd(&f[(uint32_t)ry - offset], (uint32_t)rz - (uint32_t)ry + 1);
            else if (ry < lower && rz >= lower && rz < upper)
                atomicAdd(&f[0], (uint32_t)rz - offset + 1);
            else if (ry >= lower && ry < upper && rz >= upper)
                atomicAdd(&f[(uint32_t)ry - offset], Tb);
            else if (ry < lower && rz > upper)
                atomicAdd(&f[0], (uint32_t)rz - (uint32_t)ry + 1 - ((uint32_t)rz - (offset + Tb) + 1));
        } else {
            if (ry >= lower && ry < upper && rz >= lower && rz + 1 < upper)
                atomicAdd(&f[(uint32_t)rz - offset + 1], (uint32_t)rz - (uint32_t)ry + 1);
            else if (ry < lower && rz >= lower && (rz + 1) < upper)
                atomicAdd(&f[(uint32_t)rz - offset + 1], (uint32_t)rz - offset + 1);
        }
    }
}

// RayTracer.cu:1282-1313
__global__ void __launch_bounds__(256) k_bsphere_hits(uint32_t Tb, uint32_t nodeTotal, const uint32_t* scan,
                                                      uint32_t* flags, uint32_t* hits) {
    const uint64_t n = (uint64_t)nodeTotal * Tb;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t node = (uint32_t)(x / Tb), t = (uint32_t)(x % Tb);
        int off = (int)scan[x];
        if (node > 0) off -= (int)scan[(size_t)node * Tb - 1];
        bool hit = (uint32_t)off > t;
        hits[x] = hit ? create_hit(node, t) : 0u;
        flags[x] = hit ? 0u : 1u;
    }
}

// RayTracer.cu:1315-1405 (rah = 1: every (node, triangle) pair)
__global__ void __launch_bounds__(256) k_level0_hits(const float4* h, const float4* tp, uint32_t Tb, uint32_t offset,
                                                     uint32_t hitTotal, uint32_t nodeOffset, uint32_t rah,
                                                     uint32_t* flags, uint32_t* hits, const uint32_t* thits) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < hitTotal; x += gridDim.x * blockDim.x) {
        uint32_t hit = rah ? create_hit(x / Tb, x % Tb) : thits[x];
        uint32_t node = hit_node(hit), t = hit_tri(hit);
        float4 ts = tri_sphere_at(tp, offset + t);
        flags[x] = node_test(h, nodeOffset + node, ts) ? 0u : 1u;
        hits[x] = hit;
    }
}

// RayTracer.cu:1407-1462
__global__ void __launch_bounds__(256) k_levelN_hits(const float4* h, const float4* tp, uint32_t offset,
                                                     uint32_t hitTotal, uint32_t nodeOffset, uint32_t nodeWriteTotal,
                                                     uint32_t branch, uint32_t* flags, uint32_t* hits,
                                                     const uint32_t* thits) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < hitTotal; x += gridDim.x * blockDim.x) {
        uint32_t hit = thits[x];
        uint32_t node = hit_node(hit), t = hit_tri(hit);
        float4 ts = tri_sphere_at(tp, offset + t);
        for (uint32_t i = 0; i < branch; i++) {
            uint32_t c = node * branch + i;
            uint32_t f = 1u;
            if (c < nodeWriteTotal) f = node_test(h, nodeOffset + c, ts) ? 0u : 1u;
            flags[(size_t)x * branch + i] = f;
            hits[(size_t)x * branch + i] = create_hit(c, t);
        }
    }
}

// RayTracer.cu:1498-1614 (hard shadows) / :1711-1774 (reference packing)
__global__ void __launch_bounds__(256) k_hits_intersect(const float* rays, const uint32_t* svalues, const uint32_t* hits,
                                                        const float4* tp, uint32_t hitTotal, uint32_t N,
                                                        uint32_t offset, uint32_t W, uint32_t Lmax, uint32_t leaf,
                                                        const float4* gpos, uint32_t shadow, uint32_t* out) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < hitTotal; x += gridDim.x * blockDim.x) {
        uint32_t h = hits[x], node = hit_node(h), t = hit_tri(h);
        v3 v0 = xyz(tp[3 * (size_t)(offset + t)]);
        v3 e1 = xyz(tp[3 * (size_t)(offset + t) + 1]) - v0;
        v3 e2 = xyz(tp[3 * (size_t)(offset + t) + 2]) - v0;
        for (uint32_t i = 0; i < leaf; i++) {
            if (node * leaf + i >= N) break;
            uint32_t r = svalues[node * leaf + i];
            const float* rp = rays + 6 * (size_t)r;
            v3 o = mk(rp[0], rp[1], rp[2]), d = mk(rp[3], rp[4], rp[5]);
            float th = ray_triangle(o + d * EPS, d, v0, e1, e2);
            if (shadow) {
                v3 P = xyz(gpos[(size_t)(r % W) + (size_t)(r / (W * Lmax)) * W]);
                if (th > EPS && th < len(o - P) - EPS * 2.0f) out[r] = MISS32;
            } else if (th > EPS) {
                uint32_t rec = (uint32_t)pack_hit(th, offset + t, 1);
                atomicMin(&out[r], rec);
            }
        }
    }
}

}  // namespace

// ================================================================ launchers
// ps: device-sized (N is the capacity the grid covers)
void launch_gather(const float* rays, const uint32_t* sorted_values, uint32_t N, const PassSizes* ps,
                   const float4* shadow_pos, const float* tmax, uint32_t W, uint32_t Lmax, uint32_t slot_shift,
                   float4* rec0, float4* rec1, hipStream_t s, const float4* light_pos) {
    if (N) k_gather<<<grid_for(N, 256, 65536), 256, 0, s>>>(rays, sorted_values, N, ps, shadow_pos, tmax, light_pos,
                                                            W, Lmax, slot_shift, rec0, rec1, fast_div(W),
                                                            fast_div(Lmax));
}

// ps: device-sized (N, counts and offsets are capacities; the kernels read
// the sizes; grids stride over them)
void launch_hierarchy(const float4* rec0, const float4* rec1, const PassSizes* ps, const float* rays,
                      const uint32_t* sorted_values, uint32_t N, uint32_t leaf, uint32_t branch, uint32_t depth,
                      float r0, float s0, uint32_t mode, float4* nodes, float2* aux, const uint32_t* counts,
                      const uint32_t* offsets, hipStream_t s) {
    if (!N) return;
    if (rec0) k_level0<true><<<grid_for(counts[0], 256), 256, 0, s>>>(rec0, rec1, nullptr, nullptr, N, ps, leaf, r0,
                                                                       s0, mode, nodes, aux);
    else k_level0<false><<<grid_for(counts[0], 256), 256, 0, s>>>(nullptr, nullptr, rays, sorted_values, N, ps, leaf,
                                                                   r0, s0, mode, nodes, aux);
    for (uint32_t k = 1; k < depth; k++)
        k_levelN<<<grid_for(counts[k], 256), 256, 0, s>>>(nodes, aux, offsets[k - 1], counts[k - 1], offsets[k],
                                                           counts[k], branch, mode, ps, k);
}

void launch_cull_count(const TravParams& p, uint32_t* npass, uint32_t* cand, unsigned long long* pmask,
                       hipStream_t s) {
    const uint32_t top = p.level_count[1];
    k_cull<false><<<top + 1, 64, 0, s>>>(p, npass, cand, nullptr, nullptr, pmask);
}

void launch_cull_list(const TravParams& p, uint4* obj_ent, uint32_t* obj_cum, unsigned long long* pmask,
                      hipStream_t s) {
    const uint32_t top = p.level_count[1];
    if (top && p.algorithm != 1)
        k_cull<true><<<top, 64, 0, s>>>(p, nullptr, nullptr, obj_ent, obj_cum, pmask);
}

void launch_item_table(const uint32_t* cand, const float2* top_aux, uint32_t ntop, uint32_t max_cand,
                       const uint32_t* npass, unsigned long long* table, hipStream_t s, uint32_t* done,
                       unsigned long long* host, uint32_t seq) {
    static_assert(ITEM_TABLE == 64, "one lane per chunk size");
    const bool fast = div_hi_ok((uint64_t)max_cand + 4 * ITEM_TABLE - 1, 4 * ITEM_TABLE);
    if (ntop)
        k_item_table<<<(ntop + TABLE_SLICE - 1) / TABLE_SLICE, 64, 0, s>>>(cand, top_aux, ntop, table, fast,
                                                                           npass, done, host, seq);
}

void launch_item_count(const uint32_t* cand, uint32_t chunk, uint32_t ntop, uint32_t* nitems, hipStream_t s) {
    k_item_count<<<ntop / 256 + 1, 256, 0, s>>>(cand, chunk, ntop, nitems);
}

void launch_cull_items(const TravParams& p, uint4* obj_ent, uint32_t* obj_cum, const unsigned long long* pmask,
                       uint4* desc, hipStream_t s) {
    const uint32_t top = p.level_count[1];
    if (top) k_cull_items<<<top, 64, 0, s>>>(p, obj_ent, obj_cum, pmask, desc);
}

void launch_item_desc(const TravParams& p, uint4* desc, hipStream_t s) {
    const uint32_t ntop = p.level_count[1];
    if (ntop) k_item_desc<<<ntop, 64, 0, s>>>(p, desc);
}

template <int LEAF>
hipError_t launch64(const TravParams& p, uint32_t blocks, hipStream_t s) {
    const bool c = p.counters != nullptr;
    switch (p.prefilter) {
        case 0: c ? k_traverse64<LEAF, 0, true><<<blocks, 64, 0, s>>>(p) : k_traverse64<LEAF, 0, false><<<blocks, 64, 0, s>>>(p); break;
        case 1: c ? k_traverse64<LEAF, 1, true><<<blocks, 64, 0, s>>>(p) : k_traverse64<LEAF, 1, false><<<blocks, 64, 0, s>>>(p); break;
        case 2: c ? k_traverse64<LEAF, 2, true><<<blocks, 64, 0, s>>>(p) : k_traverse64<LEAF, 2, false><<<blocks, 64, 0, s>>>(p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_traverse(const TravParams& p, hipStream_t s) {
    const uint32_t top = p.level_count[p.depth - 1];
    if (uses_items(p)) {
        if (p.items == 0) return hipSuccess;
        if (p.chunk == 0 || !p.item_desc) return hipErrorInvalidValue;
        if (!p.obj_ent || !p.obj_cum || (p.algorithm != 1 && !p.obj_off)) return hipErrorInvalidValue;
        if (p.items > MAX_TRAVERSE_ITEMS) return hipErrorInvalidValue;
        // one wave per workgroup: a finished item releases its LDS at once;
        // phase 1: one block per top node (its first item)
        const uint32_t blocks = p.phase == 1 ? p.level_count[1] : p.items;
        switch (p.leaf) {
            case 2: return launch64<2>(p, blocks, s);
            case 4: return launch64<4>(p, blocks, s);
            case 8: return launch64<8>(p, blocks, s);
            case 16: return launch64<16>(p, blocks, s);
            case 32: return launch64<32>(p, blocks, s);
            default: return hipErrorInvalidValue;
        }
    }
    if (p.splits == 0) return hipErrorInvalidValue;
    const uint64_t waves = (uint64_t)top * p.waves_per_top * p.splits;
    if (waves == 0) return hipSuccess;
    if (waves > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t blocks = (uint32_t)((waves + 3) / 4);
    switch (p.depth) {
        case 1: k_traverse<1><<<blocks, 256, 0, s>>>(p); break;
        case 2: k_traverse<2><<<blocks, 256, 0, s>>>(p); break;
        case 3: k_traverse<3><<<blocks, 256, 0, s>>>(p); break;
        case 4: k_traverse<4><<<blocks, 256, 0, s>>>(p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

void launch_bsphere_intersections(const float4* h, const float* spheres3, uint32_t K, uint32_t Tb, uint32_t offset,
                                  uint32_t nodeOffset, uint32_t nodeTotal, uint32_t* flags, hipStream_t s) {
    uint64_t n = (uint64_t)nodeTotal * K;
    if (n) k_bsphere_intersections<<<grid_for(n, 256), 256, 0, s>>>(h, spheres3, K, Tb, offset, nodeOffset, nodeTotal,
                                                                     flags);
}
void launch_bsphere_hits(uint32_t Tb, uint32_t nodeTotal, const uint32_t* scan, uint32_t* flags, uint32_t* hits,
                         hipStream_t s) {
    uint64_t n = (uint64_t)nodeTotal * Tb;
    if (n) k_bsphere_hits<<<grid_for(n, 256), 256, 0, s>>>(Tb, nodeTotal, scan, flags, hits);
}
void launch_level0_hits(const float4* h, const float4* tp, uint32_t Tb, uint32_t offset, uint32_t hitTotal,
                        uint32_t nodeOffset, uint32_t rah, uint32_t* flags, uint32_t* hits, const uint32_t* thits,
                        hipStream_t s) {
    if (hitTotal) k_level0_hits<<<grid_for(hitTotal, 256), 256, 0, s>>>(h, tp, Tb, offset, hitTotal, nodeOffset, rah,
                                                                         flags, hits, thits);
}
void launch_levelN_hits(const float4* h, const float4* tp, uint32_t offset, uint32_t hitTotal, uint32_t nodeOffset,
                        uint32_t nodeWriteTotal, uint32_t branch, uint32_t* flags, uint32_t* hits,
                        const uint32_t* thits, hipStream_t s) {
    if (hitTotal) k_levelN_hits<<<grid_for(hitTotal, 256), 256, 0, s>>>(h, tp, offset, hitTotal, nodeOffset,
                                                                         nodeWriteTotal, branch, flags, hits, thits);
}
void launch_hits_intersect(const float* rays, const uint32_t* svalues, const uint32_t* hits, const float4* tp,
                           uint32_t hitTotal, uint32_t N, uint32_t offset, uint32_t W, uint32_t Lmax, uint32_t leaf,
                           const float4* gpos, uint32_t shadow, uint32_t* out, hipStream_t s) {
    if (hitTotal) k_hits_intersect<<<grid_for(hitTotal, 256), 256, 0, s>>>(rays, svalues, hits, tp, hitTotal, N,
                                                                            offset, W, Lmax, leaf, gpos, shadow, out);
}

}  // namespace rsh
