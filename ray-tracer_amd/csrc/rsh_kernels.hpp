// rsh_kernels.hpp -- kernel parameter blocks and launchers of the ray,
// surface, shading, hierarchy and traversal stages.
#pragma once
#include "fastdiv.hpp"
#include "rsh_internal.hpp"

namespace rsh {

struct KeyCfg {
    uint32_t mode;        // 0 reference spherical bins, 1 octahedral
    HashParams h;         // h.segment_shift != 0: the pixel's view id is OR-ed in at that bit
};

struct SceneView {
    const float4* tri_pos;    // [3T] per-vertex positions
    const float4* tri_nrm;    // [3T] per-vertex normals
    const int32_t* tri_mat;   // [3T] material per vertex
    const float4* mat_dif;    // [M]
    const float4* mat_spec;   // [M]
    Lights lights;
};

// Views of one batch.  A batch renders n views as one image of P = sum(w*h)
// pixels laid out view after view (view i starts at pixel off); slots then
// follow the single-image formulas with W = P, H = 1.  n == 0 is a single
// view held inline in `one` (off 0, w = W, h = H).
struct ViewDesc {
    uint32_t off, w, h, pad;
    float eye[4], dir[4], right[4], up[4];
};
struct Views {
    const ViewDesc* d;   // [n] device table, sorted by off
    uint32_t n;
    ViewDesc one;
};
__device__ __forceinline__ uint32_t view_of(const Views& V, uint32_t p) {
    if (V.n <= 1) return 0;
    uint32_t a = 0, b = V.n;        // d[a].off <= p < d[b].off
    while (b - a > 1) {
        const uint32_t m = (a + b) >> 1;
        if (V.d[m].off <= p) a = m; else b = m;
    }
    return a;
}
// Batches of up to VIEW_LDS views stage the views' pixel offsets in LDS once
// per workgroup, so that a per-pixel view lookup is a search in LDS followed
// by one (cached) load, instead of a chain of dependent global loads.  Call
// from every thread of the workgroup.
constexpr uint32_t VIEW_LDS = 64;
__device__ __forceinline__ bool stage_views(const Views& V, uint32_t* s_off) {
    const bool lds = V.n > 1 && V.n <= VIEW_LDS;
    if (lds) {
        for (uint32_t i = threadIdx.x; i < V.n; i += blockDim.x) s_off[i] = V.d[i].off;
        __syncthreads();
    }
    return lds;
}
// The view of pixel p (s_off, lds: from stage_views).
__device__ __forceinline__ uint32_t view_index(const Views& V, const uint32_t* s_off, bool lds, uint32_t p) {
    if (!lds) return view_of(V, p);
    uint32_t a = 0, b = V.n;
    while (b - a > 1) {
        const uint32_t m = (a + b) >> 1;
        if (s_off[m] <= p) a = m; else b = m;
    }
    return a;
}
__device__ __forceinline__ ViewDesc view_at(const Views& V, const uint32_t* s_off, bool lds, uint32_t p) {
    return V.n ? V.d[view_index(V, s_off, lds, p)] : V.one;
}
__device__ __forceinline__ v3 view_eye(const Views& V, uint32_t v) {
    const float* e = V.n ? V.d[v].eye : V.one.eye;
    return mk(e[0], e[1], e[2]);
}

// One view held inline (eye only, or a full pinhole camera of W x H).
inline Views single_view(uint32_t W, uint32_t H, v3 eye, v3 D = v3{0, 0, 1}, v3 R = v3{1, 0, 0},
                         v3 U = v3{0, 1, 0}) {
    Views v{};
    v.d = nullptr; v.n = 0;
    v.one.off = 0; v.one.w = W; v.one.h = H;
    const v3 src[4] = {eye, D, R, U};
    float* dst[4] = {v.one.eye, v.one.dir, v.one.right, v.one.up};
    for (int i = 0; i < 4; i++) { dst[i][0] = src[i].x; dst[i][1] = src[i].y; dst[i][2] = src[i].z; dst[i][3] = 0.0f; }
    return v;
}

struct GBufView {
    float4* pos;
    float4* nrm;
    float4* dif;
    float4* spec;
};

// Diagnostic traversal counters are spread over this many slots (summed on
// the host) so that the per-wave atomics do not serialise on one address.
// Each slot holds COUNTERS_PER_SLOT values (see CounterIndex).
constexpr uint32_t COUNTER_SLOTS = 64;
constexpr uint32_t COUNTERS_PER_SLOT = 16;
enum CounterIndex : uint32_t {
    CNT_TOP_PAIRS = 0,     // (top node, triangle) candidate pairs of the reference's culling
    CNT_LEAF_PAIRS = 1,    // (leaf, triangle) pairs passing the leaf test
    CNT_RAY_TESTS = 2,     // ray-triangle tests run
    CNT_TOP_SURV = 3,      // (top node, triangle) pairs passing the top-level test
    CNT_TRI_TESTS = 4,     // top-level triangle-sphere tests actually run
    CNT_CLUSTER_TESTS = 5, // cluster pre-tests run
    CNT_C_STEPS = 6,       // 64-ray kernel: wave steps of the ray-test stage (lane use = tests / 64 steps)
    CNT_B_STEPS = 7,       // 64-ray kernel: wave steps of the leaf-test stage
    CNT_MT_TESTS = 8       // 64-ray kernel: Moller-Trumbore tests run after the ray pre-test
};
// Triangles per cluster of the traversal's cluster level.
constexpr uint32_t CLUSTER = 64;
// Sub-clusters: SUBCLUSTER consecutive triangles of a cluster, each with its own
// sphere (the cluster's sub-clusters at SUB_PER_CLUSTER * cluster + k).
constexpr uint32_t SUBCLUSTER = 8;
constexpr uint32_t SUB_PER_CLUSTER = CLUSTER / SUBCLUSTER;
static_assert(SUB_PER_CLUSTER <= 8, "sub-cluster flags are one byte per cluster");
// A sub-cluster gets a ray mask of its own only when its radius is below
// SUB_RATIO of its cluster's.
constexpr float SUB_RATIO = 0.7f;

struct TravParams {
    const float4* nodes;          // hierarchy, sphere at 2g, cone at 2g+1
    const float2* aux;            // (cos, tan) of each node's spread
    uint32_t level_offset[4];
    uint32_t level_count[4];
    uint32_t depth, leaf, branch;
    uint32_t rays_per_top, waves_per_top;
    const float4* rec0;           // sorted rays: (origin, slot bits)
    const float4* rec1;           //              (direction, tmax)
    uint32_t N;
    const float4* tri_sphere;     // [T]
    const float4* tri_mt;         // [3T] v0, e1, e2
    uint32_t T;
    const float4* obj_sphere;     // [K] (center, radius)
    const uint4* obj_range;       // [K] (first, last, excluded lo, excluded hi)
    uint32_t K;
    uint32_t algorithm;           // 0 CRSH (object culling), 1 RAH (all triangles)
    uint32_t kind;                // 0 closest hit, 1 any hit (shadow)
    uint32_t hit_mode;            // 0 exact, 1 reference packing
    unsigned long long* closest;  // [S] per slot
    uint32_t* shadow;             // [S] per slot
    unsigned long long* counters; // optional [COUNTER_SLOTS][COUNTERS_PER_SLOT] (CounterIndex);
                                  // waves add into slot wave % COUNTER_SLOTS
    uint32_t force_generic;       // 1: use the generic kernel even for 64-ray top nodes
    uint32_t splits;              // generic kernel: waves sharing one top node's triangle range
    // work items of the 64-ray kernel: (top node, chunk of candidate clusters)
    const uint4* item_desc;       // [items] (top, list entry, first cluster in it, clusters | single << 31 | first << 30)
    const uint32_t* item_first;   // [n_top + 1] first item of each top node
    const uint32_t* cand;         // [n_top + 1] candidate clusters of each top node
    const uint32_t* obj_off;      // [n_top + 1] first list entry of each top node (CRSH)
    const uint4* obj_ent;         // passing objects of each node, in triangle order, as (first, last,
                                  // first cluster, excluded-batch start or 0) (RAH: all objects, shared)
    const uint32_t* obj_cum;      // candidate clusters of the node before each listed object
    uint32_t chunk;               // candidate clusters per item
    uint32_t items;
    // cluster level: clusters of CLUSTER consecutive triangles of one object
    const float4* cl_sphere;      // [clusters] conservative sphere over the member triangle spheres
    const float4* sub_sphere;     // [clusters * SUB_PER_CLUSTER] the same per sub-cluster
    const uint8_t* sub_use;       // [clusters] sub-clusters worth a ray mask (bit k: sub-cluster k)
    const uint32_t* obj_cl_base;  // [K] first cluster of each object (RAH: also the list's obj_cum)
    uint32_t n_clusters;
    uint32_t prefilter;           // 64-ray kernel: 0 exact tests only, 1 ray-line / triangle-sphere
                                  // pre-test, 2 that and ray / cluster-sphere masks
    // device-sized passes (ps set): N and the level sizes come from *ps, the
    // item total from item_first[n_top]; level_count / N / items are then the
    // capacities the launches are sized for, and list_cap / items_cap bound
    // the list and descriptor writes (exceeding them sets *overflow)
    const PassSizes* ps;
    uint32_t list_cap, items_cap;
    uint32_t* overflow;
    // Front-to-back camera passes (64-ray kernel, closest hit, exact
    // records): obj_order lists the objects nearest the eye first, and the
    // cull kernels walk them in that order, so each top node's cluster stream
    // runs front to back.  phase 1 runs every top node's first item (grid =
    // top nodes), phase 2 the others (grid = items): they start from the
    // records phase 1 merged and skip clusters and triangles beyond the reach
    // of all their rays.  phase 0: every item in one launch.
    const uint32_t* obj_order;    // [K] or null (index order)
    uint32_t phase;
    uint32_t line_only;           // ray pre-tests test the line only, not the part a hit can use
};

// Work items of one launch of the 64-ray kernel: 64 lanes each, so the grid
// stays below 2^31 threads.
constexpr uint32_t MAX_TRAVERSE_ITEMS = 1u << 24;

// The 64-ray kernel (depth 2, leaf * branch == 64) consumes work items; every
// other shape runs the generic kernel with `splits` waves per top node.
// Work list of a pass with n_top top nodes:
//   launch_cull_count  -> npass, cand            [n_top + 1], entry n_top = 0
//   exclusive scans    -> obj_off (of npass), item_first (of nitems)
//   launch_item_count  -> nitems = ceil(cand / chunk)
//   launch_cull_list   -> obj_ent, obj_cum at obj_off
//   launch_item_desc   -> item_desc
inline bool uses_items(const TravParams& p) {
    return p.depth == 2 && p.leaf * p.branch == 64 && !p.force_generic &&
           (p.leaf == 2 || p.leaf == 4 || p.leaf == 8 || p.leaf == 16 || p.leaf == 32);
}

// pmask: ceil(K / 64) words per top node, written by the count pass, read by the list pass
void launch_cull_count(const TravParams& p, uint32_t* npass, uint32_t* cand, unsigned long long* pmask,
                       hipStream_t s);
void launch_cull_list(const TravParams& p, uint4* obj_ent, uint32_t* obj_cum, unsigned long long* pmask,
                      hipStream_t s);
// Item totals per chunk size: table[c - 1] = sum of ceil(cand / c), c = 1..ITEM_TABLE;
// table[ITEM_TABLE + c - 1] the same for chunk 2c, table[2 ITEM_TABLE + c - 1]
// for chunk 4c; table[TABLE_WIDE] = top nodes with a wide cone (spread tangent
// outside [0, WIDE_TAN]); table[TABLE_LIST] = list entries (npass summed).
constexpr uint32_t ITEM_TABLE = 64;
constexpr uint32_t TABLE_WIDE = 3 * ITEM_TABLE;
constexpr uint32_t TABLE_LIST = 3 * ITEM_TABLE + 1;
constexpr float WIDE_TAN = 0.1f;
// max_cand: an upper bound of every cand entry (the scene's cluster count);
// table: ITEM_TABLE_WORDS entries, zeroed beforehand; npass (summed into
// table[TABLE_LIST]) may be null.  host / done / seq: see k_item_table (the table
// published to host memory by the last block; done a zeroed device word).
void launch_item_table(const uint32_t* cand, const float2* top_aux, uint32_t ntop, uint32_t max_cand,
                       const uint32_t* npass, unsigned long long* table, hipStream_t s, uint32_t* done = nullptr,
                       unsigned long long* host = nullptr, uint32_t seq = 0);
constexpr uint32_t ITEM_TABLE_WORDS = 3 * ITEM_TABLE + 2;   // 64-bit entries
void launch_item_count(const uint32_t* cand, uint32_t chunk, uint32_t ntop, uint32_t* nitems, hipStream_t s);
// CRSH: the list pass and the work-item descriptors in one launch (needs
// obj_off, item_first and chunk in p)
void launch_cull_items(const TravParams& p, uint4* obj_ent, uint32_t* obj_cum, const unsigned long long* pmask,
                       uint4* desc, hipStream_t s);
void launch_item_desc(const TravParams& p, uint4* desc, hipStream_t s);

// --- rays ---
// Ray creation: F = uint8_t (the engine's slot flags, one byte per slot) or
// uint32_t (the stage API's); values == nullptr leaves the values unwritten
// (the engine's trim, run_pack, takes each live slot's index as its value).
template <typename F>
void launch_create_primary(const Views& V, uint32_t P, KeyCfg kc, float* rays, F* flags, uint32_t* keys,
                           uint32_t* values, hipStream_t s);
template <typename F>
void launch_create_shadow(const float4* gpos, const float4* gnrm, uint32_t W, uint32_t H, uint32_t Lmax,
                          const float4* light_pos, KeyCfg kc, const Views& V, uint32_t soft, float* rays, F* flags,
                          uint32_t* keys, uint32_t* values, hipStream_t s);
template <typename F>
void launch_create_reflection(const float4* gpos, const float4* gnrm, const float4* gdif, uint32_t W, uint32_t H,
                              uint32_t Lmax, const Views& V, uint32_t gate, uint32_t mark_gaps, KeyCfg kc,
                              float* rays, F* flags, uint32_t* keys, uint32_t* values, hipStream_t s);
void launch_create_primary_sorted(const Views& V, uint32_t P, float4* rec0, float4* rec1, hipStream_t s);
void launch_create_refraction(const float4* gpos, const float4* gnrm, uint32_t W, uint32_t H, uint32_t Lmax,
                              const Views& V, KeyCfg kc, float* rays, uint32_t* flags, uint32_t* keys,
                              uint32_t* values, hipStream_t s);
template <typename F>
void launch_create_bounce(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, const unsigned long long* rec,
                          uint32_t hit_mode, uint32_t gate, KeyCfg kc, const Views& V, float* rays, F* flags,
                          uint32_t* keys, uint32_t* values, hipStream_t s);
// rays == nullptr: the camera pass of views V, whose rays are recomputed
// from the views (k_create_primary_sorted writes no slot-order rays).
void launch_surface(const float* rays, const unsigned long long* rec, uint32_t hit_mode, uint32_t W, uint32_t H,
                    uint32_t slot_stride, uint32_t normalize_n, SceneView sc, GBufView g, hipStream_t s,
                    const Views* V = nullptr);
void launch_shade_primary(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, Lights L, const Views& V,
                          const uint32_t* sflags, uint32_t soft, uint32_t* pixels, hipStream_t s,
                          const uint32_t* live = nullptr);
void launch_shade_reflection(GBufView g, uint32_t W, uint32_t H, uint32_t Lmax, Lights L, const Views& V,
                             const unsigned long long* rec, uint32_t hit_mode, const uint32_t* sflags,
                             uint32_t soft, uint32_t* pixels, hipStream_t s);
void launch_triangle_prep(const float4* tri_pos, uint32_t T, float4* tri_sphere, float4* tri_mt, hipStream_t s);
void launch_sub_use(const float4* cl_sphere, const float4* sub_sphere, uint32_t nc, uint8_t* sub_use, hipStream_t s);
void launch_cluster_build(const float4* tri_sphere, const uint2* range, uint32_t n, float4* cl_sphere,
                          hipStream_t s);
// reference-only stages
void launch_update_vertex(const float* model, const float* normal, uint32_t vertexTotal, const float4* pos_in,
                          const float4* nrm_in, const int32_t* obj_ids, float4* pos_out, float4* nrm_out,
                          hipStream_t s);
void launch_update_sphere(const float* translation, const float* scale, uint32_t K, const float4* spheres_in,
                          float* spheres_out, hipStream_t s);
void launch_color_reflection_ref(const float4* tp, const float4* tn, uint32_t W, uint32_t H, uint32_t Lmax,
                                 SceneView sc, v3 eye, uint32_t create, const uint32_t* times, float* rays,
                                 uint32_t* flags, uint32_t* keys, uint32_t* values, uint32_t* pixels, hipStream_t s);

// --- hierarchy / traversal ---
// rays == nullptr: a frame's shadow pass, each ray recomputed from its slot's
// surface point (shadow_pos) and light (light_pos), as ray creation made it.
void launch_gather(const float* rays, const uint32_t* sorted_values, uint32_t N, const PassSizes* ps,
                   const float4* shadow_pos,
                   const float* tmax, uint32_t W, uint32_t Lmax, uint32_t slot_shift, float4* rec0, float4* rec1,
                   hipStream_t s, const float4* light_pos = nullptr);
// Ray queries: key, value (= index) and live flag of n caller rays (float[6n]);
// segs (optional) places a per-ray segment id at kc.h.segment_shift.
template <typename F>
void launch_query_keys(const float* rays, const uint32_t* segs, uint32_t n, KeyCfg kc, uint32_t type, F* flags,
                       uint32_t* keys, uint32_t* values, hipStream_t s);
void launch_hierarchy(const float4* rec0, const float4* rec1, const PassSizes* ps, const float* rays,
                      const uint32_t* sorted_values,
                      uint32_t N, uint32_t leaf, uint32_t branch, uint32_t depth, float r0, float s0, uint32_t mode,
                      float4* nodes, float2* aux, const uint32_t* counts, const uint32_t* offsets, hipStream_t s);
hipError_t launch_traverse(const TravParams& p, hipStream_t s);
void launch_bsphere_intersections(const float4* h, const float* spheres3, uint32_t K, uint32_t Tb, uint32_t offset,
                                  uint32_t nodeOffset, uint32_t nodeTotal, uint32_t* flags, hipStream_t s);
void launch_bsphere_hits(uint32_t Tb, uint32_t nodeTotal, const uint32_t* scan, uint32_t* flags, uint32_t* hits,
                         hipStream_t s);
void launch_level0_hits(const float4* h, const float4* tp, uint32_t Tb, uint32_t offset, uint32_t hitTotal,
                        uint32_t nodeOffset, uint32_t rah, uint32_t* flags, uint32_t* hits, const uint32_t* thits,
                        hipStream_t s);
void launch_levelN_hits(const float4* h, const float4* tp, uint32_t offset, uint32_t hitTotal, uint32_t nodeOffset,
                        uint32_t nodeWriteTotal, uint32_t branch, uint32_t* flags, uint32_t* hits,
                        const uint32_t* thits, hipStream_t s);
void launch_hits_intersect(const float* rays, const uint32_t* svalues, const uint32_t* hits, const float4* tp,
                           uint32_t hitTotal, uint32_t N, uint32_t offset, uint32_t W, uint32_t Lmax, uint32_t leaf,
                           const float4* gpos, uint32_t shadow, uint32_t* out, hipStream_t s);

// Hierarchy level sizes: level 0 has ceil(N/leaf) nodes, level k
// ceil(count[k-1]/branch); offsets are prefix sums.  Returns the node total.
inline uint32_t level_sizes(uint32_t N, uint32_t leaf, uint32_t branch, uint32_t depth, uint32_t* counts,
                            uint32_t* offsets) {
    uint32_t n = (N + leaf - 1) / leaf, total = 0;
    for (uint32_t k = 0; k < depth; k++) {
        counts[k] = n; offsets[k] = total; total += n;
        n = (n + branch - 1) / branch;
    }
    return total;
}

}  // namespace rsh
