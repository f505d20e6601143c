// oracle_api.h -- C interface of the CPU oracle (host pointers only).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  The product library never links it.
//
// Two layers:
//  * oracle_<Wrapper>(...) restate the reference's stage wrappers
//    (projects/CUDA-Engine/source/RayTracer.cu:2046-3101) on host arrays with
//    the same parameter lists; texture bindings become oracle_bind_*.
//  * oracle_render() restates the frame driver
//    (projects/CUDA-Engine/source/main.cpp:630-1054, 1234-1303) extended with
//    the configuration semantics (primary rays, gated bounces, bounce shadows,
//    octahedral keys, exact closest hit, segmented batches).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Scene description (host arrays, layouts of main.cpp:2375-2470,2596-2656).
typedef struct oracle_scene {
    uint32_t triangle_total;        // T
    const float* tri_pos;           // float4[3T]  per-vertex positions
    const float* tri_nrm;           // float4[3T]  per-vertex normals
    const int32_t* tri_mat;         // int32[3T]   material id per vertex
    uint32_t material_total;
    const float* mat_diffuse;       // float4[M]   (Kd.rgb, reflectivity)
    const float* mat_specular;      // float4[M]   (Ks.rgb, Ns)
    uint32_t object_total;          // K
    const float* object_spheres;    // float3[2K]  center | (radius, firstTri, lastTri)
    uint32_t light_total;           // L
    const float* light_pos;         // float4[L]
    const float* light_col;         // float4[L]
    const float* light_int;         // float2[L]   (diffuse, specular)
} oracle_scene;

typedef struct oracle_camera {
    float eye[3];
    float direction[3];             // Camera.cpp:120-123 (length = distance)
    float right[3];                 // Camera.cpp:125-128 (length = frame width)
    float up[3];                    // Camera.cpp:130-133 (length = frame height)
} oracle_camera;

typedef struct oracle_config {
    uint32_t width, height;
    uint32_t algorithm;             // 0 CRSH (sorted), 1 RAH (unsorted)
    uint32_t key_mode;              // 0 reference spherical bins, 1 octahedral
    uint32_t hit_mode;              // 0 exact 64-bit record, 1 reference 32-bit packing
    uint32_t cone_mode;             // 0 reference cone union, 1 guarded (no NaN nodes)
    uint32_t leaf;                  // rays per leaf node
    uint32_t branch;                // children per inner node
    uint32_t depth;                 // hierarchy levels
    uint32_t triangle_batch;        // triangles per traversal batch (5000)
    uint32_t bounces;               // mirror bounces after the primary hit
    uint32_t reflect_gate;          // 1: only materials with diffuse.w > 0 reflect
    uint32_t bounce_shadows;        // 1: shadow rays at bounce hits
    uint32_t primary_rays;          // 1: trace primary rays (else G-buffer given)
    uint32_t cell_bits, dir_bits;   // octahedral key split
    uint32_t key_layout;            // 0 type|cell|dir, 1 type|dir|cell, 2 interleaved z y x u v bits
    float scene_min[3];
    float cell_scale[3];            // cells per world unit along each axis
    uint32_t segment_shift;         // segment (view/request) id << shift
    float initial_radius, initial_spread;
    uint32_t row_begin, row_end;    // primary rays only for rows [begin, end); 0,0 = all
    uint32_t soft_shadows;          // 1: 16 shadow rays per slot from a 4x4 grid about the light
} oracle_config;

// Per-frame outputs.  Every pointer may be NULL.
typedef struct oracle_frame_out {
    uint32_t* pixels;               // W*H RGBA8 (R | G<<8 | B<<16)
    float* gbuffer_pos;             // W*H float4 primary G-buffer
    float* gbuffer_nrm;
    float* gbuffer_dif;
    float* gbuffer_spec;
    uint64_t* pass_hits[16];        // per pass: S slots; closest record or shadow flag
    uint32_t pass_count;            // filled
    uint32_t pass_kind[16];         // 0 primary, 1 shadow, 2 reflection
    uint32_t pass_slots[16];
    uint32_t pass_rays[16];         // live rays after trimming
    uint32_t pass_chunks[16];
    uint64_t pass_pairs_top[16];    // candidate (top node, triangle) pairs tested
    uint64_t pass_pairs_leaf[16];   // surviving (leaf, triangle) pairs
    uint64_t pass_ray_tests[16];    // ray-triangle tests performed
    double stage_ms[8];             // steady_clock ms per stage: create, trim, compress, sort,
                                    // decompress, hierarchy, traversal, shading
    // Optional per-pass state (capacity S slots each; nodes 8*S floats):
    uint32_t* pass_keys[16];        // slot keys as created
    uint32_t* pass_flags[16];       // slot flags (0 live, 1 empty)
    uint32_t* pass_order[16];       // slot of each sorted ray (rays entries)
    uint32_t* pass_bases[16];       // chunk first positions in the trimmed array (chunks entries)
    uint32_t* pass_ckeys[16];       // chunk keys (chunks entries)
    float* pass_nodes[16];          // hierarchy nodes, float4: sphere at 2g, cone at 2g+1
} oracle_frame_out;

int oracle_render(const oracle_scene* scene, const oracle_camera* camera,
                  const oracle_config* config, const float* gbuffer_in /* 4*W*H*4 or NULL */,
                  oracle_frame_out* out, int threads);
// A batch of views rendered as one image of sum(w*h) pixels (view i after
// views 0..i-1); config->width / height are ignored.  The view id is placed
// in the octahedral key at config->segment_shift when that is non-zero.
typedef struct oracle_view {
    oracle_camera camera;
    uint32_t width, height;
} oracle_view;
int oracle_render_views(const oracle_scene* scene, const oracle_view* views, uint32_t n,
                        const oracle_config* config, oracle_frame_out* out, int threads);
// Ray query (rsh_engine_trace): out = uint64 per ray (closest-hit record, or
// 0xFFFFFFFF occluded / 0 clear for kind 1).
int oracle_trace_rays(const oracle_scene* scene, const float* rays, const float* tmax, const uint32_t* segs,
                      uint32_t n, uint32_t kind, const oracle_config* config, uint64_t* out, int threads);

// Math and keys.
float oracle_acos(float x);
float oracle_cos(float x);
float oracle_sin(float x);
float oracle_tan(float x);
float oracle_pow(float x, float y);
uint32_t oracle_f2u_sat(float x);
uint32_t oracle_shadow_key_ref(const float* o, const float* d, uint32_t light);
uint32_t oracle_reflection_key_ref(const float* o, const float* d);
uint32_t oracle_refraction_key_ref(const float* o, const float* d);
void oracle_refract(const float* i, const float* n, float ior, float* out3);
uint32_t oracle_key_oct(const oracle_config* c, const float* o, const float* d, uint32_t type);
float oracle_ray_triangle(const float* o, const float* d, const float* v0, const float* e1, const float* e2);
// Soft shadows: the 16 grid points (float3 each) about `light` of a shadow
// ray with direction d, and the light scale of an occluded-cell mask.
void oracle_soft_grid(const float* light, const float* d, float* out48);
float oracle_soft_factor(uint32_t occluded);
int oracle_sphere_node(const float* sphere, const float* cone, const float* tri);
void oracle_triangle_sphere(const float* v0, const float* v1, const float* v2, float* out4);
void oracle_cone0(const float* cone, const float* ray, uint32_t mode, float* out4);
void oracle_coneN(const float* c1, const float* c2, float* out4);
void oracle_sphere_union(const float* s1, const float* s2, float* out4);
// Shading terms: RgbToInt (RT:100-107) and the Blinn-Phong colour of one
// light (RT:1655-1702; scaled = the reflection variant's x scale, RT:1863-1866).
uint32_t oracle_rgb_to_int(float r, float g, float b);
void oracle_light_term(const float* light_pos4, const float* light_col4, const float* light_int2, const float* P,
                       const float* N, const float* eye, const float* Kd4, const float* Ks4, float scale, int scaled,
                       float* out3);
// Key of camera ray (x, y) of a view w pixels wide: 8x8 tiles in row order,
// Morton order inside a tile, the view id at segment_shift (0: none).
uint32_t oracle_primary_key(uint32_t x, uint32_t y, uint32_t w, uint32_t view, uint32_t segment_shift);

// Traversal candidate sets (tests pin the closed form against the literal
// flag/scan transcription).  Returns the number of leaf pairs written as
// (leaf << 32 | triangle) into out (capacity cap), sorted ascending.
uint64_t oracle_leaf_pairs_literal(const float* hierarchy, const uint32_t* level_counts,
                                   uint32_t depth, uint32_t branch, const oracle_scene* s,
                                   uint32_t algorithm, uint32_t triangle_batch,
                                   uint64_t* out, uint64_t cap);
uint64_t oracle_leaf_pairs_closed(const float* hierarchy, const uint32_t* level_counts,
                                  uint32_t depth, uint32_t branch, const oracle_scene* s,
                                  uint32_t algorithm, uint32_t triangle_batch,
                                  uint64_t* out, uint64_t cap);

// ---- Reference wrapper restatements (RayTracer.cu:2046-3101) ---------------
void oracle_set_constants(uint32_t light_maximum, uint32_t hierarchy_depth,
                          uint32_t hierarchy_subdivision, uint32_t algorithm_id);
void oracle_bind_gbuffer(const float* diffuse, const float* specular,
                         const float* position, const float* normal);
void oracle_bind_scene(const oracle_scene* scene);
void oracle_bind_original(const float* tri_pos, const float* tri_nrm, const int32_t* tri_obj,
                          const float* spheres4);

void oracle_TriangleUpdateWrapper(const float* model, const float* normal, uint32_t T,
                                  float* pos_out, float* nrm_out);
void oracle_BoundingSphereUpdateWrapper(const float* translation, const float* scale,
                                        uint32_t K, float* spheres3_out);
void oracle_ShadowRayCreationWrapper(uint32_t W, uint32_t H, uint32_t L, float* rays,
                                     uint32_t* flags, uint32_t* keys, uint32_t* values);
void oracle_RefractionRayCreationWrapper(uint32_t W, uint32_t H, const float* eye, float* rays,
                                         uint32_t* flags, uint32_t* keys, uint32_t* values);
void oracle_ReflectionRayCreationWrapper(uint32_t W, uint32_t H, const float* eye, float* rays,
                                         uint32_t* flags, uint32_t* keys, uint32_t* values);
void oracle_RayTrimmingWrapper(const uint32_t* keys, const uint32_t* values, uint32_t W,
                               uint32_t H, uint32_t L, const uint32_t* flags, uint32_t* scan,
                               uint32_t* tkeys, uint32_t* tvalues, uint32_t* ray_total);
void oracle_RayCompressionWrapper(const uint32_t* tkeys, const uint32_t* tvalues, uint32_t N,
                                  uint32_t* flags, uint32_t* scan, uint32_t* bases,
                                  uint32_t* sizes, uint32_t* ckeys, uint32_t* cvalues,
                                  uint32_t* chunk_total);
void oracle_RaySortingWrapper(const uint32_t* ckeys, const uint32_t* cvalues, uint32_t C,
                              uint32_t* skeys, uint32_t* svalues);
void oracle_RayDecompressionWrapper(const uint32_t* bases, const uint32_t* sizes,
                                    const uint32_t* skeys, const uint32_t* svalues,
                                    const uint32_t* tkeys, const uint32_t* tvalues,
                                    uint32_t* skeleton, uint32_t* scan, uint32_t C,
                                    uint32_t* rkeys, uint32_t* rvalues);
void oracle_HierarchyCreationWrapper(const float* rays, const uint32_t* skeys,
                                     const uint32_t* svalues, uint32_t N, float r0, float s0,
                                     float* hierarchy);
void oracle_HierarchyTraversalWarmUpWrapper(const float* hierarchy, const float* spheres3,
                                            const float* tri_pos, uint32_t K, uint32_t N,
                                            uint32_t Tb, uint32_t offset, uint32_t* flags,
                                            uint32_t* scan, uint32_t* hits, uint32_t* thits,
                                            uint32_t* hit_total);
void oracle_HierarchyTraversalWrapper(const float* hierarchy, const float* tri_pos, uint32_t N,
                                      uint32_t Tb, uint32_t offset, uint32_t* flags,
                                      uint32_t* scan, uint32_t* hits, uint32_t* thits,
                                      uint32_t* hit_total);
void oracle_ShadowRayPreparationWrapper(uint32_t W, uint32_t H, uint32_t L, uint32_t* sflags);
void oracle_ShadowRayIntersectionWrapper(const float* rays, const uint32_t* skeys,
                                         const uint32_t* svalues, const uint32_t* hits,
                                         const float* tri_pos, uint32_t hit_total, uint32_t N,
                                         uint32_t offset, uint32_t W, uint32_t H,
                                         uint32_t* sflags);
void oracle_ShadowRayColoringWrapper(uint32_t W, uint32_t H, uint32_t L, const float* eye,
                                     const uint32_t* sflags, uint32_t* pixels);
void oracle_ReflectionRayPreparationWrapper(uint32_t W, uint32_t H, uint32_t* times);
void oracle_ReflectionRayIntersectionWrapper(const float* rays, const uint32_t* skeys,
                                             const uint32_t* svalues, const uint32_t* hits,
                                             const float* tri_pos, uint32_t hit_total,
                                             uint32_t N, uint32_t offset, uint32_t W,
                                             uint32_t H, uint32_t* times);
void oracle_ReflectionRayColoringWrapper(const float* tri_pos, const float* tri_nrm, uint32_t W,
                                         uint32_t H, uint32_t L, const float* eye,
                                         int create_rays, const uint32_t* times, float* rays,
                                         uint32_t* flags, uint32_t* keys, uint32_t* values,
                                         uint32_t* pixels);

#ifdef __cplusplus
}
#endif
