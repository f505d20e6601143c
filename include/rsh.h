/* rsh.h -- native C ABI of the MI355X ray-space-hierarchy engine (gfx950).
 *
 * The engine owns device-resident scenes and per-pass scratch and runs a whole
 * frame (primary rays, shadow rays, gated mirror bounces, shading) on one HIP
 * stream.  It replaces the per-frame driver of the reference
 * (projects/CUDA-Engine/source/main.cpp:630-1054 createShadowRays / castRays /
 * colorShadowRays / createReflectionRays / colorReflectionRays, and the frame
 * loop main.cpp:1234-1303), and the OpenGL G-buffer (main.cpp:1082-1108) by a
 * primary-ray stage.  The stage-by-stage drop-in for the reference's
 * extern "C" wrappers is include/rsh_engine.h.
 *
 * Plain C types only: device pointers are void*, streams are void*
 * (hipStream_t), sizes are explicit.  Every call returns an rsh_status; the
 * message of the last failure is rsh_last_error().
 */
#ifndef RSH_H_
#define RSH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rsh_status {
    RSH_OK = 0,
    RSH_INVALID_ARG = 1,
    RSH_CAPACITY = 2,
    RSH_HIP_ERROR = 3,
    RSH_INTERNAL = 4
} rsh_status;

/* Scene in the reference's upload layout (main.cpp:2375-2470, :2596-2656).
 * Host pointers; the engine copies them to device memory. */
typedef struct rsh_scene_desc {
    uint32_t triangle_total;        /* T */
    const float* tri_pos;           /* float4[3T] per-vertex position          */
    const float* tri_nrm;           /* float4[3T] per-vertex normal            */
    const int32_t* tri_mat;         /* int32[3T]  material id per vertex        */
    uint32_t material_total;        /* M */
    const float* mat_diffuse;       /* float4[M]  (Kd.rgb, reflectivity)       */
    const float* mat_specular;      /* float4[M]  (Ks.rgb, Ns)                 */
    uint32_t object_total;          /* K */
    const float* object_spheres;    /* float3[2K] center | (radius, first, last); objects tile [0,T) in order */
    uint32_t light_total;           /* L */
    const float* light_pos;         /* float4[L] */
    const float* light_col;         /* float4[L] */
    const float* light_int;         /* float2[L] (diffuse, specular intensity) */
} rsh_scene_desc;

/* Pinhole basis of Camera::loadView (projects/CUDA-Engine/source/Camera.cpp:108-133);
 * primary ray (x, y) = normalize(direction + right*(x/W-0.5) + up*(y/H-0.5)). */
typedef struct rsh_camera {
    float eye[3];
    float direction[3];
    float right[3];
    float up[3];
} rsh_camera;

enum { RSH_CRSH = 0, RSH_RAH = 1 };
enum { RSH_KEY_REF = 0, RSH_KEY_OCT = 1 };
enum { RSH_HIT_EXACT = 0, RSH_HIT_REF = 1 };
enum { RSH_CONE_REF = 0, RSH_CONE_GUARDED = 1 };

/* Frame configuration.  Defaults of the reference (Constants.h:9-29):
 * leaf = branch = 8 (HIERARCHY_SUBDIVISION), depth = 2, triangle_batch = 5000. */
typedef struct rsh_frame_config {
    uint32_t width, height;
    uint32_t algorithm;             /* RSH_CRSH sorted hierarchy, RSH_RAH unsorted       */
    uint32_t key_mode;              /* RSH_KEY_REF spherical bins, RSH_KEY_OCT octahedral */
    uint32_t hit_mode;              /* RSH_HIT_EXACT 64-bit record, RSH_HIT_REF 32-bit   */
    uint32_t cone_mode;             /* RSH_CONE_REF reference union, RSH_CONE_GUARDED    */
    uint32_t leaf;                  /* rays per leaf node                                */
    uint32_t branch;                /* children per inner node                           */
    uint32_t depth;                 /* hierarchy levels (1..4)                           */
    uint32_t triangle_batch;        /* batch size whose boundary case the culling keeps  */
    uint32_t bounces;               /* mirror bounces after the primary hit              */
    uint32_t reflect_gate;          /* 1: only diffuse.w > 0 materials reflect           */
    uint32_t bounce_shadows;        /* 1: shadow rays at bounce hits                     */
    uint32_t primary_rays;          /* must be 1 (the G-buffer comes from primary rays)  */
    uint32_t cell_bits, dir_bits;   /* octahedral key: 3*cell_bits + 2*dir_bits + 2 <= 32 */
    uint32_t key_layout;            /* 0 type|cell|dir, 1 type|dir|cell, 2 type|bits of
                                       cell z,y,x and direction u,v interleaved          */
    float scene_min[3];
    float cell_scale[3];            /* cells per world unit                              */
    uint32_t segment_shift;         /* batches: key bit of the view id (0 = not in key) */
    float initial_radius, initial_spread;
    uint32_t traversal_kernel;      /* 0 automatic, 1 generic kernel for every shape     */
    uint32_t traversal_splits;      /* generic kernel: waves per top node; 0 = fill the GPU */
    uint32_t traversal_chunk;       /* 64-ray kernel: candidate clusters (64 consecutive
                                       triangles of one object) per work item; 0 = sized
                                       per pass                                           */
    uint32_t traversal_prefilter;   /* 64-ray kernel: ray pre-tests before the exact test.
                                       1 none; 2 each ray's line against the triangle's
                                       bounding sphere (except for a frame's camera
                                       rays, also the sphere's parameter range against
                                       the part of the ray a hit can still use: beyond
                                       EPS, short of a shadow ray's length or the closest
                                       hit so far); 3 that and, per passing cluster,
                                       the rays' lines against the cluster's sphere (for
                                       spread-out rays); 0 automatic: 2 for a frame's
                                       camera rays; otherwise 3 when more than a
                                       sixteenth of the top nodes' cones have tan(spread)
                                       outside [0, 0.1], else 2.  Hit records
                                       are the same in every mode; with mode 3 and stats
                                       counters (flag bit 1) the cluster masks are off so
                                       that the counts stay the reference's.             */
    uint32_t soft_shadows;          /* 1: soft shadows (RayTracer.cu:1550-1603, 1680-1697,
                                       compiled out by Constants.h:32): 16 rays per shadow
                                       slot from a 4x4 grid about the light, each an exact
                                       any-hit query; shading scales both light terms by
                                       log2(1 + 2 s), s the grid weights of the open cells */
} rsh_frame_config;

#define RSH_MAX_PASSES 16

typedef struct rsh_frame_stats {
    uint32_t pass_count;
    uint32_t pass_kind[RSH_MAX_PASSES];      /* 0 primary, 1 shadow, 2 reflection */
    uint32_t pass_slots[RSH_MAX_PASSES];
    uint32_t pass_rays[RSH_MAX_PASSES];      /* live rays after trimming */
    uint32_t pass_chunks[RSH_MAX_PASSES];    /* packets after compression */
    uint64_t pass_pairs_top[RSH_MAX_PASSES]; /* (top node, triangle) pairs tested */
    uint64_t pass_pairs_leaf[RSH_MAX_PASSES];/* surviving (leaf, triangle) pairs */
    uint64_t pass_ray_tests[RSH_MAX_PASSES]; /* ray-triangle tests */
    float stage_ms[8];                       /* create, trim, compress, sort, decompress,
                                                hierarchy, traversal, shading (summed) */
    uint32_t rays_total;
    uint64_t pass_top_survivors[RSH_MAX_PASSES]; /* top-level (node, triangle) pairs passing */
    float pass_traverse_ms[RSH_MAX_PASSES];      /* traversal kernel time per pass (bit 2 or 4) */
    uint32_t pass_splits[RSH_MAX_PASSES];        /* generic kernel: waves per top node       */
    uint32_t pass_items[RSH_MAX_PASSES];         /* 64-ray kernel: work items launched       */
    uint32_t pass_chunk[RSH_MAX_PASSES];         /* 64-ray kernel: candidate clusters per item */
    uint64_t pass_tri_tests[RSH_MAX_PASSES];     /* top-level triangle tests actually run (bit 1) */
    uint64_t pass_cluster_tests[RSH_MAX_PASSES]; /* cluster pre-tests run (bit 1)            */
    uint64_t pass_leaf_steps[RSH_MAX_PASSES];    /* 64-ray kernel: wave steps of leaf tests (bit 1) */
    uint64_t pass_ray_steps[RSH_MAX_PASSES];     /* 64-ray kernel: wave steps of ray tests (bit 1)  */
    uint32_t graph_replayed;                     /* 1: the call replayed a captured graph (bit 3) */
    uint64_t pass_mt_tests[RSH_MAX_PASSES];      /* Moller-Trumbore tests run after the ray pre-test
                                                    (bit 1; equals pass_ray_tests without it)   */
    uint32_t pass_mode[RSH_MAX_PASSES];          /* 64-ray kernel mode used: 0 exact tests only,
                                                    1 triangle pre-test, 2 and cluster masks   */
} rsh_frame_stats;

typedef struct rsh_engine rsh_engine;

/* Engine lifetime.  `device` is the HIP device ordinal. */
rsh_status rsh_engine_create(rsh_engine** out, int device);
void rsh_engine_destroy(rsh_engine* e);
const char* rsh_last_error(const rsh_engine* e);
rsh_status rsh_engine_set_stream(rsh_engine* e, void* hip_stream);
/* Deadline of each host wait of a frame for a pass size the device publishes
 * (default 60000 ms).  Past it the call fails with RSH_INTERNAL and a message
 * instead of spinning on a hung device or a blocked caller stream.  No
 * counterpart in the reference, whose blocking cudaMemcpy read-backs
 * (RayTracer.cu:2316,2364,2649,2720,2832) wait without a bound. */
rsh_status rsh_engine_set_wait_timeout(rsh_engine* e, uint32_t ms);

/* Upload a scene (host arrays) and precompute per-triangle data. */
rsh_status rsh_engine_load_scene(rsh_engine* e, const rsh_scene_desc* scene);

/* Move the loaded scene's objects, as the reference does every frame
 * (main.cpp:1131-1196 driving RayTracer.cu:448-570): vertex v of object k
 * becomes model_k * (x, y, z, w) of its position as loaded (w: the object id
 * afterwards, as the reference stores it), its normal normal_k * (nx, ny, nz, 0),
 * normalised; object k's sphere centre model_k * (c, 1) and radius r times the
 * largest |diagonal entry| of scale_k's upper 3x3.  Host arrays of 16 floats per
 * object, row-major (entry (i, j) at 4i + j).  Every call starts from the scene
 * as loaded.  The per-triangle, cluster and sub-cluster data are rebuilt on
 * the device and captured frame graphs are dropped. */
rsh_status rsh_engine_update_objects(rsh_engine* e, const float* model, const float* normal, const float* scale);

/* Object bounding spheres for a scene (host code, no device needed): the
 * smallest ball enclosing each object's vertices (BoundingSphere.cpp:10-62,
 * the reference's Gaertner Miniball; here Welzl's move-to-front algorithm in
 * double precision).  Object k spans triangles ranges[2k]..ranges[2k+1] of
 * tri_pos (float4 per vertex, 3 per triangle).  out: float3[2K] in the
 * reference's layout (centre | radius, first, last), the form
 * rsh_scene_desc.object_spheres takes.  mode RSH_SPHERE_MINIMAL: centre
 * rounded to float, radius the largest vertex distance from it, rounded up;
 * RSH_SPHERE_REFERENCE: the reference's radius min(1e5, max(r^4, largest
 * vertex distance)), r^2 the ball's squared radius (its quirk Q12). */
enum { RSH_SPHERE_MINIMAL = 0, RSH_SPHERE_REFERENCE = 1 };
rsh_status rsh_object_spheres(const float* tri_pos, uint32_t triangle_total, const uint32_t* ranges,
                              uint32_t object_total, uint32_t mode, float* out);

/* Render one frame into the engine's pixel buffer.  The shadow passes run on
 * a second stream of the engine (and, eagerly, a host thread of their own),
 * concurrently with the primary / reflection passes on the engine's stream;
 * the call returns when both are done.  `flags`:
 *   bit 0  keep every pass's per-slot hit records for rsh_engine_read_pass
 *   bit 1  collect traversal counters (pairs / tests) into the stats
 *   bit 2  time stages with HIP events into stats->stage_ms
 *   bit 4  time only the traversal kernel launches (stats->pass_traverse_ms and
 *          stage_ms[6]), two events per pass instead of two per stage
 *   bit 5  serial passes: each shadow pass is issued as soon as its hits are,
 *          one pass at a time (the timings of bits 2 and 4 are then free of
 *          the other chain's kernels)
 *   bit 6  keep every pass's keys, sort order, packets and hierarchy nodes for
 *          rsh_engine_read_pass_order (a debug read-back for parity checks)
 *   bit 3  graph replay (alone, without bits 0-2, 4, 6): the first call for a given
 *          set of view sizes, configuration and scene renders eagerly and
 *          captures a HIP graph of the frame; later calls replay it with no
 *          host read-back inside.  CRSH frames with the 64-ray traversal
 *          (depth 2, leaf x branch = 64) keep every pass's sizes in device
 *          memory, with launches sized by capacities from the eager frame: the
 *          graph replays for any cameras of those sizes, and the stats carry
 *          the replay's own ray counts.  A replay whose work list outgrows the
 *          capacities writes nothing past them and the call renders that frame
 *          eagerly instead.  Other configurations fix every pass's sizes to the
 *          eager frame's and replay only for the same cameras.  The graph is
 *          dropped when any device buffer of the library is reallocated, the
 *          scene is reloaded or the view sizes or configuration change.  Needs
 *          a non-null stream (the engine's own by default). */
rsh_status rsh_engine_render(rsh_engine* e, const rsh_camera* cam, const rsh_frame_config* cfg,
                             uint32_t flags, rsh_frame_stats* stats);

/* One view of a batch. */
typedef struct rsh_view {
    rsh_camera camera;
    uint32_t width, height;
} rsh_view;

/* Render n views as one batch: every pass traces the rays of all views
 * together (one trim, one sort, one hierarchy, one traversal launch).  The
 * pixels of view i follow those of views 0..i-1 (rsh_engine_read_pixels
 * reads sum(w*h) values).  cfg->width / height are ignored.  With
 * cfg->segment_shift = s > 0 the view id is placed in key bits s..31 (the
 * octahedral key must fit below s, and n <= 2^(32-s)), so the sort orders
 * rays by view first and hierarchy nodes rarely mix views.  Flags as for
 * rsh_engine_render.  Replaces running the frame driver once per view
 * (main.cpp:1234-1303). */
rsh_status rsh_engine_render_views(rsh_engine* e, const rsh_view* views, uint32_t n,
                                   const rsh_frame_config* cfg, uint32_t flags, rsh_frame_stats* stats);

/* Ray query: trace n caller rays against the loaded scene through the same
 * pipeline (octahedral or reference key, trim, compress, sort, decompress,
 * hierarchy, traversal).  Device buffers:
 *   rays      float[6n]: origin xyz, direction xyz per ray (the T1 layout)
 *   tmax      float[n], any hit only: a hit counts when eps < t < tmax
 *   segments  optional uint32[n]: id placed in key bits cfg->segment_shift..31
 *             (merged requests are sorted apart; octahedral key below the shift)
 *   out       kind 0 closest hit: uint64[n] (hit_mode packing, miss = all ones
 *             in EXACT mode, 0xFFFFFFFF in REF mode); kind 1 any hit: uint32[n],
 *             0xFFFFFFFF occluded, 0 clear.
 * Only cfg's hierarchy, key, hit and culling fields are used; t is measured
 * from origin + direction * 0.01 as for frame rays.  flags: bits 1, 2. */
rsh_status rsh_engine_trace(rsh_engine* e, const float* rays, const float* tmax, const uint32_t* segments,
                            uint32_t n, uint32_t kind, const rsh_frame_config* cfg, void* out,
                            uint32_t flags, rsh_frame_stats* stats);

rsh_status rsh_engine_read_pixels(rsh_engine* e, uint32_t* host_rgba, size_t count);
/* The frame's pixels copied to host memory without waiting: the copy is
 * queued behind the frame on a copy stream of the engine and the call returns
 * at once; rsh_engine_wait_pixels waits for it.  The next frame clears its
 * pixels only once the copy has ended (right before its first shading), so the
 * copy runs beside that frame's first passes.  host_rgba must stay valid, and
 * is not to be read, until the wait returns; page-locked memory
 * (rsh_host_alloc) makes the copy one DMA transfer.  A graph-replayed frame
 * (flag bit 3) first waits for a pending copy. */
rsh_status rsh_engine_read_pixels_async(rsh_engine* e, uint32_t* host_rgba, size_t count);
rsh_status rsh_engine_wait_pixels(rsh_engine* e);
/* Primary G-buffer: diffuse, specular, position, normal, each float4[W*H]. */
rsh_status rsh_engine_read_gbuffer(rsh_engine* e, float* host_dif, float* host_spec,
                                   float* host_pos, float* host_nrm, size_t count);
/* Per-slot records of pass i (closest: exact or reference packing; shadow:
 * 0 or 0xFFFFFFFF), widened to uint64.  Needs flag bit 0 at render. */
rsh_status rsh_engine_read_pass(rsh_engine* e, uint32_t pass, uint64_t* host, size_t count);
/* Pass i's hash / sort / compress state, kept with flag bit 6 (host buffers,
 * any may be NULL): the slot keys and flags as ray creation wrote them
 * (slots; none for camera rays created in key order), the sorted order
 * (for each of the rays sorted rays, its slot: the permutation the
 * hierarchy is built over), the packets as compression found them (chunk
 * first positions in the trimmed array and chunk keys; CRSH passes that
 * compress), and the hierarchy nodes (node_count float4: sphere at 2g, cone
 * at 2g+1, level 0 first).  The reference's counterparts are the buffers
 * its wrappers fill (RayTracer.cu:2276-2535). */
rsh_status rsh_engine_read_pass_order(rsh_engine* e, uint32_t pass, uint32_t* keys, uint32_t* flags, size_t slots,
                                      uint32_t* order, size_t rays, uint32_t* bases, uint32_t* ckeys, size_t chunks,
                                      float* nodes, size_t node_count);
/* Device pointer of the pixel buffer (uint32[W*H], RGBA8 R|G<<8|B<<16). */
void* rsh_engine_pixels_device(rsh_engine* e);

/* ------------------------------------------------------------- serving --
 * A request loop over device-resident scenes (no counterpart in the
 * reference, whose GLUT loop renders one frame at a time, main.cpp:1234-1303).
 * Each scene gets its own engine and HIP stream and one worker thread.
 * Requests (ray queries, as rsh_engine_trace) are admitted in arrival order
 * while the sum of in-flight requests' device bytes (256 per ray) stays within
 * the budget; a scene's worker merges the next slice (rsh_server_set_slice)
 * of every admitted request of the head request's kind (up to max_batch_rays
 * rays and 64 slices) into one traced batch; a request's slice must fit
 * max_batch_rays.  Each slice's index in the batch is placed above the 26-bit octahedral
 * key so the sort keeps requests apart.  Requests arriving while a batch runs
 * join the next one.  Host buffers passed to submit must stay valid until
 * rsh_server_wait returns for that request (pinned memory from
 * rsh_host_alloc copies fastest). */
typedef struct rsh_server rsh_server;
typedef struct rsh_request_info {
    double queue_ms;         /* submit -> start of the first batch that traced it */
    double latency_ms;       /* submit -> results in the caller's buffer           */
    uint32_t batch_requests; /* most request slices in one batch that traced it    */
} rsh_request_info;
typedef struct rsh_server_info {
    uint64_t batches, requests, rays;
    double busy_ms;          /* summed over scenes (scenes overlap) */
    uint32_t pending;        /* waiting for budget */
    uint64_t admitted_bytes;
} rsh_server_info;

rsh_status rsh_server_create(rsh_server** out, int device, uint64_t hbm_budget_bytes, uint32_t max_batch_rays);
/* cfg: hierarchy / key / hit fields used for the scene's queries; the
 * octahedral key must fit 26 bits.  Before the first submit. */
rsh_status rsh_server_add_scene(rsh_server* s, const rsh_scene_desc* scene, const rsh_frame_config* cfg,
                                uint32_t* scene_id);
/* kind 0 closest hit (host_out uint64[n]); kind 1 any hit with host_tmax
 * (host_out uint32[n]).  Non-blocking. */
rsh_status rsh_server_submit(rsh_server* s, uint32_t scene_id, uint32_t kind, const float* host_rays,
                             const float* host_tmax, uint32_t n, void* host_out, uint64_t* request_id);
/* Block until the request has completed; info may be NULL.  Each request is
 * waited for once. */
rsh_status rsh_server_wait(rsh_server* s, uint64_t request_id, rsh_request_info* info);
/* Largest slice of one request traced in one batch (default 2^20 rays): a
 * larger request is traced over several batches, each shared with the
 * requests queued meanwhile, so small requests do not wait behind all of a
 * large one.  Each ray's result does not depend on its batch. */
rsh_status rsh_server_set_slice(rsh_server* s, uint32_t slice_rays);
rsh_status rsh_server_stats(rsh_server* s, rsh_server_info* info);
const char* rsh_server_last_error(const rsh_server* s);
/* Stops the workers after the admitted requests; wait for every request first. */
void rsh_server_destroy(rsh_server* s);
void* rsh_host_alloc(size_t bytes);
void rsh_host_free(void* p);

/* Device self-check (no counterpart in the reference; used by the test
 * suite): compares the kernels' correctly rounded reciprocal (one Newton step
 * on the hardware estimate) with IEEE division over every float in its
 * estimate range.  Returns a HIP error code; *mismatches must be 0. */
int rsh_selfcheck_rcp(int device, uint64_t* tested, uint64_t* mismatches);
/* Self-check of the bounded host wait (no counterpart in the reference; used
 * by the test suite): waits with a deadline of wait_ms for a host word that is
 * never published, on a stream a host function keeps busy for block_ms.
 * Returns RSH_INTERNAL when the deadline ended the wait (expected for
 * block_ms > wait_ms), RSH_HIP_ERROR when the stream went idle first. */
rsh_status rsh_selfcheck_wait_deadline(uint32_t wait_ms, uint32_t block_ms, uint32_t* elapsed_ms);

/* Device memory helpers (plain HIP calls, for callers without a HIP runtime
 * binding of their own). */
void* rsh_device_alloc(size_t bytes);
void rsh_device_free(void* p);
int rsh_memcpy_htod(void* dst, const void* src, size_t bytes);
int rsh_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int rsh_memset_d32(void* dst, uint32_t value, size_t count);
int rsh_device_synchronize(void);
int rsh_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* RSH_H_ */
