// engine.hip -- native frame driver of the RSH pipeline on one HIP stream.
//
// Restates the reference's per-frame orchestration
// (projects/CUDA-Engine/source/main.cpp:630-1054 createShadowRays, castRays,
// colorShadowRays, createReflectionRays, colorReflectionRays; frame loop
// main.cpp:1234-1303) with these MI355X-side changes:
//  * the scene stays resident; per-triangle bounding spheres and edge
//    vectors are computed once at load instead of once per candidate pair;
//  * a pass reads back two scalars (live rays, packets) to size its
//    launches, where the reference blocks on a readback after every scan
//    of every triangle batch (RayTracer.cu:2316,2364,2649,2720,2832);
//  * traversal and intersection run as one fused kernel per pass over a
//    work list of (top node, chunk of candidate clusters) items built by an
//    object-culling kernel, with no triangle batching and no candidate pair
//    arrays;
//  * the OpenGL G-buffer is produced by a primary-ray pass through the same
//    pipeline.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rsh.h"
#include "rsh_kernels.hpp"

using namespace rsh;

// Graph replay (render flag bit 3).  An eager frame records every pass's
// sizes (the values the host otherwise reads back); a later run replays
// those sizes without read-backs or synchronisation under stream capture.
// Camera rays' pre-tests test the line only (the range test prunes little
// where every ray starts at the eye): 1 in every launch, 2 in the front-to-back
// pass's first launch only (the second starts from the first's hits), 0 never.
#ifndef RSH_CAMERA_LINE_ONLY
#define RSH_CAMERA_LINE_ONLY 1
#endif
// A single view's camera pass runs front to back (trace_pass), in two phases
// (RSH_CAMERA_PHASES) or in one; with RSH_SHADOW_FRONT so do a single light's
// shadow passes, ordered from the light (off: on the 1080p atrium the shadow
// passes run 0.86 -> 1.05, 0.54 -> 0.72 ms and the frame 5.44 -> 5.65 ms,
// while the 64-view batch gains 1 %).  1080p atrium, camera-pass traversal alone /
// concurrent frame: index order 0.87 / 5.55 ms, front to back in two phases
// 0.80 / 5.55, in one 0.86 / 5.62; the camera rays' line-only pre-test 0.80
// against 0.83 ms with the range test.
#ifndef RSH_CAMERA_FRONT
#define RSH_CAMERA_FRONT 1
#endif
#ifndef RSH_CAMERA_PHASES
#define RSH_CAMERA_PHASES 1
#endif
#ifndef RSH_SHADOW_FRONT
#define RSH_SHADOW_FRONT 0
#endif
// Traversal mode of a frame's camera pass (TravParams::prefilter), and with
// RSH_SHADOW_MODE1 every shadow pass in mode 1 (1080p atrium frame 6.30 against
// 5.78 ms with the choice below).
#ifndef RSH_CAMERA_MODE
#define RSH_CAMERA_MODE 1u
#endif
#ifndef RSH_SHADOW_MODE1
#define RSH_SHADOW_MODE1 0
#endif
// Mode 2 when more than 1/RSH_WIDE_DIV of a pass's top nodes are wide.
#ifndef RSH_WIDE_DIV
#define RSH_WIDE_DIV 16
#endif

struct PassPlan { uint32_t N = 0, C = 0, chunk = 0, list_total = 0, items = 0, mode = 0; };

// One pass pipeline's stream and scratch.  A frame runs its shadow passes on a
// second lane, concurrently with the primary / reflection chain of the first,
// so one chain's latency-bound steps (read-backs, one-workgroup scans, the
// tail of a traversal launch) overlap the other's work.
struct Lane {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    size_t cap_slots = 0;
    DevBuf flags, keys, values, tvalues, bases, ckeys, cvalues, skeys, svalues;
    DevBuf rvalues, soff, rec0, rec1, nodes, aux, counters, dcounts;
    // front-to-back passes: the objects nearest the pass's common ray origin
    // first, on the device and its page-locked host staging (K words)
    DevBuf obj_order;
    uint32_t* ord_host = nullptr;
    uint32_t* hstage = nullptr;        // page-locked host words for the per-pass read-backs
    uint32_t* hsize = nullptr;         // coherent page-locked words the pack scan writes: seq, N, C
    uint32_t* hsize_dev = nullptr;     // their device address
    uint32_t hseq = 0;                 // sequence number of the last pack
    unsigned long long* htable = nullptr;      // coherent page-locked: seq, then the item table
    unsigned long long* htable_dev = nullptr;  // its device address
    uint32_t tseq = 0;                         // sequence number of the last item table
    DevBuf w_done;                             // block counter of k_item_table (kept at zero)
    hipEvent_t rb_ev = nullptr;        // marks a read-back's completion (polled)
    uint32_t wait_ms = 60000;          // deadline of one host wait for a read-back or a published word
    bool table_clean = false;          // w_table is zero (the item scan clears it after each read-back)
    // work list of the 64-ray traversal kernel
    DevBuf w_npass, w_cand, w_nitems, w_off, w_first, w_list, w_cum, w_items, w_table, w_pmask;
    DevBuf psz;                        // PassSizes of a device-sized pass
    ScanScratch sc;
    std::vector<PassPlan> plan;        // sizes of this lane's passes (recorded / replayed in order)
    size_t plan_pos = 0;
    // timing (flag bits 2 and 4)
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timed;
    size_t ev_used = 0;
    std::vector<DevBuf*> bufs() {
        return {&flags, &keys, &values, &tvalues, &bases, &ckeys, &cvalues, &skeys, &svalues, &rvalues, &soff,
                &rec0, &rec1, &nodes, &aux, &counters, &dcounts, &w_npass, &w_cand, &w_nitems, &w_off, &w_first,
                &w_list, &w_cum, &w_items, &w_table, &w_pmask, &psz, &obj_order};
    }
};

struct rsh_engine {
    int device = 0;
    hipStream_t stream = nullptr;      // the caller's stream (lane 0)
    bool own_stream = false;
    // rsh_engine_read_pixels_async: its copy stream, the frame's end, the
    // copy's end, and whether a copy may still run
    hipStream_t copy_stream = nullptr;
    hipEvent_t frame_done = nullptr, copy_done = nullptr;
    bool copy_pending = false;
    std::string err;

    // scene
    uint32_t T = 0, K = 0, M = 0, L = 0;
    DevBuf tri_pos, tri_nrm, tri_mat, mat_dif, mat_spec, light_pos, light_col, light_int;
    DevBuf tri_sphere, tri_mt, obj_sphere, obj_range;
    std::vector<float> host_light_pos;   // float4[L] as loaded
    // cluster level: CLUSTER consecutive triangles of one object per cluster
    uint32_t NC = 0;
    DevBuf cl_sphere, sub_sphere, sub_use, cl_base, all_list;   // float4[NC], uint32[K] first cluster, uint4[K] all objects as
                                           // list entries (first, last, first cluster, 0) for RAH
    std::vector<float> host_spheres;   // float3[2K] as loaded
    DevBuf crange, srange;             // uint2 triangle ranges of the clusters and sub-clusters
    // the scene as loaded, kept for rsh_engine_update_objects (made on its first call)
    bool orig_ready = false;
    DevBuf tri_pos0, tri_nrm0, tri_obj, sph0, mats, sph_tmp;
    uint32_t ranges_tb = 0xFFFFFFFFu;  // triangle batch the exclusion ranges were built for

    // frame buffers: closest hits and the bounce G-buffer alternate between
    // consecutive reflection passes (a pass's shading reads them while the
    // next pass traces)
    uint32_t W = 0, H = 0, Lmax = 1;
    size_t cap_slots = 0, cap_pixels = 0;   // slots (pixels x lights) and pixels the buffers hold
    size_t cap_shadow_slots = 0;            // shadow-pass flags `shadow` holds
    // shadow[k % 2]: the shadow flags of G-buffer k (a frame's shadow pass k
    // may run while G-buffer k - 1 still waits for its shading)
    DevBuf rays_a, rays_b, closest[2], shadow[2], pixels;
    DevBuf views_d;                    // view table of a batch (more than one view)
    uint64_t scene_epoch = 0;          // counts scene loads
    Lane lane[2];
    std::vector<hipEvent_t> sync_ev;   // cross-lane dependencies of a frame

    bool recording = false, replaying = false;
    // Graphs that replay for any camera of the same view sizes: the capture
    // keeps the shadow and reflection passes' sizes on the device (dev_sized),
    // sizes their launches by capacities from the recorded frame, and records
    // each pass's (N, C) and an overflow word in frame_rec, read after a
    // replay (hrec); an overflow re-renders the frame eagerly.
    bool dev_sized = false;            // set while such a frame is captured
    bool gdev = false;                 // the current graph is device-sized
    DevBuf frame_rec;                  // uint32[2 * RSH_MAX_PASSES + 1]: (N, C) per pass, overflow
    uint32_t* hrec = nullptr;          // page-locked copy of frame_rec
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> gsig;         // views + config + scene epoch the graph was captured for
    uint64_t gepoch = 0;               // allocation epoch at capture
    rsh_frame_stats gstats{};          // statistics of the captured frame
    DevBuf g0[4], gb[2][4];

    // kept passes (flag bit 0), by pass index
    struct Kept {
        DevBuf buf; uint32_t kind = 0; uint32_t slots = 0; bool set = false;
        // flag bit 6: the pass's slot keys and flags, sorted order (the slot of
        // each sorted ray), chunk bases and keys, hierarchy nodes
        DevBuf keys, flags, rec, bases, ckeys, nodes;
        uint32_t o_slots = 0, o_rays = 0, o_chunks = 0, o_nodes = 0;
        bool have_keys = false, have_chunks = false, order_set = false;
        void release() {
            for (DevBuf* b : {&buf, &keys, &flags, &rec, &bases, &ckeys, &nodes}) b->release();
            set = order_set = false;
        }
    };
    std::vector<Kept> kept;
};

namespace {

bool fail(rsh_engine* e, const std::string& msg) {
    e->err = msg;
    return false;
}

#define HIP_OK(call)                                                                       \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            e->err = std::string(#call) + ": " + hipGetErrorString(_e);                    \
            return RSH_HIP_ERROR;                                                          \
        }                                                                                  \
    } while (0)

// The same inside a lane's pipeline: the error goes to the lane (a frame's two
// lanes may fail on their own threads) and the frame copies it out.
#define LANE_OK(call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            L.err = std::string(#call) + ": " + hipGetErrorString(_e);                     \
            return RSH_HIP_ERROR;                                                          \
        }                                                                                  \
    } while (0)

// A bounded host wait (readback_wait, wait_word): past the deadline the pass
// fails with RSH_INTERNAL and says so.
#define WAIT_OK(call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e == hipErrorLaunchTimeOut) {                                                 \
            L.err = std::string(#call) + ": no result within " + std::to_string(L.wait_ms) + \
                    " ms (device hung or stream blocked)";                                 \
            return RSH_INTERNAL;                                                           \
        }                                                                                  \
        if (_e != hipSuccess) {                                                            \
            L.err = std::string(#call) + ": " + hipGetErrorString(_e);                     \
            return RSH_HIP_ERROR;                                                          \
        }                                                                                  \
    } while (0)

inline float4* f4p(const DevBuf& b) { return b.as<float4>(); }
inline uint32_t* u32p(const DevBuf& b) { return b.as<uint32_t>(); }
inline uint8_t* u8p(const DevBuf& b) { return b.as<uint8_t>(); }
inline unsigned long long* u64p(const DevBuf& b) { return b.as<unsigned long long>(); }

GBufView view(rsh_engine* e, DevBuf* g) { return GBufView{f4p(g[0]), f4p(g[1]), f4p(g[2]), f4p(g[3])}; }

SceneView scene_view(rsh_engine* e) {
    SceneView s;
    s.tri_pos = f4p(e->tri_pos); s.tri_nrm = f4p(e->tri_nrm); s.tri_mat = e->tri_mat.as<int32_t>();
    s.mat_dif = f4p(e->mat_dif); s.mat_spec = f4p(e->mat_spec);
    s.lights = Lights{f4p(e->light_pos), f4p(e->light_col), e->light_int.as<float2>(), e->L};
    return s;
}

struct StageTimer {
    Lane& L;
    int stage;
    bool on;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(Lane& ln, int st, bool enabled) : L(ln), stage(st), on(enabled) {
        if (!on) return;
        if (L.ev_used + 2 > L.ev_pool.size()) {
            for (int i = 0; i < 64; i++) {
                hipEvent_t ev;
                if (hipEventCreate(&ev) == hipSuccess) L.ev_pool.push_back(ev);
            }
        }
        if (L.ev_used + 2 > L.ev_pool.size()) { on = false; return; }
        a = L.ev_pool[L.ev_used++];
        b = L.ev_pool[L.ev_used++];
        (void)hipEventRecord(a, L.stream);
    }
    ~StageTimer() {
        if (!on) return;
        (void)hipEventRecord(b, L.stream);
        L.timed.push_back({stage, {a, b}});
    }
};

enum { ST_CREATE = 0, ST_TRIM, ST_COMPRESS, ST_SORT, ST_DECOMPRESS, ST_HIER, ST_TRAVERSE, ST_SHADE };
constexpr int ST_PASS_TRAVERSE = 100;   // + pass index: per-pass traversal time
constexpr uint64_t TRAVERSE_TARGET_WAVES = 16384;
// 64-ray kernel work items: a pass starts at ITEM_CHUNK candidate clusters
// (of CLUSTER triangles) per item and shrinks the chunk (not below
// ITEM_CHUNK_MIN) until it has at least ITEM_TARGET items, about 13 per
// resident wave slot (5120 on 256 CUs): item lengths vary widely, and short
// passes otherwise end in a long tail, while every item repeats the node's
// set-up.  With the shadow passes overlapping the reflection chain (which
// fills most tails), measured on the 1080p atrium frame: 32768 -> 7.57 ms,
// 49152 -> 7.53, 65536 -> 7.52, 98304 -> 7.64, 131072 -> 7.81 ms per frame.
constexpr uint32_t ITEM_CHUNK = 64;
// Shadow passes also take the longer items: alone the first shadow pass of
// the 1080p atrium runs 0.85 ms with 64-cluster items against 1.05 ms with
// 128, but it overlaps the reflection chain, and with longer items the frame
// is no slower (6.02 against 6.06 ms) and the 64-view batch faster (66.3
// against 67.8 ms).
#ifndef RSH_SHADOW_LONG_ITEMS
#define RSH_SHADOW_LONG_ITEMS 1
#endif
// Passes with at least RSH_LONGER_ITEMS * ITEM_TARGET items of 2 ITEM_CHUNK
// clusters take items of 4 ITEM_CHUNK (0: never).  On the 1080p atrium (its
// camera pass alone qualifies) and its 64-view batch (most passes do), median
// frame / graph-replayed batch: never 5.77 / 62.0 ms, 4 5.65 / 59.9, 2 5.76 /
// 60.0 (the first reflection pass also takes them and runs 0.97 -> 1.09 ms).
#ifndef RSH_LONGER_ITEMS
#define RSH_LONGER_ITEMS 4
#endif
constexpr size_t HSTAGE_WORDS = 2 + 2 * ITEM_TABLE_WORDS;   // N / C or the item total, then the item table
constexpr uint32_t ITEM_CHUNK_MIN = 4;
#ifndef RSH_ITEM_TARGET
#define RSH_ITEM_TARGET 65536
#endif
constexpr uint32_t ITEM_TARGET = RSH_ITEM_TARGET;

// Host waits are bounded: past the lane's deadline (wait_ms) a wait gives up
// with WAIT_TIMEOUT and the pass fails with RSH_INTERNAL, instead of spinning
// for ever on a stream that never drains (a hung device, or a caller's stream
// blocked on outside work).
constexpr hipError_t WAIT_TIMEOUT = hipErrorLaunchTimeOut;

inline bool past_deadline(std::chrono::steady_clock::time_point t0, uint32_t ms) {
    return std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(ms);
}

// Wait for a lane's read-back copies: an event after them, polled.  A
// read-back is a few microseconds of GPU work, and a blocking wait would add
// the host thread's wake-up latency to every pass.
hipError_t readback_wait(Lane& L) {
    hipError_t r = hipEventRecord(L.rb_ev, L.stream);
    if (r != hipSuccess) return r;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 1;; spin++) {
        r = hipEventQuery(L.rb_ev);
        if (r != hipErrorNotReady) return r;
        if ((spin & 1023u) == 0 && past_deadline(t0, L.wait_ms)) return WAIT_TIMEOUT;
        __builtin_ia32_pause();
    }
}

// Wait for a kernel's sequence word `seq` in coherent host memory (the pack
// scan's sizes, the item table).  Every 1024 polls the stream is queried: an
// error ends the wait, and so does an idle stream without the word (then the
// values never came: an internal error), and so does the deadline.
template <class T>
hipError_t wait_word(Lane& L, const T* word, uint32_t seq) {
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 1;; spin++) {
        if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == (T)seq) return hipSuccess;
        if ((spin & 1023u) == 0) {
            const hipError_t r = hipStreamQuery(L.stream);
            if (r == hipSuccess)
                return __atomic_load_n(word, __ATOMIC_ACQUIRE) == (T)seq ? hipSuccess : hipErrorUnknown;
            if (r != hipErrorNotReady) return r;
            if (past_deadline(t0, L.wait_ms)) return WAIT_TIMEOUT;
        }
        __builtin_ia32_pause();
    }
}
hipError_t wait_sizes(Lane& L, uint32_t seq) { return wait_word(L, L.hsize, seq); }

// Lane 1's stream and page-locked words, created on first use; lane 0 runs
// on the engine's stream.
rsh_status ensure_lanes(rsh_engine* e) {
    e->lane[0].stream = e->stream;
    Lane& L = e->lane[1];
    if (!L.stream) {
        HIP_OK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        L.own_stream = true;
    }
    for (Lane& l : e->lane) {
        if (!l.hstage) HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&l.hstage), HSTAGE_WORDS * 4, hipHostMallocDefault));
        if (!l.hsize) {
            HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&l.hsize), 64, hipHostMallocCoherent | hipHostMallocMapped));
            std::memset(l.hsize, 0, 64);
            l.hseq = 0;
            HIP_OK(hipHostGetDevicePointer(reinterpret_cast<void**>(&l.hsize_dev), l.hsize, 0));
        }
        if (!l.htable) {
            const size_t bytes = (1 + ITEM_TABLE_WORDS) * sizeof(unsigned long long);
            HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&l.htable), bytes, hipHostMallocCoherent | hipHostMallocMapped));
            std::memset(l.htable, 0, bytes);
            l.tseq = 0;
            HIP_OK(hipHostGetDevicePointer(reinterpret_cast<void**>(&l.htable_dev), l.htable, 0));
            HIP_OK(l.w_done.reserve(64));
            launch_fill_u32(u32p(l.w_done), 0u, 16, l.stream);     // stream-ordered before its first use
        }
        if (!l.rb_ev) HIP_OK(hipEventCreateWithFlags(&l.rb_ev, hipEventDisableTiming));
    }
    while (e->sync_ev.size() < 3 * (RSH_MAX_PASSES + 2)) {
        hipEvent_t ev;
        HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e->sync_ev.push_back(ev);
    }
    return RSH_OK;
}

rsh_status ensure_lane(rsh_engine* e, Lane& L, size_t S) {
    if (S <= L.cap_slots) return RSH_OK;
    const size_t u = S * 4;
    for (DevBuf* b : {&L.flags, &L.keys, &L.values, &L.tvalues, &L.ckeys, &L.cvalues, &L.skeys, &L.svalues,
                      &L.rvalues})
        HIP_OK(b->reserve(u));
    HIP_OK(L.bases.reserve(u + 4));   // bases[C] = N closes the last chunk
    HIP_OK(L.soff.reserve(u + 4));
    HIP_OK(L.rec0.reserve(S * 16)); HIP_OK(L.rec1.reserve(S * 16));
    HIP_OK(L.counters.reserve(COUNTER_SLOTS * COUNTERS_PER_SLOT * 8)); HIP_OK(L.dcounts.reserve(64));
    HIP_OK(L.sc.reserve(S));
    L.cap_slots = S;
    return RSH_OK;
}

// shadow_mult: rays per slot of the shadow passes (16 with soft shadows);
// they run on lane 1 and write the shadow flags (their rays are recomputed by
// the gather, so they have no slot-order ray buffer).  rays_b holds the camera
// rays of a frame created in slot order (one per pixel).
rsh_status ensure_frame(rsh_engine* e, uint32_t W, uint32_t H, uint32_t Lmax, uint32_t shadow_mult = 1) {
    size_t S = (size_t)W * H * Lmax, P = (size_t)W * H, Ss = S * shadow_mult;
    e->W = W; e->H = H; e->Lmax = Lmax;
    rsh_status rs = ensure_lanes(e);
    if (rs != RSH_OK) return rs;
    if ((rs = ensure_lane(e, e->lane[0], S)) != RSH_OK) return rs;
    if ((rs = ensure_lane(e, e->lane[1], Ss)) != RSH_OK) return rs;
    // per-slot and per-pixel buffers grow separately: a scene with more lights
    // can leave enough slots for a later frame with more pixels
    if (S > e->cap_slots) {
        HIP_OK(e->rays_a.reserve(S * 24));
        HIP_OK(e->closest[0].reserve(S * 8)); HIP_OK(e->closest[1].reserve(S * 8));
        e->cap_slots = S;
    }
    if (Ss > e->cap_shadow_slots) {
        HIP_OK(e->shadow[0].reserve(Ss * 4)); HIP_OK(e->shadow[1].reserve(Ss * 4));
        e->cap_shadow_slots = Ss;
    }
    if (P > e->cap_pixels) {
        HIP_OK(e->rays_b.reserve(P * 24));
        HIP_OK(e->pixels.reserve(P * 4));
        for (int i = 0; i < 4; i++) {
            HIP_OK(e->g0[i].reserve(P * 16)); HIP_OK(e->gb[0][i].reserve(P * 16));
            HIP_OK(e->gb[1][i].reserve(P * 16));
        }
        e->cap_pixels = P;
    }
    return RSH_OK;
}

// Hierarchy levels of N rays on both lanes.
rsh_status ensure_nodes(rsh_engine* e, uint32_t N, const rsh_frame_config& cfg) {
    uint32_t cnt[4], off[4];
    const size_t total = level_sizes(N, cfg.leaf, cfg.branch, cfg.depth, cnt, off);
    for (Lane& L : e->lane) {
        HIP_OK(L.nodes.reserve(total * 32 + 64));
        HIP_OK(L.aux.reserve(total * 8 + 64));
    }
    return RSH_OK;
}

rsh_status build_ranges(rsh_engine* e, uint32_t Tb) {
    if (Tb == e->ranges_tb) return RSH_OK;
    std::vector<uint32_t> r(4 * (size_t)e->K);
    for (uint32_t k = 0; k < e->K; k++) {
        uint32_t y = (uint32_t)e->host_spheres[6 * k + 4], z = (uint32_t)e->host_spheres[6 * k + 5];
        uint32_t lo = 0, hi = 0;
        // RayTracer.cu:1192-1244 has no branch for an object that starts before a
        // batch and ends exactly at the batch's upper bound: the batch's pairs of
        // that object are never candidates.
        if (Tb && z != 0 && z % Tb == 0 && y < z - Tb) { lo = z - Tb; hi = z; }
        r[4 * k] = y; r[4 * k + 1] = z; r[4 * k + 2] = lo; r[4 * k + 3] = hi;
    }
    HIP_OK(hipMemcpyAsync(e->obj_range.p, r.data(), r.size() * 4, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->ranges_tb = Tb;
    return RSH_OK;
}

uint32_t key_bits(const rsh_frame_config& c) {
    if (c.key_mode == RSH_KEY_REF) return 32;
    uint32_t b = 2 + 3 * c.cell_bits + 2 * c.dir_bits;
    if (c.segment_shift) b = 32;
    return b > 32 ? 32 : b;
}

KeyCfg key_cfg(const rsh_frame_config& c) {
    KeyCfg k{};
    k.mode = c.key_mode;
    k.h.cell_bits = c.cell_bits; k.h.dir_bits = c.dir_bits; k.h.layout = c.key_layout;
    k.h.segment_shift = c.segment_shift;
    k.h.min_x = c.scene_min[0]; k.h.min_y = c.scene_min[1]; k.h.min_z = c.scene_min[2];
    k.h.scale_x = c.cell_scale[0]; k.h.scale_y = c.cell_scale[1]; k.h.scale_z = c.cell_scale[2];
    return k;
}

// Cluster and sub-cluster spheres and the sub-cluster flags from the triangle
// spheres (scene load and object updates).
void build_clusters(rsh_engine* e, hipStream_t st) {
    launch_cluster_build(f4p(e->tri_sphere), e->crange.as<uint2>(), e->NC, f4p(e->cl_sphere), st);
    launch_cluster_build(f4p(e->tri_sphere), e->srange.as<uint2>(), e->NC * SUB_PER_CLUSTER, f4p(e->sub_sphere), st);
    launch_sub_use(f4p(e->cl_sphere), f4p(e->sub_sphere), e->NC, e->sub_use.as<uint8_t>(), st);
}

struct PassOut {
    uint32_t N = 0, C = 0, splits = 0, items = 0, chunk = 0;
    uint64_t top = 0, leaf = 0, tests = 0, surv = 0, tri_tests = 0, cl_tests = 0, b_steps = 0, c_steps = 0;
    uint64_t mt_tests = 0;
    uint32_t mode = 0;     // traversal mode (TravParams::prefilter)
};

// Work list of the 64-ray kernel (see uses_items): cull, size, scan, fill.
// Work-list buffers of a pass with ntop top nodes, list_total list entries
// and items work items.
hipError_t reserve_items(Lane& L, uint32_t ntop, uint32_t K, uint64_t list_total, uint64_t items) {
    const size_t nb = ((size_t)ntop + 1) * 4;
    hipError_t r = hipSuccess;
    for (DevBuf* b : {&L.w_npass, &L.w_cand, &L.w_nitems, &L.w_off, &L.w_first})
        if (r == hipSuccess) r = b->reserve(nb);
    if (r == hipSuccess) r = L.w_pmask.reserve(((size_t)ntop + 1) * ((K + 63) / 64) * 8 + 8);
    if (r == hipSuccess) r = L.w_list.reserve((size_t)list_total * 16 + 16);
    if (r == hipSuccess) r = L.w_cum.reserve((size_t)list_total * 4 + 4);
    if (r == hipSuccess) r = L.w_items.reserve((size_t)items * 16 + 16);
    return r;
}

// Capacities of a device-sized pass's work list from the recorded pass.
uint64_t list_capacity(const PassPlan& pp) { return (uint64_t)pp.list_total * 3 / 2 + 4096; }
uint64_t items_capacity(const PassPlan& pp) {
    const uint64_t c = (uint64_t)pp.items * 3 / 2 + 4096;
    return c < MAX_TRAVERSE_ITEMS ? c : MAX_TRAVERSE_ITEMS;
}

rsh_status build_items(rsh_engine* e, Lane& L, const rsh_frame_config& cfg, TravParams& p, PassOut& out,
                       bool camera) {
    hipStream_t s = L.stream;
    const uint32_t ntop = p.level_count[1];
    const size_t nb = ((size_t)ntop + 1) * 4;
    for (DevBuf* b : {&L.w_npass, &L.w_cand, &L.w_nitems, &L.w_off, &L.w_first}) LANE_OK(b->reserve(nb));
    uint32_t* npass = u32p(L.w_npass);
    uint32_t* cand = u32p(L.w_cand);
    uint32_t* nitems = u32p(L.w_nitems);
    LANE_OK(L.w_pmask.reserve(((size_t)ntop + 1) * ((e->K + 63) / 64) * 8 + 8));
    unsigned long long* pmask = L.w_pmask.as<unsigned long long>();
    launch_cull_count(p, npass, cand, pmask, s);
    // the list offsets: exclusive scan of npass, done by the item scan's launch
    const uint32_t* np_scan = p.algorithm != 1 ? npass : nullptr;
    uint32_t chunk = cfg.traversal_chunk ? cfg.traversal_chunk : ITEM_CHUNK;
    uint32_t tot[2] = {0, 0};   // list entries, items
    if (e->replaying && p.ps) {
        // device-sized: the recorded pass's chunk and mode; the list and the
        // item total stay on the device, bounded by capacities
        const PassPlan& pp = L.plan[L.plan_pos - 1];
        if (pp.chunk) { chunk = pp.chunk; p.prefilter = pp.mode; }
        tot[0] = (uint32_t)list_capacity(pp); tot[1] = (uint32_t)items_capacity(pp);
        p.list_cap = tot[0]; p.items_cap = tot[1];
        LANE_OK(run_item_scan(cand, ntop, chunk, e->NC, nitems, u32p(L.w_first), L.sc, s, nullptr, 0, np_scan,
                              u32p(L.w_off), p.ps));
    } else if (e->replaying) {
        const PassPlan& pp = L.plan[L.plan_pos - 1];
        chunk = pp.chunk; tot[0] = pp.list_total; tot[1] = pp.items; p.prefilter = pp.mode;
        LANE_OK(run_item_scan(cand, ntop, chunk, e->NC, nitems, u32p(L.w_first), L.sc, s, nullptr, 0, np_scan,
                              u32p(L.w_off)));
    }
    if (!e->replaying) {
        // One read-back sizes the pass: the list total and, from the table of
        // item totals per chunk size, the chunk (shrunk from ITEM_CHUNK until the
        // pass has ITEM_TARGET items) and the item total.
        const bool tabled = chunk <= ITEM_TABLE || (chunk <= 2 * ITEM_TABLE && chunk % 2 == 0) ||
                            (chunk <= 4 * ITEM_TABLE && chunk % 4 == 0);
        unsigned long long ht[ITEM_TABLE_WORDS] = {};
        // c tabled: c <= ITEM_TABLE, or even up to 2 ITEM_TABLE, or a multiple of 4 up to 4 ITEM_TABLE
        auto items_at = [&](uint32_t c) {
            return c <= ITEM_TABLE ? ht[c - 1] : c <= 2 * ITEM_TABLE ? ht[ITEM_TABLE + c / 2 - 1]
                                                                      : ht[2 * ITEM_TABLE + c / 4 - 1];
        };
        auto tabled_below = [&](uint32_t c) {    // the largest tabled chunk size <= c
            return c <= ITEM_TABLE ? c : c <= 2 * ITEM_TABLE ? c & ~1u : c & ~3u;
        };
        uint32_t* first = u32p(L.w_first);
        LANE_OK(L.w_table.reserve(ITEM_TABLE_WORDS * sizeof(unsigned long long)));
        unsigned long long* table = L.w_table.as<unsigned long long>();
        if (!L.table_clean) launch_fill_u64(table, 0ull, ITEM_TABLE_WORDS, s);   // new, or a pass that failed
        L.table_clean = false;
        // the tabled case reads the table from the last k_item_table block's
        // host copy; otherwise the item scan's total is needed too
        const bool publish = tabled && ntop > 0;
        launch_item_table(cand, p.aux + p.level_offset[1], ntop, e->NC, np_scan, table, s,
                          publish ? u32p(L.w_done) : nullptr, publish ? L.htable_dev : nullptr,
                          publish ? ++L.tseq : 0u);
        if (!publish) LANE_OK(hipMemcpyAsync(L.hstage + 2, table, sizeof(ht), hipMemcpyDeviceToHost, s));
        if (!tabled) {
            LANE_OK(run_item_scan(cand, ntop, chunk, e->NC, nitems, first, L.sc, s, nullptr, 0, np_scan, u32p(L.w_off)));
            LANE_OK(hipMemcpyAsync(L.hstage + 1, first + ntop, 4, hipMemcpyDeviceToHost, s));
        }
        if (publish) {
            WAIT_OK(wait_word(L, reinterpret_cast<uint64_t*>(L.htable), L.tseq));
            for (uint32_t i = 0; i < ITEM_TABLE_WORDS; i++) ht[i] = __atomic_load_n(L.htable + 1 + i, __ATOMIC_RELAXED);
        } else {
            WAIT_OK(readback_wait(L));
            std::memcpy(ht, L.hstage + 2, sizeof(ht));
        }
        if (p.algorithm != 1) tot[0] = (uint32_t)ht[TABLE_LIST];
        if (!tabled) {
            tot[1] = L.hstage[1];
            launch_fill_u64(table, 0ull, ITEM_TABLE_WORDS, s);    // for the next pass
            L.table_clean = true;
        }
        // Traversal mode: the cluster ray masks pay when many top nodes are wide
        // (spread-out secondary rays) and cost where rays are coherent.  A frame's
        // camera rays never take them (measured faster without on the 1080p
        // atrium frame and on both 64-view batches, whose small views leave many
        // primary nodes above WIDE_TAN); otherwise more than a sixteenth of the
        // nodes with tan(spread) outside [0, WIDE_TAN] picks the best mode for
        // every pass of the 1080p atrium frame (first shadow pass without, every
        // later pass with).
        if (cfg.traversal_prefilter == 0)
            p.prefilter = camera ? RSH_CAMERA_MODE
                          : (RSH_SHADOW_MODE1 && p.kind == 1) ? 1u
                          : RSH_WIDE_DIV * ht[TABLE_WIDE] > ntop ? 2u : 1u;
        if (tabled) {
            if (!cfg.traversal_chunk) {
                // passes with plenty of items take longer items (less per-item
                // set-up: on the 1080p atrium, primary 1.98 -> 1.83 ms, first
                // reflection pass 1.13 -> 1.05 ms)
                if (items_at(ITEM_CHUNK) >= 2 * ITEM_TARGET && (RSH_SHADOW_LONG_ITEMS || p.kind != 1))
                    chunk = 2 * ITEM_CHUNK;
                if (RSH_LONGER_ITEMS && chunk == 2 * ITEM_CHUNK &&
                    items_at(2 * ITEM_CHUNK) >= (uint64_t)RSH_LONGER_ITEMS * ITEM_TARGET)
                    chunk = 4 * ITEM_CHUNK;
                while (items_at(chunk) < ITEM_TARGET && chunk > ITEM_CHUNK_MIN && items_at(chunk) != 0) {
                    const uint64_t want = (uint64_t)chunk * items_at(chunk) / ITEM_TARGET;
                    chunk = want < ITEM_CHUNK_MIN ? ITEM_CHUNK_MIN : tabled_below((uint32_t)want);
                }
            }
            if (items_at(chunk) > MAX_TRAVERSE_ITEMS) {
                L.err = "traversal work list of " + std::to_string(items_at(chunk)) + " items exceeds " +
                         std::to_string(MAX_TRAVERSE_ITEMS) + " (raise traversal_chunk)";
                return RSH_CAPACITY;
            }
            tot[1] = (uint32_t)items_at(chunk);
            // the item scan also clears the table for the next pass
            LANE_OK(run_item_scan(cand, ntop, chunk, e->NC, nitems, first, L.sc, s, table, ITEM_TABLE_WORDS, np_scan,
                                  u32p(L.w_off)));
            L.table_clean = true;
        }
    }
    LANE_OK(L.w_list.reserve((size_t)tot[0] * 16 + 16));
    LANE_OK(L.w_cum.reserve((size_t)tot[0] * 4 + 4));
    LANE_OK(L.w_items.reserve((size_t)tot[1] * 16 + 16));
    p.item_desc = L.w_items.as<uint4>(); p.item_first = u32p(L.w_first); p.cand = cand;
    p.obj_off = u32p(L.w_off); p.obj_ent = L.w_list.as<uint4>(); p.obj_cum = u32p(L.w_cum);
    if (p.algorithm == 1) { p.obj_ent = e->all_list.as<uint4>(); p.obj_cum = u32p(e->cl_base); }
    p.chunk = chunk; p.items = tot[1];
    if (p.algorithm != 1)
        launch_cull_items(p, L.w_list.as<uint4>(), u32p(L.w_cum), pmask, L.w_items.as<uint4>(), s);   // at p.obj_off
    else
        launch_item_desc(p, L.w_items.as<uint4>(), s);
    LANE_OK(hipGetLastError());
    out.items = tot[1]; out.chunk = chunk; out.mode = p.prefilter;
    if (e->recording) {
        PassPlan& pp = L.plan.back();
        pp.chunk = chunk; pp.list_total = tot[0]; pp.items = tot[1]; pp.mode = p.prefilter;
    }
    return RSH_OK;
}

// One ray batch: trim -> (compress -> sort -> decompress) -> gather ->
// hierarchy -> fused traversal + intersection.  kind 0: closest hit into
// closest; kind 1: shadow flags (u32 per slot) into `closest`.
rsh_status trace_pass(rsh_engine* e, Lane& L, const rsh_frame_config& cfg, const float* rays, uint32_t S,
                      uint32_t kind, const float4* shadow_pos, uint32_t flags_mask, PassOut& out, uint32_t pass_idx,
                      DevBuf& closest, const float* tmax = nullptr, bool camera = false, uint32_t sort_bits = 0,
                      bool presorted = false, const float* front_point = nullptr) {
    out = PassOut{};
    hipStream_t s = L.stream;
    const bool timing = flags_mask & 4u, counting = flags_mask & 2u;
    const bool trav_timing = timing || (flags_mask & 16u);     // bit 4: traversal launches only
    uint32_t* dc = u32p(L.dcounts);
    // A frame's camera rays are all live and their keys (screen tiles) are
    // unique: trimming and compression are the identity, so the slots are
    // sorted directly and no sizes are read back.
    const bool all_unique = camera && kind == 0;
    if (all_unique) {
        // hit records cleared: no trim or compression to time
        launch_fill_u64(u64p(closest), cfg.hit_mode == 0 ? MISS64 : (unsigned long long)MISS32, S, s);
    } else {
        // trimming and compression in one walk over the slots (the compress
        // stage time is part of the trim stage's).  Eager passes take N and C
        // from the pack scan's host words, published before the chunk writes,
        // so the next launches are issued while those run.
        StageTimer t(L, ST_TRIM, timing);
        const unsigned long long miss = cfg.hit_mode == 0 ? MISS64 : (unsigned long long)MISS32;
        const bool poll = !e->replaying && S > 0;
        // slot flags are bytes; a live slot's value is its index (creation writes no values)
        LANE_OK(run_pack(u8p(L.flags), u32p(L.keys), nullptr, S, u32p(L.tvalues), u32p(L.bases),
                        u32p(L.ckeys), u32p(L.cvalues), dc, kind == 0 ? u64p(closest) : nullptr, miss,
                        kind == 1 ? u32p(closest) : nullptr, L.sc, s, poll ? L.hsize_dev : nullptr,
                        poll ? ++L.hseq : 0u));
    }
    // Device-sized pass (a capture for any camera): N and C stay in device
    // memory; the launches below cover the pass's S slots.  The camera pass's
    // sizes are fixed (every slot live, one chunk per ray), but its work list
    // depends on the camera, so it takes the same path.
    const bool dev = e->dev_sized && cfg.algorithm == RSH_CRSH && S > 0;
    PassSizes* psz = dev ? L.psz.as<PassSizes>() : nullptr;
    if (dev && all_unique)
        launch_fixed_sizes(S, S, cfg.leaf, cfg.branch, cfg.depth, psz, u32p(e->frame_rec) + 2 * (size_t)pass_idx, s);
    else if (dev)
        launch_pass_sizes(dc, cfg.leaf, cfg.branch, cfg.depth, psz, u32p(e->frame_rec) + 2 * (size_t)pass_idx, s);
    uint32_t hc[2] = {0, 0};
    if (e->replaying) {
        const PassPlan& pp = L.plan[L.plan_pos++];
        hc[0] = pp.N; hc[1] = pp.C;
        if (dev) hc[0] = hc[1] = S;       // capacities
    } else if (all_unique) {
        hc[0] = hc[1] = S;
        if (e->recording) { L.plan.emplace_back(); L.plan.back().N = S; L.plan.back().C = S; }
    } else if (S == 0) {
        if (e->recording) { L.plan.emplace_back(); L.plan.back().N = 0; L.plan.back().C = 0; }
    } else {
        WAIT_OK(wait_sizes(L, L.hseq));
        hc[0] = __atomic_load_n(L.hsize + 1, __ATOMIC_RELAXED);
        hc[1] = __atomic_load_n(L.hsize + 2, __ATOMIC_RELAXED);
        if (e->recording) { L.plan.emplace_back(); L.plan.back().N = hc[0]; L.plan.back().C = hc[1]; }
    }
    const uint32_t N = hc[0];
    out.N = N;
    if (N == 0) return RSH_OK;
    const uint32_t* sorted_values = all_unique ? u32p(L.values) : u32p(L.tvalues);
    if (presorted) {
        // camera rays created in key order, straight into the hierarchy
        // records (k_create_primary_sorted): nothing to sort or gather
        out.C = N;
    } else if (cfg.algorithm == RSH_CRSH && all_unique) {
        out.C = N;
        StageTimer t(L, ST_SORT, timing);
        LANE_OK(run_radix_sort(u32p(L.keys), u32p(L.values), N, sort_bits ? sort_bits : key_bits(cfg), u32p(L.skeys),
                               u32p(L.svalues), L.sc, s));
        sorted_values = u32p(L.svalues);
    } else if (cfg.algorithm == RSH_CRSH) {
        const uint32_t C = hc[1];
        out.C = C;
        const uint32_t bits = sort_bits ? sort_bits : key_bits(cfg);
        {
            StageTimer t(L, ST_SORT, timing);
            if (dev)
                LANE_OK(run_radix_sort_dev(u32p(L.ckeys), u32p(L.cvalues), C, &psz->C, bits, u32p(L.skeys),
                                           u32p(L.svalues), L.sc, s));
            else
                LANE_OK(run_radix_sort(u32p(L.ckeys), u32p(L.cvalues), C, bits, u32p(L.skeys), u32p(L.svalues), L.sc,
                                       s));
        }
        {
            StageTimer t(L, ST_DECOMPRESS, timing);
            // the sorted keys are not needed after the sort: their buffer holds the chunk bases
            if (dev)
                LANE_OK(run_unpack_dev(u32p(L.bases), u32p(L.svalues), u32p(L.tvalues), C, psz, u32p(L.soff),
                                       u32p(L.skeys), u32p(L.rvalues), L.sc, s));
            else
                LANE_OK(run_unpack(u32p(L.bases), u32p(L.svalues), u32p(L.tvalues), C, N, u32p(L.soff),
                                   u32p(L.skeys), u32p(L.rvalues), L.sc, s));
        }
        sorted_values = u32p(L.rvalues);
    } else {
        out.C = N;
    }
    uint32_t counts[4] = {0, 0, 0, 0}, offs[4] = {0, 0, 0, 0};
    level_sizes(N, cfg.leaf, cfg.branch, cfg.depth, counts, offs);
    {
        StageTimer t(L, ST_HIER, timing);
        if (!presorted)
            launch_gather(rays, sorted_values, N, psz, kind == 1 ? shadow_pos : nullptr, kind == 1 ? tmax : nullptr,
                          e->W, e->Lmax, kind == 1 && shadow_pos && cfg.soft_shadows ? 4u : 0u, f4p(L.rec0),
                          f4p(L.rec1), s, f4p(e->light_pos));
        launch_hierarchy(f4p(L.rec0), f4p(L.rec1), psz, nullptr, nullptr, N, cfg.leaf, cfg.branch, cfg.depth,
                         cfg.initial_radius, cfg.initial_spread, cfg.cone_mode, f4p(L.nodes), L.aux.as<float2>(),
                         counts, offs, s);
        LANE_OK(hipGetLastError());
    }
    if ((flags_mask & 64u) && pass_idx < e->kept.size()) {
        // debug read-back of the pass's keys, order, packets and nodes
        rsh_engine::Kept& k = e->kept[pass_idx];
        uint32_t nn = 0;
        for (uint32_t d = 0; d < cfg.depth; d++) nn += counts[d];
        k.o_slots = S; k.o_rays = N; k.o_chunks = out.C; k.o_nodes = 2 * nn;
        k.have_keys = !presorted;
        k.have_chunks = !presorted && cfg.algorithm == RSH_CRSH && !all_unique;
        auto keep = [&](DevBuf& dst, const void* src, size_t bytes) -> hipError_t {
            hipError_t r = dst.reserve(bytes);
            return r == hipSuccess ? hipMemcpyAsync(dst.p, src, bytes, hipMemcpyDeviceToDevice, s) : r;
        };
        if (k.have_keys) {
            LANE_OK(keep(k.keys, L.keys.p, (size_t)S * 4));
            LANE_OK(keep(k.flags, L.flags.p, (size_t)S));        // one byte per slot
        }
        if (k.have_chunks) {
            LANE_OK(keep(k.bases, L.bases.p, (size_t)out.C * 4));
            LANE_OK(keep(k.ckeys, L.ckeys.p, (size_t)out.C * 4));
        }
        LANE_OK(keep(k.rec, L.rec0.p, (size_t)N * 16));
        LANE_OK(keep(k.nodes, L.nodes.p, (size_t)2 * nn * 16));
        k.order_set = true;
    }
    TravParams p{};
    p.nodes = f4p(L.nodes); p.aux = L.aux.as<float2>();
    for (int k = 0; k < 4; k++) { p.level_offset[k] = offs[k]; p.level_count[k] = counts[k]; }
    p.depth = cfg.depth; p.leaf = cfg.leaf; p.branch = cfg.branch;
    uint32_t rpt = cfg.leaf;
    for (uint32_t k = 1; k < cfg.depth; k++) rpt *= cfg.branch;
    p.rays_per_top = rpt;
    p.waves_per_top = (rpt + 63) / 64;
    p.rec0 = f4p(L.rec0); p.rec1 = f4p(L.rec1); p.N = N;
    p.tri_sphere = f4p(e->tri_sphere); p.tri_mt = f4p(e->tri_mt); p.T = e->T;
    p.obj_sphere = f4p(e->obj_sphere); p.obj_range = e->obj_range.as<uint4>(); p.K = e->K;
    p.cl_sphere = f4p(e->cl_sphere); p.sub_sphere = f4p(e->sub_sphere); p.sub_use = e->sub_use.as<uint8_t>(); p.obj_cl_base = u32p(e->cl_base); p.n_clusters = e->NC;
    p.algorithm = cfg.algorithm; p.kind = kind; p.hit_mode = cfg.hit_mode;
    p.closest = u64p(closest); p.shadow = u32p(closest);       // by kind
    p.counters = counting ? u64p(L.counters) : nullptr;
    p.force_generic = cfg.traversal_kernel == 1 ? 1u : 0u;
    // Traversal mode: 0 exact tests only, 1 ray / triangle-sphere pre-test, 2 that
    // and ray / cluster-sphere masks.  Automatic (0): 1 or 2, picked per pass from
    // the spread of its top nodes when the work list is sized.
    p.prefilter = cfg.traversal_prefilter == 1 ? 0u : cfg.traversal_prefilter == 3 ? 2u : 1u;
    p.splits = 1;
    p.line_only = RSH_CAMERA_LINE_ONLY && camera && kind == 0 ? 1u : 0u;
    p.ps = psz;
    p.overflow = dev ? u32p(e->frame_rec) + 2 * RSH_MAX_PASSES : nullptr;
    if (counting) launch_fill_u64(u64p(L.counters), 0ull, COUNTER_SLOTS * COUNTERS_PER_SLOT, s);
    // Front to back: a single view's camera pass (exact records, automatic
    // mode) from the eye, a single light's shadow pass from the light
    // (front_point): objects nearest it first in every node's cluster stream,
    // and the traversal in two phases (first items, then the rest, which start
    // from what the first merged and prune against it).  Counting runs keep
    // one phase and index order.
    const bool front_kind = (kind == 0 && camera && RSH_CAMERA_FRONT && cfg.hit_mode == 0 &&
                             cfg.traversal_prefilter == 0) ||
                            (kind == 1 && RSH_SHADOW_FRONT && e->L == 1);
    const bool front = front_kind && front_point && !counting && cfg.algorithm == RSH_CRSH && e->K > 1 &&
                       L.ord_host && uses_items(p);
    if (uses_items(p)) {
        // cull + work list; its time (host read-backs included) counts as traversal
        StageTimer t(L, ST_TRAVERSE, timing);
        if (front) {
            // by the distance from the point to each object's sphere as loaded
            // (moved objects keep their loaded place: the order only steers
            // the pruning, never a result)
            std::vector<std::pair<float, uint32_t>> dk(e->K);
            for (uint32_t k = 0; k < e->K; k++) {
                const float* h = &e->host_spheres[6 * (size_t)k];
                const float dx = h[0] - front_point[0], dy = h[1] - front_point[1], dz = h[2] - front_point[2];
                const float d = std::sqrt(dx * dx + dy * dy + dz * dz) - h[3];
                dk[k] = {d > 0.0f ? d : 0.0f, k};
            }
            std::sort(dk.begin(), dk.end());
            for (uint32_t k = 0; k < e->K; k++) L.ord_host[k] = dk[k].second;
            LANE_OK(hipMemcpyAsync(L.obj_order.p, L.ord_host, (size_t)e->K * 4, hipMemcpyHostToDevice, s));
            p.obj_order = u32p(L.obj_order);
        }
        rsh_status r = build_items(e, L, cfg, p, out, camera);
        if (r != RSH_OK) return r;
    } else {
        // Enough waves to fill the GPU: passes with few top nodes split each top
        // node's triangle range across waves (at least 64 triangles per wave).
        const uint64_t base = (uint64_t)counts[cfg.depth - 1] * p.waves_per_top;
        uint64_t sp = (TRAVERSE_TARGET_WAVES + base - 1) / (base ? base : 1);
        const uint64_t max_sp = e->T / 64 ? e->T / 64 : 1;
        if (sp > max_sp) sp = max_sp;
        if (sp > 1024) sp = 1024;
        if (cfg.traversal_splits) sp = cfg.traversal_splits;
        p.splits = (uint32_t)(sp ? sp : 1);
        out.splits = p.splits;
    }
    {
        StageTimer t(L, ST_PASS_TRAVERSE + (int)pass_idx, trav_timing);
        if (front && RSH_CAMERA_PHASES) {
            p.phase = 1;
            LANE_OK(launch_traverse(p, s));
            p.phase = 2;
            if (RSH_CAMERA_LINE_ONLY == 2) p.line_only = 0;
        } else if (front) {
            p.phase = 3;
        }
        LANE_OK(launch_traverse(p, s));
    }
    if (counting) {
        unsigned long long c[COUNTER_SLOTS * COUNTERS_PER_SLOT];
        LANE_OK(hipMemcpyAsync(c, L.counters.p, sizeof(c), hipMemcpyDeviceToHost, s));
        LANE_OK(hipStreamSynchronize(s));
        unsigned long long sum[COUNTERS_PER_SLOT] = {};
        for (uint32_t k = 0; k < COUNTER_SLOTS; k++)
            for (uint32_t i = 0; i < COUNTERS_PER_SLOT; i++) sum[i] += c[(size_t)k * COUNTERS_PER_SLOT + i];
        out.top = sum[CNT_TOP_PAIRS]; out.leaf = sum[CNT_LEAF_PAIRS]; out.tests = sum[CNT_RAY_TESTS];
        out.surv = sum[CNT_TOP_SURV]; out.tri_tests = sum[CNT_TRI_TESTS]; out.cl_tests = sum[CNT_CLUSTER_TESTS];
        out.b_steps = sum[CNT_B_STEPS]; out.c_steps = sum[CNT_C_STEPS]; out.mt_tests = sum[CNT_MT_TESTS];
    }
    return RSH_OK;
}

// Statistics of pass `idx` (its own slots of st, so the two lanes never write
// the same field) and, with flag bit 0, a copy of its results on its lane.
// kind: 0 primary, 1 shadow, 2 reflection.
rsh_status record(rsh_engine* e, Lane& L, rsh_frame_stats* st, uint32_t idx, uint32_t kind, uint32_t S,
                  const PassOut& po, uint32_t fl, const DevBuf& closest) {
    if (st && idx < RSH_MAX_PASSES) {
        const uint32_t i = idx;
        st->pass_splits[i] = po.splits;
        st->pass_items[i] = po.items; st->pass_chunk[i] = po.chunk;
        st->pass_kind[i] = kind; st->pass_slots[i] = S; st->pass_rays[i] = po.N; st->pass_chunks[i] = po.C;
        st->pass_pairs_top[i] = po.top; st->pass_pairs_leaf[i] = po.leaf; st->pass_ray_tests[i] = po.tests;
        st->pass_top_survivors[i] = po.surv;
        st->pass_tri_tests[i] = po.tri_tests; st->pass_cluster_tests[i] = po.cl_tests;
        st->pass_leaf_steps[i] = po.b_steps; st->pass_ray_steps[i] = po.c_steps; st->pass_mt_tests[i] = po.mt_tests;
        st->pass_mode[i] = po.mode;
    }
    if ((fl & 1u) && idx < e->kept.size()) {
        rsh_engine::Kept& k = e->kept[idx];
        k.kind = kind; k.slots = S; k.set = true;
        size_t bytes = kind == 1 ? (size_t)S * 4 : (size_t)S * 8;
        LANE_OK(k.buf.reserve(bytes));
        LANE_OK(hipMemcpyAsync(k.buf.p, closest.p, bytes, hipMemcpyDeviceToDevice,
                               L.stream));
    }
    return RSH_OK;
}

}  // namespace

extern "C" {

rsh_status rsh_engine_create(rsh_engine** out, int device) {
    if (!out) return RSH_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return RSH_HIP_ERROR;
    if (hipSetDevice(device) != hipSuccess) return RSH_HIP_ERROR;
    rsh_engine* e = new rsh_engine();
    e->device = device;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) { delete e; return RSH_HIP_ERROR; }
    e->own_stream = true;
    *out = e;
    return RSH_OK;
}

void rsh_engine_destroy(rsh_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->copy_stream) {
        (void)hipStreamSynchronize(e->copy_stream);
        (void)hipStreamDestroy(e->copy_stream);
    }
    if (e->frame_done) (void)hipEventDestroy(e->frame_done);
    if (e->copy_done) (void)hipEventDestroy(e->copy_done);
    for (Lane& L : e->lane) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        for (DevBuf* b : L.bufs()) b->release();
        L.sc.release();
        if (L.hstage) (void)hipHostFree(L.hstage);
        if (L.ord_host) (void)hipHostFree(L.ord_host);
        if (L.hsize) (void)hipHostFree(L.hsize);
        if (L.htable) (void)hipHostFree(L.htable);
        L.w_done.release();
        if (L.rb_ev) (void)hipEventDestroy(L.rb_ev);
        for (hipEvent_t ev : L.ev_pool) (void)hipEventDestroy(ev);
        if (L.own_stream && L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (hipEvent_t ev : e->sync_ev) (void)hipEventDestroy(ev);
    for (DevBuf* b : {&e->tri_pos, &e->tri_nrm, &e->tri_mat, &e->mat_dif, &e->mat_spec, &e->light_pos, &e->light_col,
                      &e->light_int, &e->tri_sphere, &e->tri_mt, &e->obj_sphere, &e->obj_range, &e->cl_sphere,
                      &e->sub_sphere, &e->sub_use, &e->cl_base, &e->all_list, &e->rays_a, &e->rays_b, &e->closest[0],
                      &e->closest[1], &e->shadow[0], &e->shadow[1], &e->pixels, &e->views_d, &e->crange, &e->srange, &e->tri_pos0,
                      &e->tri_nrm0, &e->tri_obj, &e->sph0, &e->mats, &e->sph_tmp})
        b->release();
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (int i = 0; i < 4; i++) { e->g0[i].release(); e->gb[0][i].release(); e->gb[1][i].release(); }
    for (auto& k : e->kept) k.release();
    e->frame_rec.release();
    if (e->hrec) (void)hipHostFree(e->hrec);
    if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* rsh_last_error(const rsh_engine* e) { return e ? e->err.c_str() : "null engine"; }

rsh_status rsh_engine_set_stream(rsh_engine* e, void* stream) {
    if (!e) return RSH_INVALID_ARG;
    if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
    e->stream = static_cast<hipStream_t>(stream);
    e->own_stream = false;
    return RSH_OK;
}

rsh_status rsh_engine_set_wait_timeout(rsh_engine* e, uint32_t ms) {
    if (!e || ms == 0) return e ? (fail(e, "set_wait_timeout: the deadline must be positive"), RSH_INVALID_ARG)
                                : RSH_INVALID_ARG;
    for (Lane& l : e->lane) l.wait_ms = ms;
    return RSH_OK;
}

namespace {
void sleep_host_fn(void* ms) { std::this_thread::sleep_for(std::chrono::milliseconds(*static_cast<uint32_t*>(ms))); }
}  // namespace

// The bounded wait on a word that is never published: a stream held busy by
// a host function for block_ms, a deadline of wait_ms.  The host function
// ends by itself, so nothing is left running.
rsh_status rsh_selfcheck_wait_deadline(uint32_t wait_ms, uint32_t block_ms, uint32_t* elapsed_ms) {
    if (wait_ms == 0 || !elapsed_ms) return RSH_INVALID_ARG;
    Lane L;
    L.wait_ms = wait_ms;
    if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess) return RSH_HIP_ERROR;
    static uint32_t block;
    block = block_ms;
    uint32_t* word = nullptr;
    rsh_status st = RSH_HIP_ERROR;
    if (hipHostMalloc(reinterpret_cast<void**>(&word), 64, hipHostMallocCoherent) == hipSuccess) {
        word[0] = 0;
        if (hipLaunchHostFunc(L.stream, sleep_host_fn, &block) == hipSuccess) {
            const auto t0 = std::chrono::steady_clock::now();
            const hipError_t r = wait_word(L, word, 1u);
            *elapsed_ms = (uint32_t)std::chrono::duration_cast<std::chrono::milliseconds>(
                              std::chrono::steady_clock::now() - t0).count();
            st = r == WAIT_TIMEOUT ? RSH_INTERNAL : r == hipSuccess ? RSH_OK : RSH_HIP_ERROR;
        }
        (void)hipStreamSynchronize(L.stream);
        (void)hipHostFree(word);
    }
    (void)hipStreamDestroy(L.stream);
    L.stream = nullptr;
    return st;
}

rsh_status rsh_engine_load_scene(rsh_engine* e, const rsh_scene_desc* s) {
    if (!e || !s || !s->tri_pos || !s->tri_nrm || !s->tri_mat || !s->mat_diffuse || !s->mat_specular ||
        !s->object_spheres || !s->light_pos || !s->light_col || !s->light_int)
        return e ? (fail(e, "load_scene: null argument"), RSH_INVALID_ARG) : RSH_INVALID_ARG;
    if (s->triangle_total == 0 || s->object_total == 0 || s->light_total == 0 || s->material_total == 0) {
        e->err = "load_scene: empty scene";
        return RSH_INVALID_ARG;
    }
    // Objects must tile [0, T) in order (the culling closed form relies on it).
    uint32_t next = 0;
    for (uint32_t k = 0; k < s->object_total; k++) {
        float fy = s->object_spheres[6 * k + 4], fz = s->object_spheres[6 * k + 5];
        if (!(fy >= 0.0f) || !(fz >= fy) || (uint32_t)fy != next || fz >= (float)s->triangle_total) {
            e->err = "load_scene: object triangle ranges must tile [0, T) in order";
            return RSH_INVALID_ARG;
        }
        next = (uint32_t)fz + 1;
    }
    if (next != s->triangle_total) { e->err = "load_scene: objects do not cover every triangle"; return RSH_INVALID_ARG; }
    for (uint32_t i = 0; i < 3 * s->triangle_total; i++) {
        if (s->tri_mat[i] < 0 || (uint32_t)s->tri_mat[i] >= s->material_total) {
            e->err = "load_scene: material id out of range";
            return RSH_INVALID_ARG;
        }
    }
    HIP_OK(hipSetDevice(e->device));
    const uint32_t T = s->triangle_total, K = s->object_total, M = s->material_total, L = s->light_total;
    e->T = T; e->K = K; e->M = M; e->L = L;
    HIP_OK(e->tri_pos.reserve((size_t)T * 48)); HIP_OK(e->tri_nrm.reserve((size_t)T * 48));
    HIP_OK(e->tri_mat.reserve((size_t)T * 12));
    HIP_OK(e->mat_dif.reserve((size_t)M * 16)); HIP_OK(e->mat_spec.reserve((size_t)M * 16));
    HIP_OK(e->light_pos.reserve((size_t)L * 16)); HIP_OK(e->light_col.reserve((size_t)L * 16));
    HIP_OK(e->light_int.reserve((size_t)L * 8));
    HIP_OK(e->tri_sphere.reserve((size_t)T * 16)); HIP_OK(e->tri_mt.reserve((size_t)T * 48));
    HIP_OK(e->obj_sphere.reserve((size_t)K * 16)); HIP_OK(e->obj_range.reserve((size_t)K * 16));
    for (Lane& l : e->lane) {
        HIP_OK(l.obj_order.reserve((size_t)K * 4 + 4));
        if (l.ord_host) { (void)hipHostFree(l.ord_host); l.ord_host = nullptr; }
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&l.ord_host), (size_t)K * 4 + 4, hipHostMallocDefault));
    }
    hipStream_t st = e->stream;
    HIP_OK(hipMemcpyAsync(e->tri_pos.p, s->tri_pos, (size_t)T * 48, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->tri_nrm.p, s->tri_nrm, (size_t)T * 48, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->tri_mat.p, s->tri_mat, (size_t)T * 12, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->mat_dif.p, s->mat_diffuse, (size_t)M * 16, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->mat_spec.p, s->mat_specular, (size_t)M * 16, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->light_pos.p, s->light_pos, (size_t)L * 16, hipMemcpyHostToDevice, st));
    e->host_light_pos.assign(s->light_pos, s->light_pos + 4 * (size_t)L);
    HIP_OK(hipMemcpyAsync(e->light_col.p, s->light_col, (size_t)L * 16, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->light_int.p, s->light_int, (size_t)L * 8, hipMemcpyHostToDevice, st));
    e->host_spheres.assign(s->object_spheres, s->object_spheres + 6 * (size_t)K);
    std::vector<float> s4(4 * (size_t)K);
    for (uint32_t k = 0; k < K; k++)
        for (int i = 0; i < 4; i++) s4[4 * k + i] = s->object_spheres[6 * k + i];
    HIP_OK(hipMemcpyAsync(e->obj_sphere.p, s4.data(), s4.size() * 4, hipMemcpyHostToDevice, st));
    launch_triangle_prep(f4p(e->tri_pos), T, f4p(e->tri_sphere), f4p(e->tri_mt), st);
    // clusters of CLUSTER consecutive triangles per object, in object order
    std::vector<uint32_t> base(K), ident(4 * (size_t)K);
    std::vector<uint32_t> crange;
    uint32_t nc = 0;
    for (uint32_t k = 0; k < K; k++) {
        const uint32_t first = (uint32_t)s->object_spheres[6 * k + 4], last = (uint32_t)s->object_spheres[6 * k + 5];
        base[k] = nc;
        ident[4 * k] = first; ident[4 * k + 1] = last; ident[4 * k + 2] = nc; ident[4 * k + 3] = 0;
        for (uint32_t t = first; t <= last; t += CLUSTER) {
            crange.push_back(t);
            crange.push_back(t + CLUSTER - 1 < last ? t + CLUSTER - 1 : last);
            nc++;
        }
    }
    e->NC = nc;
    // sub-clusters of each cluster; missing ones (short clusters) get an empty
    // range, are never tested and hold a NaN sphere
    std::vector<uint32_t> srange((size_t)nc * SUB_PER_CLUSTER * 2);
    for (uint32_t c = 0; c < nc; c++)
        for (uint32_t k = 0; k < SUB_PER_CLUSTER; k++) {
            const uint32_t a = crange[2 * c] + k * SUBCLUSTER, b = crange[2 * c + 1];
            srange[2 * ((size_t)c * SUB_PER_CLUSTER + k)] = a <= b ? a : 1u;
            srange[2 * ((size_t)c * SUB_PER_CLUSTER + k) + 1] = a <= b ? (a + SUBCLUSTER - 1 < b ? a + SUBCLUSTER - 1 : b) : 0u;
        }
    HIP_OK(e->cl_sphere.reserve((size_t)nc * 16)); HIP_OK(e->cl_base.reserve((size_t)K * 4));
    HIP_OK(e->sub_sphere.reserve((size_t)nc * SUB_PER_CLUSTER * 16)); HIP_OK(e->sub_use.reserve((size_t)nc + 4));
    HIP_OK(e->all_list.reserve((size_t)K * 16));
    HIP_OK(e->crange.reserve(crange.size() * 4));
    HIP_OK(e->srange.reserve(srange.size() * 4));
    HIP_OK(hipMemcpyAsync(e->crange.p, crange.data(), crange.size() * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->srange.p, srange.data(), srange.size() * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->cl_base.p, base.data(), (size_t)K * 4, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(e->all_list.p, ident.data(), (size_t)K * 16, hipMemcpyHostToDevice, st));
    build_clusters(e, st);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
    e->orig_ready = false;
    e->ranges_tb = 0xFFFFFFFFu;
    e->scene_epoch++;
    return RSH_OK;
}

}  // extern "C"

namespace {

// One frame over n views (n == 1: a W x H image; n > 1: a batch laid out as
// one image of P = sum(w*h) pixels, W = P, H = 1).  hv[i].off must be the
// prefix sums of w*h.
rsh_status render_frame(rsh_engine* e, const ViewDesc* hv, uint32_t nv, const rsh_frame_config& cfg, uint32_t fl,
                        rsh_frame_stats* st) {
    uint64_t P64 = 0;
    for (uint32_t i = 0; i < nv; i++) P64 += (uint64_t)hv[i].w * hv[i].h;
    const uint32_t width = nv == 1 ? hv[0].w : (uint32_t)P64, height = nv == 1 ? hv[0].h : 1u;
    if (width == 0 || height == 0 || cfg.leaf == 0 || cfg.branch == 0 || cfg.depth == 0 || cfg.depth > 4 ||
        cfg.key_layout > 2 ||
        cfg.primary_rays != 1 || cfg.algorithm > 1 || cfg.key_mode > 1 || cfg.hit_mode > 1 || cfg.cone_mode > 1 ||
        cfg.soft_shadows > 1) {
        e->err = "render: invalid configuration";
        return RSH_INVALID_ARG;
    }
    if (cfg.key_mode == RSH_KEY_OCT && (cfg.cell_bits > 8 || cfg.dir_bits > 10 || cfg.cell_bits == 0 ||
                                        cfg.dir_bits == 0 || 2 + 3 * cfg.cell_bits + 2 * cfg.dir_bits > 32)) {
        e->err = "render: octahedral key does not fit 32 bits";
        return RSH_INVALID_ARG;
    }
    if (cfg.hit_mode == RSH_HIT_REF && e->T > (1u << 20)) {
        e->err = "render: reference hit packing holds 20-bit triangle ids";
        return RSH_CAPACITY;
    }
    if (cfg.segment_shift) {
        const uint32_t hb = cfg.key_mode == RSH_KEY_OCT ? 2 + 3 * cfg.cell_bits + 2 * cfg.dir_bits : 32;
        if (cfg.key_mode != RSH_KEY_OCT || cfg.segment_shift > 31 || hb > cfg.segment_shift ||
            (nv > 1 && ((uint64_t)(nv - 1) >> (32 - cfg.segment_shift)) != 0)) {
            e->err = "render: the view id does not fit above the octahedral key at segment_shift";
            return RSH_INVALID_ARG;
        }
    }
    // primary keys: 8x8 tiles of a view (Morton order inside), the view id at
    // segment_shift; their sort needs only the bits they use
    uint64_t tile_keys = 1;
    for (uint32_t i = 0; i < nv; i++)
        tile_keys = std::max<uint64_t>(tile_keys, (uint64_t)((hv[i].w + 7) / 8) * ((hv[i].h + 7) / 8) * 64);
    if (cfg.segment_shift && tile_keys > (1ull << cfg.segment_shift)) {
        e->err = "render: a view's primary tile keys do not fit below segment_shift";
        return RSH_INVALID_ARG;
    }
    auto bit_length = [](uint64_t v) { uint32_t b = 0; while (v) { b++; v >>= 1; } return b; };
    uint32_t primary_bits = std::max<uint32_t>(1u, bit_length(tile_keys - 1));
    if (cfg.segment_shift && nv > 1) primary_bits = std::min<uint32_t>(32u, cfg.segment_shift + bit_length(nv - 1));
    const uint64_t S64 = (uint64_t)width * height * e->L;
    const uint32_t smult = cfg.soft_shadows ? 16u : 1u;    // rays per shadow slot
    if (S64 * smult > (uint64_t)MAX_TILES * TILE || S64 * smult > 0xFFFFFFFFull) {
        e->err = "render: frame exceeds slot capacity";
        return RSH_CAPACITY;
    }
    HIP_OK(hipSetDevice(e->device));
    rsh_status rs = ensure_frame(e, width, height, e->L, smult);
    if (rs != RSH_OK) return rs;
    rs = build_ranges(e, cfg.triangle_batch);
    if (rs != RSH_OK) return rs;
    if ((rs = ensure_nodes(e, (uint32_t)(S64 * smult), cfg)) != RSH_OK) return rs;
    Views V{};
    if (nv == 1 && !e->dev_sized) {
        V.n = 0; V.one = hv[0];
    } else {      // a device table (a capture for any camera reads the cameras from it)
        if (!e->replaying) {
            HIP_OK(e->views_d.reserve((size_t)nv * sizeof(ViewDesc)));
            HIP_OK(hipMemcpyAsync(e->views_d.p, hv, (size_t)nv * sizeof(ViewDesc), hipMemcpyHostToDevice, e->stream));
        }
        V.d = e->views_d.as<ViewDesc>(); V.n = nv;
    }
    if (!e->replaying) {
        for (auto& k : e->kept) k.release();
        e->kept.clear();
        if (fl & 65u) e->kept.resize(RSH_MAX_PASSES);
    }
    for (Lane& l : e->lane) { l.timed.clear(); l.ev_used = 0; l.err.clear(); }
    if (st) std::memset(st, 0, sizeof(*st));
    Lane& A = e->lane[0];
    Lane& B = e->lane[1];
    hipStream_t sa = A.stream, sb = B.stream;
    const uint32_t W = width, H = height, Lm = e->L, P = W * H, S = W * H * Lm, Ssh = S * smult;
    // every key of a shadow or reflection pass has the same type (the top two
    // bits of an octahedral key), so their sorts skip those bits
    const uint32_t pass_bits = cfg.key_mode == RSH_KEY_OCT && cfg.segment_shift == 0 ? key_bits(cfg) - 2 : 0;
    const KeyCfg kc = key_cfg(cfg);
    const SceneView sv = scene_view(e);
    const bool timing = fl & 4u;
    const bool bs = cfg.bounce_shadows != 0;
    const uint32_t bounces = cfg.bounces < (RSH_MAX_PASSES - 2) / 2 ? cfg.bounces : (RSH_MAX_PASSES - 2) / 2;
    GBufView G0 = view(e, e->g0), GB[2] = {view(e, e->gb[0]), view(e, e->gb[1])};
    float* rays_a = e->rays_a.as<float>();
    float* rays_b = e->rays_b.as<float>();
    // cross-lane events: evG[k] G-buffer k written on lane 0 (k = 0: primary
    // hits, k = b: bounce b), evS[k] shadow pass of G-buffer k done on lane 1,
    // evD[k] G-buffer k shaded on lane 0 (its shadow flags free again)
    hipEvent_t* evG = e->sync_ev.data();
    hipEvent_t* evS = e->sync_ev.data() + RSH_MAX_PASSES + 2;
    hipEvent_t* evD = e->sync_ev.data() + 2 * (RSH_MAX_PASSES + 2);
    // Pass indices (statistics, kept results): primary 0, shadow of the primary
    // hits 1, then per bounce b its reflection pass and, with bounce shadows,
    // the shadow pass of its hits.
    auto idx_refl = [&](uint32_t b) { return 2u + (b - 1u) * (bs ? 2u : 1u); };
    auto idx_shadow = [&](uint32_t k) { return k == 0 ? 1u : idx_refl(k) + 1u; };
    uint32_t last_idx = 0;
    // Lane 1 runs on a thread of its own for eager frames (its read-backs then
    // block only that thread); under capture, or with flag bit 5, each shadow
    // pass is issued as soon as its G-buffer is (the order of a serial frame).
    const bool threaded = !e->replaying && !(fl & 32u);
    struct Shared {
        std::mutex m;
        std::condition_variable cv;
        uint32_t g_ready = 0, s_done = 0;
        bool a_end = false, b_end = false;
        rsh_status b_rs = RSH_OK;
        PassOut sh[RSH_MAX_PASSES];
    } X;

    auto shadow_step = [&](uint32_t k) -> rsh_status {
        Lane& L = B;
        const GBufView& G = k == 0 ? G0 : GB[k % 2];
        LANE_OK(hipStreamWaitEvent(sb, evG[k], 0));
        // shadow[k % 2] holds G-buffer k - 2's flags until that is shaded
        // (recorded before G-buffer k was published)
        if (k >= 2) LANE_OK(hipStreamWaitEvent(sb, evD[k - 2], 0));
        {
            StageTimer t(B, ST_CREATE, timing);
            // no slot-order rays: the gather recomputes each live ray
            launch_create_shadow(G.pos, G.nrm, W, H, Lm, f4p(e->light_pos), kc, V, cfg.soft_shadows, nullptr,
                                 u8p(B.flags), u32p(B.keys), nullptr, sb);
        }
        PassOut po;
        rsh_status r = trace_pass(e, B, cfg, nullptr, Ssh, 1, G.pos, fl, po, idx_shadow(k), e->shadow[k % 2],
                                  nullptr, false, pass_bits, false,
                                  e->host_light_pos.size() >= 4 ? e->host_light_pos.data() : nullptr);
        if (r != RSH_OK) return r;
        if ((r = record(e, B, st, idx_shadow(k), 1, Ssh, po, fl, e->shadow[k % 2])) != RSH_OK) return r;
        LANE_OK(hipEventRecord(evS[k], sb));
        X.sh[k] = po;
        return RSH_OK;
    };
    auto b_loop = [&]() {
        rsh_status r = RSH_OK;
        for (uint32_t k = 0;; k++) {
            {
                std::unique_lock<std::mutex> lk(X.m);
                X.cv.wait(lk, [&] { return X.g_ready > k || X.a_end; });
                if (X.g_ready <= k) break;     // lane 0 ended without G-buffer k
            }
            r = shadow_step(k);
            std::lock_guard<std::mutex> lk(X.m);
            if (r != RSH_OK) break;
            X.s_done = k + 1;
            X.cv.notify_all();
        }
        std::lock_guard<std::mutex> lk(X.m);
        X.b_rs = r;
        X.b_end = true;
        X.cv.notify_all();
    };
    // lane 0: G-buffer k is written
    auto publish = [&](uint32_t k) -> rsh_status {
        Lane& L = A;
        LANE_OK(hipEventRecord(evG[k], sa));
        last_idx = std::max(last_idx, idx_shadow(k));
        if (threaded) {
            std::lock_guard<std::mutex> lk(X.m);
            X.g_ready = k + 1;
            X.cv.notify_all();
            return RSH_OK;
        }
        const rsh_status r = shadow_step(k);
        if (r != RSH_OK) return r;
        X.s_done = k + 1;
        // serial passes: lane 0 resumes once the shadow pass has ended
        if (fl & 32u) LANE_OK(hipStreamWaitEvent(sa, evS[k], 0));
        return RSH_OK;
    };
    // lane 0: wait for the shadow pass of G-buffer k
    auto await_shadow = [&](uint32_t k, PassOut& po) -> rsh_status {
        Lane& L = A;
        if (threaded) {
            std::unique_lock<std::mutex> lk(X.m);
            X.cv.wait(lk, [&] { return X.s_done > k || X.b_end; });
            if (X.s_done <= k) return X.b_rs != RSH_OK ? X.b_rs : RSH_INTERNAL;
        }
        po = X.sh[k];
        LANE_OK(hipStreamWaitEvent(sa, evS[k], 0));
        return RSH_OK;
    };
    // lane 0: shading of G-buffer k's hits, after its shadow pass
    auto shade = [&](uint32_t k) -> rsh_status {
        PassOut ps;
        if (k == 0) {
            rsh_status r = await_shadow(0, ps);
            if (r != RSH_OK) return r;
            if (ps.N > 0) {
                StageTimer t(A, ST_SHADE, timing);
                // device-sized capture: the shadow pass's live count gates the shading
                launch_shade_primary(G0, W, H, Lm, sv.lights, V, u32p(e->shadow[0]), cfg.soft_shadows, u32p(e->pixels),
                                     sa, e->dev_sized ? u32p(e->frame_rec) + 2 * (size_t)idx_shadow(0) : nullptr);
            }
            return RSH_OK;
        }
        const uint32_t* sflags = nullptr;
        if (bs) {
            rsh_status r = await_shadow(k, ps);
            if (r != RSH_OK) return r;
            sflags = u32p(e->shadow[k % 2]);
        }
        StageTimer t(A, ST_SHADE, timing);
        launch_shade_reflection(GB[k % 2], W, H, Lm, sv.lights, V, u64p(e->closest[k % 2]), cfg.hit_mode, sflags,
                                cfg.soft_shadows,
                                u32p(e->pixels), sa);
        return RSH_OK;
    };
    // lane 0: primary pass, then the mirror bounces; bounce b's hits alternate
    // between closest[b % 2] / GB[b % 2], and G-buffer k is shaded just before
    // bounce k + 2 starts (its shadow pass runs meanwhile, beside bounce k + 1)
    auto run_a = [&]() -> rsh_status {
        Lane& L = A;
        if (e->dev_sized) launch_fill_u32(u32p(e->frame_rec), 0u, 2 * RSH_MAX_PASSES + 1, sa);
        // CRSH camera rays come out of creation in key order when every key
        // is unique (one view, or the view id in the key)
        const bool presort = cfg.algorithm == RSH_CRSH && (nv == 1 || cfg.segment_shift != 0);
        {
            StageTimer t(A, ST_CREATE, timing);
            if (presort)
                launch_create_primary_sorted(V, P, f4p(A.rec0), f4p(A.rec1), sa);
            else
                launch_create_primary(V, P, kc, rays_b, u8p(A.flags), u32p(A.keys), u32p(A.values), sa);
        }
        PassOut po;
        rsh_status r = trace_pass(e, A, cfg, rays_b, P, 0, nullptr, fl, po, 0, e->closest[0], nullptr, true,
                                  primary_bits, presort, nv == 1 ? hv[0].eye : nullptr);
        if (r != RSH_OK) return r;
        if ((r = record(e, A, st, 0, 0, P, po, fl, e->closest[0])) != RSH_OK) return r;
        {
            StageTimer t(A, ST_SHADE, timing);
            // key-ordered camera rays exist only as hierarchy records: the
            // surface stage recomputes them from the views
            launch_surface(presort ? nullptr : rays_b, u64p(e->closest[0]), cfg.hit_mode, W, H, 1, 1, sv, G0, sa,
                           &V);
        }
        if ((r = publish(0)) != RSH_OK) return r;
        // G-buffers are shaded in order, each once its shadow pass has run
        // and no later than bounce k + 2, which reuses G-buffer k's hit
        // records and G-buffer (closest[k % 2], GB[k % 2]) and whose shadow
        // pass reuses its shadow flags (shadow[k % 2], released by evD[k]):
        // the reflection chain runs a pass ahead of the shadow passes.
        int64_t made = 0, shaded = -1;     // last G-buffer written, shaded
        auto shade_upto = [&](int64_t j) -> rsh_status {
            if (shaded < 0 && j >= 0) {
                // the pixels are cleared right before the first shading, once
                // an asynchronous copy of the last frame's has ended
                if (e->copy_pending) LANE_OK(hipStreamWaitEvent(sa, e->copy_done, 0));
                launch_fill_u32(u32p(e->pixels), 0u, P, sa);
            }
            while (shaded < j && shaded < made) {
                const rsh_status rr = shade((uint32_t)(shaded + 1));
                if (rr != RSH_OK) return rr;
                shaded++;
                LANE_OK(hipEventRecord(evD[shaded], sa));
            }
            return RSH_OK;
        };
        for (uint32_t b = 1; b <= bounces; b++) {
            if ((r = shade_upto((int64_t)b - 2)) != RSH_OK) return r;
            DevBuf& cur = e->closest[b % 2];
            {
                StageTimer t(A, ST_CREATE, timing);
                if (b == 1)
                    launch_create_reflection(G0.pos, G0.nrm, G0.dif, W, H, Lm, V, cfg.reflect_gate, 1, kc, rays_a,
                                             u8p(A.flags), u32p(A.keys), nullptr, sa);
                else
                    launch_create_bounce(GB[(b - 1) % 2], W, H, Lm, u64p(e->closest[(b - 1) % 2]), cfg.hit_mode,
                                         cfg.reflect_gate, kc, V, rays_a, u8p(A.flags), u32p(A.keys), nullptr,
                                         sa);
            }
            PassOut pr;
            if ((r = trace_pass(e, A, cfg, rays_a, S, 0, nullptr, fl, pr, idx_refl(b), cur, nullptr, false,
                                pass_bits)) != RSH_OK)
                return r;
            if ((r = record(e, A, st, idx_refl(b), 2, S, pr, fl, cur)) != RSH_OK) return r;
            last_idx = std::max(last_idx, idx_refl(b));
            if (pr.N == 0) break;
            {
                StageTimer t(A, ST_SHADE, timing);
                launch_surface(rays_a, u64p(cur), cfg.hit_mode, W, H, Lm, 0, sv, GB[b % 2], sa);
            }
            made = b;
            if (bs && (r = publish(b)) != RSH_OK) return r;
        }
        if ((r = shade_upto(made)) != RSH_OK) return r;
        LANE_OK(hipGetLastError());
        return RSH_OK;
    };

    std::thread tb;
    if (threaded) tb = std::thread(b_loop);
    const rsh_status ra = run_a();
    if (threaded) {
        {
            std::lock_guard<std::mutex> lk(X.m);
            X.a_end = true;
            X.cv.notify_all();
        }
        tb.join();
    }
    if (ra != RSH_OK || (threaded && X.b_rs != RSH_OK)) {
        e->err = !A.err.empty() ? A.err : !B.err.empty() ? B.err : "render: a pass failed";
        if (!e->replaying) (void)hipDeviceSynchronize();     // nothing of this frame left in flight
        return ra != RSH_OK ? ra : X.b_rs;
    }
    if (st) {
        st->pass_count = last_idx + 1 < RSH_MAX_PASSES ? last_idx + 1 : RSH_MAX_PASSES;
        for (uint32_t i = 0; i < st->pass_count; i++) st->rays_total += st->pass_rays[i];
    }
    if (!e->replaying && (fl & 65u)) {
        uint32_t n = 0;
        while (n < e->kept.size() && (e->kept[n].set || e->kept[n].order_set)) n++;
        for (size_t i = n; i < e->kept.size(); i++) e->kept[i].release();
        e->kept.resize(n);
    }
    if (e->replaying) return RSH_OK;     // under stream capture: no synchronisation
    HIP_OK(hipStreamSynchronize(sa));
    if (st && (timing || (fl & 16u))) {
        for (Lane& l : e->lane)
            for (auto& t : l.timed) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, t.second.first, t.second.second) != hipSuccess) continue;
                if (t.first >= ST_PASS_TRAVERSE) {
                    st->stage_ms[ST_TRAVERSE] += ms;
                    if (t.first - ST_PASS_TRAVERSE < RSH_MAX_PASSES) st->pass_traverse_ms[t.first - ST_PASS_TRAVERSE] += ms;
                } else {
                    st->stage_ms[t.first] += ms;
                }
            }
    }
    return RSH_OK;
}

ViewDesc view_desc(const rsh_camera& c, uint32_t off, uint32_t w, uint32_t h) {
    ViewDesc v{};
    v.off = off; v.w = w; v.h = h;
    for (int i = 0; i < 3; i++) {
        v.eye[i] = c.eye[i]; v.dir[i] = c.direction[i]; v.right[i] = c.right[i]; v.up[i] = c.up[i];
    }
    return v;
}

void drop_graph(rsh_engine* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr; e->graph = nullptr;
    e->gsig.clear();
}

// Shadow and reflection passes of this configuration can keep their sizes on
// the device: CRSH with the 64-ray traversal kernel, single-block item scans.
bool dev_sizable(const rsh_engine* e, const rsh_frame_config& cfg, uint64_t slots) {
    if (cfg.algorithm != RSH_CRSH || cfg.depth != 2 || cfg.leaf * cfg.branch != 64 || cfg.traversal_kernel == 1)
        return false;
    if (cfg.leaf != 2 && cfg.leaf != 4 && cfg.leaf != 8 && cfg.leaf != 16 && cfg.leaf != 32) return false;
    const uint64_t n0 = (slots + cfg.leaf - 1) / cfg.leaf, ntop = (n0 + cfg.branch - 1) / cfg.branch;
    return ntop + 1 <= SMALL_SCAN && e->K > 0;
}

// Flag bit 3: replay a captured graph of the frame when one exists for the
// same configuration, scene and view sizes (device-sized graphs: any cameras;
// otherwise the same cameras) and no device buffer has been reallocated since;
// otherwise render eagerly (recording the pass sizes), then capture a replay
// for the next call.  A device-sized graph keeps the shadow and reflection
// passes' sizes in device memory and sizes their launches by capacities from
// the recorded frame; a replay whose work list outgrows them reports an
// overflow and the frame is rendered eagerly instead (the next call captures
// again).  Other graphs fix every pass's sizes to the recorded frame's, which
// only the same cameras repeat.
rsh_status render(rsh_engine* e, const ViewDesc* hv, uint32_t nv, const rsh_frame_config& cfg, uint32_t fl,
                  rsh_frame_stats* st) {
    if (!(fl & 8u)) return render_frame(e, hv, nv, cfg, fl, st);
    // a graph clears the pixels without waiting for an asynchronous copy
    if (e->copy_pending) {
        HIP_OK(hipEventSynchronize(e->copy_done));
        e->copy_pending = false;
    }
    if (fl & 87u) {
        e->err = "render: graph replay keeps no passes, orders, counters or timers (flag bits 0-2, 4, 6)";
        return RSH_INVALID_ARG;
    }
    if (e->stream == nullptr) { e->err = "render: graph replay needs a non-null stream"; return RSH_INVALID_ARG; }
    uint64_t P = 0;
    for (uint32_t i = 0; i < nv; i++) P += (uint64_t)hv[i].w * hv[i].h;
    const uint32_t smult = cfg.soft_shadows ? 16u : 1u;
    const bool dev = dev_sizable(e, cfg, P * e->L * smult);
    // signatures: configuration, scene and views; a device-sized graph's
    // leaves the cameras out (sizes and offsets only)
    auto signature = [&](bool cameras) {
        std::vector<uint8_t> sg(sizeof(cfg) + sizeof(e->scene_epoch) + 4 + (size_t)nv * sizeof(ViewDesc));
        uint8_t* q = sg.data();
        std::memcpy(q, &cfg, sizeof(cfg)); q += sizeof(cfg);
        std::memcpy(q, &e->scene_epoch, sizeof(e->scene_epoch)); q += sizeof(e->scene_epoch);
        std::memcpy(q, &nv, 4); q += 4;
        for (uint32_t i = 0; i < nv; i++, q += sizeof(ViewDesc)) {
            ViewDesc v = hv[i];
            if (!cameras) {
                std::memset(v.eye, 0, sizeof(v.eye)); std::memset(v.dir, 0, sizeof(v.dir));
                std::memset(v.right, 0, sizeof(v.right)); std::memset(v.up, 0, sizeof(v.up));
            }
            std::memcpy(q, &v, sizeof(ViewDesc));
        }
        return sg;
    };
    const std::vector<uint8_t> sig = signature(!(e->gexec && e->gdev));
    if (e->gexec && sig == e->gsig && alloc_epoch().load() == e->gepoch) {
        HIP_OK(hipSetDevice(e->device));
        if (e->gdev)
            HIP_OK(hipMemcpyAsync(e->views_d.p, hv, (size_t)nv * sizeof(ViewDesc), hipMemcpyHostToDevice, e->stream));
        HIP_OK(hipGraphLaunch(e->gexec, e->stream));
        if (e->gdev)
            HIP_OK(hipMemcpyAsync(e->hrec, e->frame_rec.p, (2 * RSH_MAX_PASSES + 1) * 4, hipMemcpyDeviceToHost,
                                  e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
        if (e->gdev && e->hrec[2 * RSH_MAX_PASSES]) {
            // the work list outgrew its capacities: render this frame eagerly
            drop_graph(e);
            return render_frame(e, hv, nv, cfg, 0, st);
        }
        if (st) {
            *st = e->gstats;
            st->graph_replayed = 1;
            if (e->gdev) {
                st->rays_total = 0;
                for (uint32_t i = 0; i < st->pass_count; i++) {
                    if (i > 0 || cfg.algorithm != RSH_CRSH) {      // the camera pass is fixed by the view sizes
                        st->pass_rays[i] = e->hrec[2 * i];
                        st->pass_chunks[i] = e->hrec[2 * i + 1];
                    }
                    st->rays_total += st->pass_rays[i];
                }
            }
        }
        return RSH_OK;
    }
    drop_graph(e);
    rsh_frame_stats local{};
    for (Lane& l : e->lane) { l.plan.clear(); l.plan_pos = 0; }
    e->recording = true;
    rsh_status rs = render_frame(e, hv, nv, cfg, 0, &local);
    e->recording = false;
    if (rs != RSH_OK) return rs;
    if (st) *st = local;
    // every pass recorded (no early end), so a capture can issue all of them
    const uint32_t full = 2 + cfg.bounces * (cfg.bounce_shadows ? 2u : 1u);
    const bool use_dev = dev && local.pass_count == full && cfg.bounces < (RSH_MAX_PASSES - 2) / 2;
    if (use_dev) {
        // buffers the capture uses, at their capacities, before it starts
        HIP_OK(e->frame_rec.reserve((2 * RSH_MAX_PASSES + 1) * 4));
        if (!e->hrec) HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&e->hrec), (2 * RSH_MAX_PASSES + 1) * 4,
                                           hipHostMallocDefault));
        HIP_OK(e->views_d.reserve((size_t)nv * sizeof(ViewDesc)));
        HIP_OK(hipMemcpyAsync(e->views_d.p, hv, (size_t)nv * sizeof(ViewDesc), hipMemcpyHostToDevice, e->stream));
        for (int li = 0; li < 2; li++) {
            Lane& L = e->lane[li];
            const uint64_t S = li == 0 ? P * e->L : P * e->L * smult;
            const uint64_t n0 = (S + cfg.leaf - 1) / cfg.leaf, ntop = (n0 + cfg.branch - 1) / cfg.branch;
            uint64_t lc = 0, ic = 0;
            for (const PassPlan& pp : L.plan) {
                lc = std::max<uint64_t>(lc, std::max<uint64_t>(list_capacity(pp), pp.list_total));
                ic = std::max<uint64_t>(ic, std::max<uint64_t>(items_capacity(pp), pp.items));
            }
            HIP_OK(L.psz.reserve(sizeof(PassSizes)));
            HIP_OK(reserve_items(L, (uint32_t)ntop, e->K, lc, ic));
        }
        HIP_OK(hipStreamSynchronize(e->stream));
    }
    const uint64_t epoch = alloc_epoch().load();
    e->replaying = true;
    e->dev_sized = use_dev;
    for (Lane& l : e->lane) l.plan_pos = 0;
    HIP_OK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
    rs = render_frame(e, hv, nv, cfg, 0, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    e->replaying = false;
    e->dev_sized = false;
    if (rs != RSH_OK) { if (g) (void)hipGraphDestroy(g); return rs; }
    HIP_OK(ce);
    if (alloc_epoch().load() != epoch || e->lane[0].plan_pos != e->lane[0].plan.size() ||
        e->lane[1].plan_pos != e->lane[1].plan.size()) {
        (void)hipGraphDestroy(g);
        e->err = "render: the captured frame did not repeat the recorded one";
        return RSH_INTERNAL;
    }
    const hipError_t ie = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) { (void)hipGraphDestroy(g); e->gexec = nullptr; HIP_OK(ie); }
    e->graph = g; e->gsig = signature(!use_dev); e->gepoch = epoch; e->gstats = local; e->gdev = use_dev;
    return RSH_OK;
}

}  // namespace

extern "C" {

rsh_status rsh_engine_render(rsh_engine* e, const rsh_camera* cam, const rsh_frame_config* cfgp, uint32_t fl,
                             rsh_frame_stats* st) {
    if (!e || !cam || !cfgp) return RSH_INVALID_ARG;
    if (e->T == 0) { e->err = "render: no scene loaded"; return RSH_INVALID_ARG; }
    const ViewDesc v = view_desc(*cam, 0, cfgp->width, cfgp->height);
    return render(e, &v, 1, *cfgp, fl, st);
}

rsh_status rsh_engine_render_views(rsh_engine* e, const rsh_view* views, uint32_t n, const rsh_frame_config* cfgp,
                                   uint32_t fl, rsh_frame_stats* st) {
    if (!e || !views || !cfgp || n == 0) return RSH_INVALID_ARG;
    if (e->T == 0) { e->err = "render_views: no scene loaded"; return RSH_INVALID_ARG; }
    std::vector<ViewDesc> hv(n);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (views[i].width == 0 || views[i].height == 0) { e->err = "render_views: empty view"; return RSH_INVALID_ARG; }
        if (off > 0xFFFFFFFFull) { e->err = "render_views: batch exceeds 2^32 pixels"; return RSH_CAPACITY; }
        hv[i] = view_desc(views[i].camera, (uint32_t)off, views[i].width, views[i].height);
        off += (uint64_t)views[i].width * views[i].height;
    }
    if (off > 0xFFFFFFFFull) { e->err = "render_views: batch exceeds 2^32 pixels"; return RSH_CAPACITY; }
    return render(e, hv.data(), n, *cfgp, fl, st);
}

rsh_status rsh_engine_trace(rsh_engine* e, const float* rays, const float* tmax, const uint32_t* segments, uint32_t n,
                            uint32_t kind, const rsh_frame_config* cfgp, void* out, uint32_t fl, rsh_frame_stats* st) {
    if (!e || !rays || !cfgp || !out || n == 0 || kind > 1 || (kind == 1 && !tmax)) {
        if (e) e->err = "trace: null buffer, empty query, kind > 1 or any-hit without tmax";
        return RSH_INVALID_ARG;
    }
    if (fl & 9u) { e->err = "trace: flags take counters (bit 1) and timing (bit 2) only"; return RSH_INVALID_ARG; }
    if (e->T == 0) { e->err = "trace: no scene loaded"; return RSH_INVALID_ARG; }
    const rsh_frame_config cfg = *cfgp;
    if (cfg.leaf == 0 || cfg.branch == 0 || cfg.depth == 0 || cfg.depth > 4 || cfg.algorithm > 1 ||
        cfg.key_mode > 1 || cfg.hit_mode > 1 || cfg.cone_mode > 1 || cfg.key_layout > 2) {
        e->err = "trace: invalid configuration";
        return RSH_INVALID_ARG;
    }
    if (cfg.key_mode == RSH_KEY_OCT && (cfg.cell_bits > 8 || cfg.dir_bits > 10 || cfg.cell_bits == 0 ||
                                        cfg.dir_bits == 0 || 2 + 3 * cfg.cell_bits + 2 * cfg.dir_bits > 32)) {
        e->err = "trace: octahedral key does not fit 32 bits";
        return RSH_INVALID_ARG;
    }
    if (segments && (cfg.key_mode != RSH_KEY_OCT || cfg.segment_shift == 0 || cfg.segment_shift > 31 ||
                     2 + 3 * cfg.cell_bits + 2 * cfg.dir_bits > cfg.segment_shift)) {
        e->err = "trace: segment ids need an octahedral key below segment_shift";
        return RSH_INVALID_ARG;
    }
    if (cfg.hit_mode == RSH_HIT_REF && e->T > (1u << 20)) {
        e->err = "trace: reference hit packing holds 20-bit triangle ids";
        return RSH_CAPACITY;
    }
    if ((uint64_t)n > (uint64_t)MAX_TILES * TILE) { e->err = "trace: query exceeds slot capacity"; return RSH_CAPACITY; }
    HIP_OK(hipSetDevice(e->device));
    rsh_status rs = ensure_frame(e, n, 1, 1);
    if (rs != RSH_OK) return rs;
    if ((rs = build_ranges(e, cfg.triangle_batch)) != RSH_OK) return rs;
    if ((rs = ensure_nodes(e, n, cfg)) != RSH_OK) return rs;
    Lane& L = e->lane[0];
    L.timed.clear();
    L.ev_used = 0;
    L.err.clear();
    if (st) std::memset(st, 0, sizeof(*st));
    hipStream_t s = L.stream;
    const bool timing = fl & 4u;
    {
        StageTimer t(L, ST_CREATE, timing);
        launch_query_keys(rays, segments, n, key_cfg(cfg), kind == 1 ? 1u : 0u, u8p(L.flags), u32p(L.keys), nullptr,
                          s);
        HIP_OK(hipGetLastError());
    }
    PassOut po;
    if ((rs = trace_pass(e, L, cfg, rays, n, kind, nullptr, fl, po, 0, e->closest[0], tmax)) != RSH_OK) {
        e->err = L.err;
        return rs;
    }
    if (st && st->pass_count < RSH_MAX_PASSES) {
        const uint32_t i = st->pass_count++;
        st->pass_kind[i] = kind; st->pass_slots[i] = n; st->pass_rays[i] = po.N; st->pass_chunks[i] = po.C;
        st->pass_pairs_top[i] = po.top; st->pass_pairs_leaf[i] = po.leaf; st->pass_ray_tests[i] = po.tests;
        st->pass_top_survivors[i] = po.surv; st->pass_items[i] = po.items; st->pass_chunk[i] = po.chunk;
        st->pass_splits[i] = po.splits; st->pass_tri_tests[i] = po.tri_tests; st->pass_cluster_tests[i] = po.cl_tests;
        st->pass_leaf_steps[i] = po.b_steps; st->pass_ray_steps[i] = po.c_steps; st->pass_mt_tests[i] = po.mt_tests;
        st->pass_mode[i] = po.mode;
        st->rays_total = po.N;
    }
    HIP_OK(hipMemcpyAsync(out, e->closest[0].p, (size_t)n * (kind == 1 ? 4 : 8),
                          hipMemcpyDeviceToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    if (st && (timing || (fl & 16u))) {
        for (auto& t : L.timed) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, t.second.first, t.second.second) != hipSuccess) continue;
            const int k = t.first >= ST_PASS_TRAVERSE ? ST_TRAVERSE : t.first;
            st->stage_ms[k] += ms;
            if (t.first >= ST_PASS_TRAVERSE) st->pass_traverse_ms[0] += ms;
        }
    }
    return RSH_OK;
}

rsh_status rsh_engine_read_pixels_async(rsh_engine* e, uint32_t* host, size_t count) {
    if (!e || !host || count > (size_t)e->W * e->H) return RSH_INVALID_ARG;
    HIP_OK(hipSetDevice(e->device));
    if (!e->copy_stream) HIP_OK(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
    if (!e->frame_done) HIP_OK(hipEventCreateWithFlags(&e->frame_done, hipEventDisableTiming));
    if (!e->copy_done) HIP_OK(hipEventCreateWithFlags(&e->copy_done, hipEventDisableTiming));
    // behind the frame (and any earlier copy, on the same copy stream)
    HIP_OK(hipEventRecord(e->frame_done, e->stream));
    HIP_OK(hipStreamWaitEvent(e->copy_stream, e->frame_done, 0));
    HIP_OK(hipMemcpyAsync(host, e->pixels.p, count * 4, hipMemcpyDeviceToHost, e->copy_stream));
    HIP_OK(hipEventRecord(e->copy_done, e->copy_stream));
    e->copy_pending = true;
    return RSH_OK;
}

rsh_status rsh_engine_wait_pixels(rsh_engine* e) {
    if (!e) return RSH_INVALID_ARG;
    if (e->copy_pending) {
        HIP_OK(hipEventSynchronize(e->copy_done));
        e->copy_pending = false;
    }
    return RSH_OK;
}

rsh_status rsh_engine_read_pixels(rsh_engine* e, uint32_t* host, size_t count) {
    if (!e || !host || count > (size_t)e->W * e->H) return RSH_INVALID_ARG;
    // on the engine's stream, after the frame's shading (page-locked host
    // memory from rsh_host_alloc makes this one DMA transfer)
    HIP_OK(hipMemcpyAsync(host, e->pixels.p, count * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    return RSH_OK;
}

rsh_status rsh_engine_read_gbuffer(rsh_engine* e, float* dif, float* spec, float* pos, float* nrm, size_t count) {
    if (!e || count > (size_t)e->W * e->H) return RSH_INVALID_ARG;
    float* dst[4] = {pos, nrm, dif, spec};
    for (int i = 0; i < 4; i++)
        if (dst[i]) HIP_OK(hipMemcpy(dst[i], e->g0[i].p, count * 16, hipMemcpyDeviceToHost));
    return RSH_OK;
}

rsh_status rsh_engine_read_pass(rsh_engine* e, uint32_t pass, uint64_t* host, size_t count) {
    if (!e || !host || pass >= e->kept.size()) return RSH_INVALID_ARG;
    const auto& k = e->kept[pass];
    if (count > k.slots) return RSH_INVALID_ARG;
    if (k.kind == 1) {
        std::vector<uint32_t> tmp(count);
        HIP_OK(hipMemcpy(tmp.data(), k.buf.p, count * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < count; i++) host[i] = tmp[i];
    } else {
        HIP_OK(hipMemcpy(host, k.buf.p, count * 8, hipMemcpyDeviceToHost));
    }
    return RSH_OK;
}

rsh_status rsh_engine_read_pass_order(rsh_engine* e, uint32_t pass, uint32_t* keys, uint32_t* flags, size_t slots,
                                      uint32_t* order, size_t rays, uint32_t* bases, uint32_t* ckeys, size_t chunks,
                                      float* nodes, size_t node_count) {
    if (!e || pass >= e->kept.size() || !e->kept[pass].order_set) return RSH_INVALID_ARG;
    const auto& k = e->kept[pass];
    if ((keys || flags) && (!k.have_keys || slots > k.o_slots)) return RSH_INVALID_ARG;
    if ((bases || ckeys) && (!k.have_chunks || chunks > k.o_chunks)) return RSH_INVALID_ARG;
    if ((order && rays > k.o_rays) || (nodes && node_count > k.o_nodes)) return RSH_INVALID_ARG;
    HIP_OK(hipStreamSynchronize(e->stream));
    if (keys) HIP_OK(hipMemcpy(keys, k.keys.p, slots * 4, hipMemcpyDeviceToHost));
    if (flags) {       // kept as one byte per slot
        std::vector<uint8_t> fb(slots);
        HIP_OK(hipMemcpy(fb.data(), k.flags.p, slots, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < slots; i++) flags[i] = fb[i];
    }
    if (bases) HIP_OK(hipMemcpy(bases, k.bases.p, chunks * 4, hipMemcpyDeviceToHost));
    if (ckeys) HIP_OK(hipMemcpy(ckeys, k.ckeys.p, chunks * 4, hipMemcpyDeviceToHost));
    if (nodes) HIP_OK(hipMemcpy(nodes, k.nodes.p, node_count * 16, hipMemcpyDeviceToHost));
    if (order) {
        std::vector<float> rec(rays * 4);
        HIP_OK(hipMemcpy(rec.data(), k.rec.p, rays * 16, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < rays; i++) std::memcpy(order + i, &rec[4 * i + 3], 4);
    }
    return RSH_OK;
}

void* rsh_engine_pixels_device(rsh_engine* e) { return e ? e->pixels.p : nullptr; }

void* rsh_device_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}
void rsh_device_free(void* p) { if (p) (void)hipFree(p); }
int rsh_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
int rsh_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
int rsh_memset_d32(void* dst, uint32_t value, size_t count) {
    launch_fill_u32(static_cast<uint32_t*>(dst), value, count, nullptr);
    if (hipGetLastError() != hipSuccess) return -1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}
int rsh_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }
int rsh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// RayTracer.cu:448-570 (UpdateVertex, UpdateBoundingSphere) as driven per
// frame by main.cpp:1131-1196, on the engine's scene: positions and normals
// from the scene as loaded, then the per-triangle, cluster and sub-cluster
// data rebuilt; captured frame graphs are dropped.
rsh_status rsh_engine_update_objects(rsh_engine* e, const float* model, const float* normal, const float* scale) {
    if (!e || !model || !normal || !scale) return e ? (fail(e, "update_objects: null argument"), RSH_INVALID_ARG)
                                                    : RSH_INVALID_ARG;
    if (e->T == 0) { e->err = "update_objects: no scene loaded"; return RSH_INVALID_ARG; }
    HIP_OK(hipSetDevice(e->device));
    hipStream_t st = e->stream;
    const uint32_t T = e->T, K = e->K;
    if (!e->orig_ready) {
        HIP_OK(e->tri_pos0.reserve((size_t)T * 48)); HIP_OK(e->tri_nrm0.reserve((size_t)T * 48));
        HIP_OK(hipMemcpyAsync(e->tri_pos0.p, e->tri_pos.p, (size_t)T * 48, hipMemcpyDeviceToDevice, st));
        HIP_OK(hipMemcpyAsync(e->tri_nrm0.p, e->tri_nrm.p, (size_t)T * 48, hipMemcpyDeviceToDevice, st));
        std::vector<int32_t> ids(3 * (size_t)T);
        std::vector<float> s8(8 * (size_t)K);     // (centre, 1), (radius, first, last, 1) per object
        for (uint32_t k = 0; k < K; k++) {
            const float* h = &e->host_spheres[6 * (size_t)k];
            const uint32_t first = (uint32_t)h[4], last = (uint32_t)h[5];
            for (uint32_t t = first; t <= last; t++) ids[3 * (size_t)t] = ids[3 * (size_t)t + 1] = ids[3 * (size_t)t + 2] = (int32_t)k;
            float* o = &s8[8 * (size_t)k];
            o[0] = h[0]; o[1] = h[1]; o[2] = h[2]; o[3] = 1.0f;
            o[4] = h[3]; o[5] = h[4]; o[6] = h[5]; o[7] = 1.0f;
        }
        HIP_OK(e->tri_obj.reserve(ids.size() * 4)); HIP_OK(e->sph0.reserve(s8.size() * 4));
        HIP_OK(hipMemcpyAsync(e->tri_obj.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(e->sph0.p, s8.data(), s8.size() * 4, hipMemcpyHostToDevice, st));
        HIP_OK(hipStreamSynchronize(st));     // the host vectors go out of scope
        e->orig_ready = true;
    }
    const size_t mb = (size_t)K * 64;
    HIP_OK(e->mats.reserve(3 * mb)); HIP_OK(e->sph_tmp.reserve((size_t)K * 24));
    float* dm = e->mats.as<float>();
    HIP_OK(hipMemcpyAsync(dm, model, mb, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dm + 16 * (size_t)K, normal, mb, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dm + 32 * (size_t)K, scale, mb, hipMemcpyHostToDevice, st));
    launch_update_vertex(dm, dm + 16 * (size_t)K, 3 * T, f4p(e->tri_pos0), f4p(e->tri_nrm0), e->tri_obj.as<int32_t>(),
                         f4p(e->tri_pos), f4p(e->tri_nrm), st);
    // the reference's translation matrix is the model matrix (main.cpp:1172-1179)
    launch_update_sphere(dm, dm + 32 * (size_t)K, K, f4p(e->sph0), e->sph_tmp.as<float>(), st);
    HIP_OK(hipMemcpy2DAsync(e->obj_sphere.p, 16, e->sph_tmp.p, 24, 16, K, hipMemcpyDeviceToDevice, st));
    launch_triangle_prep(f4p(e->tri_pos), T, f4p(e->tri_sphere), f4p(e->tri_mt), st);
    build_clusters(e, st);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
    drop_graph(e);
    e->scene_epoch++;
    return RSH_OK;
}

}  // extern "C"
