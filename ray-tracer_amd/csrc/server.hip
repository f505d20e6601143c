// server.hip -- request loop over device-resident scenes (continuous batching).
//
// The reference renders one frame at a time from its GLUT loop
// (projects/CUDA-Engine/source/main.cpp:1234-1303).  Serving keeps several
// scenes, each with its own engine (device-resident scene, cluster tables,
// hierarchy and pass scratch) and HIP stream, and answers ray queries:
//  * submit() checks a request against the HBM budget and queues it; requests
//    are admitted in arrival order while the budget allows;
//  * one worker thread per scene takes, from every admitted request of the
//    head request's kind (closest / any hit), its next slice of at most
//    slice_rays rays, up to a ray and request cap, and traces them as one
//    batch: each slice's rays get its index in the batch as the key's segment
//    id, so the sort keeps slices apart.  A request larger than a slice stays
//    queued with the rest of its rays, so requests that arrive meanwhile join
//    the next batch beside its next slice instead of waiting for all of it;
//  * results are copied back to each caller's host buffer, the request's
//    bytes are returned to the budget (admitting waiting requests) and
//    waiters are woken.
// Requests that arrive while a batch runs join the next one; batches of
// different scenes run concurrently on their own streams.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/rsh.h"
#include "rsh_kernels.hpp"

using namespace rsh;

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes one in-flight ray may occupy: staging (ray 24, tmax 4, segment 4,
// result 8) plus the engine's per-slot pass scratch (about 200).
constexpr uint64_t BYTES_PER_RAY = 256;
// Segment ids sit above a 26-bit octahedral key: at most 64 requests per batch.
constexpr uint32_t SEGMENT_SHIFT = 26;
constexpr uint32_t MAX_BATCH_REQUESTS = 1u << (32 - SEGMENT_SHIFT);

struct Request {
    uint64_t id = 0;
    uint32_t scene = 0, kind = 0, n = 0;
    uint32_t taken = 0, finished = 0;   // rays handed to batches / traced
    const float* rays = nullptr;     // caller's host memory, valid until the request completes
    const float* tmax = nullptr;
    void* out = nullptr;
    uint64_t bytes = 0;
    bool admitted = false, done = false;
    rsh_status status = RSH_OK;
    Clock::time_point t_submit, t_start, t_done;
    bool started = false;
    uint32_t batch_requests = 0;     // most slices traced in one batch with one of this request's
};

// A slice of a request in a batch: rays [off, off + n) of the request.
struct Slice {
    Request* r = nullptr;
    uint32_t off = 0, n = 0;
};

struct SceneSlot {
    rsh_engine* eng = nullptr;
    rsh_frame_config cfg{};
    std::deque<Request*> queue;      // admitted, not yet taken by the worker
    std::condition_variable cv;
    std::thread worker;
    DevBuf rays, tmax, segs, out;    // batch staging on the device
    hipStream_t stage = nullptr;     // staging copies (synchronised before and after the trace)
    uint64_t batches = 0, rays_done = 0, requests_done = 0;
    double busy_ms = 0.0;
};

}  // namespace

struct rsh_server {
    int device = 0;
    uint64_t budget = 0, admitted_bytes = 0;
    uint32_t max_batch_rays = 0;
    uint32_t slice_rays = 1u << 20;
    std::mutex mu;
    std::condition_variable done_cv;
    std::deque<Request*> pending;                // waiting for budget, in arrival order
    std::unordered_map<uint64_t, Request*> reqs;
    std::vector<SceneSlot*> scenes;
    uint64_t next_id = 1;
    bool stop = false, started = false;
    std::string err;
};

namespace {

// Admit waiting requests in arrival order while the budget allows (mu held).
void admit_locked(rsh_server* s) {
    while (!s->pending.empty()) {
        Request* r = s->pending.front();
        if (s->admitted_bytes + r->bytes > s->budget) break;
        s->pending.pop_front();
        s->admitted_bytes += r->bytes;
        r->admitted = true;
        SceneSlot* sc = s->scenes[r->scene];
        sc->queue.push_back(r);
        sc->cv.notify_one();
    }
}

void finish_locked(rsh_server* s, std::vector<Slice>& batch, rsh_status st, SceneSlot* sc) {
    const auto now = Clock::now();
    for (Slice& b : batch) {
        Request* r = b.r;
        if (st != RSH_OK) r->status = st;
        r->finished += b.n;
        r->batch_requests = std::max(r->batch_requests, (uint32_t)batch.size());
        sc->rays_done += b.n;
        if (r->finished == r->n || st != RSH_OK) {
            if (r->done) continue;
            if (st != RSH_OK && r->taken < r->n) {     // a failed request takes no more slices
                auto it = std::find(sc->queue.begin(), sc->queue.end(), r);
                if (it != sc->queue.end()) sc->queue.erase(it);
            }
            r->done = true;
            r->t_done = now;
            s->admitted_bytes -= r->bytes;
            sc->requests_done++;
        }
    }
    admit_locked(s);
    s->done_cv.notify_all();
}

rsh_status trace_batch(rsh_server* s, SceneSlot* sc, std::vector<Slice>& batch) {
    uint64_t total = 0;
    for (const Slice& b : batch) total += b.n;
    const uint32_t kind = batch[0].r->kind;
    if (hipSetDevice(s->device) != hipSuccess) return RSH_HIP_ERROR;
    if (sc->rays.reserve(total * 24) != hipSuccess || sc->segs.reserve(total * 4) != hipSuccess ||
        sc->out.reserve(total * 8) != hipSuccess || (kind == 1 && sc->tmax.reserve(total * 4) != hipSuccess))
        return RSH_HIP_ERROR;
    hipStream_t st = sc->stage;
    rsh_engine* e = sc->eng;
    uint64_t off = 0;
    hipError_t he = hipSuccess;
    for (uint32_t j = 0; j < batch.size() && he == hipSuccess; j++) {
        const Slice& b = batch[j];
        he = hipMemcpyAsync(sc->rays.as<float>() + 6 * off, b.r->rays + 6 * (size_t)b.off, (size_t)b.n * 24,
                            hipMemcpyHostToDevice, st);
        if (he == hipSuccess && kind == 1)
            he = hipMemcpyAsync(sc->tmax.as<float>() + off, b.r->tmax + b.off, (size_t)b.n * 4, hipMemcpyHostToDevice,
                                st);
        if (he == hipSuccess) launch_fill_u32(sc->segs.as<uint32_t>() + off, j, b.n, st);
        off += b.n;
    }
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    rsh_status rs = he == hipSuccess ? RSH_OK : RSH_HIP_ERROR;
    if (rs == RSH_OK)
        rs = rsh_engine_trace(e, sc->rays.as<float>(), kind == 1 ? sc->tmax.as<float>() : nullptr,
                              sc->segs.as<uint32_t>(), (uint32_t)total, kind, &sc->cfg, sc->out.p, 0, nullptr);
    if (rs == RSH_OK) {
        off = 0;
        const size_t esz = kind == 1 ? 4 : 8;
        for (const Slice& b : batch) {
            he = hipMemcpyAsync(static_cast<char*>(b.r->out) + (size_t)b.off * esz,
                                static_cast<char*>(sc->out.p) + off * esz, (size_t)b.n * esz, hipMemcpyDeviceToHost,
                                st);
            if (he != hipSuccess) break;
            off += b.n;
        }
        if (he == hipSuccess) he = hipStreamSynchronize(st);
        if (he != hipSuccess) rs = RSH_HIP_ERROR;
    }
    return rs;
}

void worker_loop(rsh_server* s, SceneSlot* sc) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        sc->cv.wait(lk, [&] { return s->stop || !sc->queue.empty(); });
        if (s->stop && sc->queue.empty()) return;
        // the head request's kind; the next slice of every queued request of
        // that kind, in arrival order, up to the caps
        const uint32_t kind = sc->queue.front()->kind;
        std::vector<Slice> batch;
        uint64_t rays = 0;
        const auto t0 = Clock::now();
        for (auto it = sc->queue.begin(); it != sc->queue.end() && batch.size() < MAX_BATCH_REQUESTS;) {
            Request* r = *it;
            const uint32_t left = r->n - r->taken;
            const uint32_t want = left < s->slice_rays ? left : s->slice_rays;
            if (r->kind != kind || (!batch.empty() && rays + want > s->max_batch_rays)) {
                ++it;
                continue;
            }
            batch.push_back(Slice{r, r->taken, want});
            r->taken += want;
            rays += want;
            if (!r->started) { r->started = true; r->t_start = t0; }
            if (r->taken == r->n) it = sc->queue.erase(it);
            else ++it;
        }
        lk.unlock();
        const rsh_status st = trace_batch(s, sc, batch);
        lk.lock();
        sc->batches++;
        sc->busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        finish_locked(s, batch, st, sc);
    }
}

}  // namespace

extern "C" {

rsh_status rsh_server_create(rsh_server** out, int device, uint64_t hbm_budget_bytes, uint32_t max_batch_rays) {
    if (!out || hbm_budget_bytes == 0 || max_batch_rays == 0) return RSH_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return RSH_HIP_ERROR;
    rsh_server* s = new rsh_server();
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->max_batch_rays = max_batch_rays;
    *out = s;
    return RSH_OK;
}

const char* rsh_server_last_error(const rsh_server* s) { return s ? s->err.c_str() : "null server"; }

rsh_status rsh_server_set_slice(rsh_server* s, uint32_t slice_rays) {
    if (!s || slice_rays == 0) return RSH_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->slice_rays = slice_rays;
    return RSH_OK;
}

rsh_status rsh_server_add_scene(rsh_server* s, const rsh_scene_desc* scene, const rsh_frame_config* cfg,
                                uint32_t* scene_id) {
    if (!s || !scene || !cfg || !scene_id) return RSH_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) { s->err = "add_scene: scenes are added before the first submit"; return RSH_INVALID_ARG; }
    if (cfg->key_mode != RSH_KEY_OCT || 2 + 3 * cfg->cell_bits + 2 * cfg->dir_bits > SEGMENT_SHIFT) {
        s->err = "add_scene: serving needs an octahedral key of at most 26 bits (request ids sit above it)";
        return RSH_INVALID_ARG;
    }
    SceneSlot* sc = new SceneSlot();
    rsh_status rs = rsh_engine_create(&sc->eng, s->device);
    if (rs != RSH_OK) { delete sc; s->err = "add_scene: engine creation failed"; return rs; }
    if (hipStreamCreateWithFlags(&sc->stage, hipStreamNonBlocking) != hipSuccess) {
        rsh_engine_destroy(sc->eng);
        delete sc;
        s->err = "add_scene: stream creation failed";
        return RSH_HIP_ERROR;
    }
    rs = rsh_engine_load_scene(sc->eng, scene);
    if (rs != RSH_OK) {
        s->err = std::string("add_scene: ") + rsh_last_error(sc->eng);
        (void)hipStreamDestroy(sc->stage);
        rsh_engine_destroy(sc->eng);
        delete sc;
        return rs;
    }
    sc->cfg = *cfg;
    sc->cfg.segment_shift = SEGMENT_SHIFT;
    *scene_id = (uint32_t)s->scenes.size();
    s->scenes.push_back(sc);
    return RSH_OK;
}

rsh_status rsh_server_submit(rsh_server* s, uint32_t scene_id, uint32_t kind, const float* host_rays,
                             const float* host_tmax, uint32_t n, void* host_out, uint64_t* request_id) {
    if (!s || !host_rays || !host_out || !request_id || n == 0 || kind > 1 || (kind == 1 && !host_tmax))
        return RSH_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (scene_id >= s->scenes.size()) { s->err = "submit: unknown scene"; return RSH_INVALID_ARG; }
    // a request is traced in slices of at most slice_rays, each within a batch
    const uint32_t piece = n < s->slice_rays ? n : s->slice_rays;
    if (piece > s->max_batch_rays) { s->err = "submit: request slice larger than the batch cap"; return RSH_CAPACITY; }
    const uint64_t bytes = (uint64_t)n * BYTES_PER_RAY;
    if (bytes > s->budget) { s->err = "submit: request larger than the HBM budget"; return RSH_CAPACITY; }
    if (!s->started) {
        for (SceneSlot* sc : s->scenes) sc->worker = std::thread(worker_loop, s, sc);
        s->started = true;
    }
    Request* r = new Request();
    r->id = s->next_id++;
    r->scene = scene_id; r->kind = kind; r->n = n;
    r->rays = host_rays; r->tmax = host_tmax; r->out = host_out;
    r->bytes = bytes;
    r->t_submit = Clock::now();
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    admit_locked(s);
    *request_id = r->id;
    return RSH_OK;
}

rsh_status rsh_server_wait(rsh_server* s, uint64_t request_id, rsh_request_info* info) {
    if (!s) return RSH_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) { s->err = "wait: unknown request"; return RSH_INVALID_ARG; }
    Request* r = it->second;
    s->done_cv.wait(lk, [&] { return r->done; });
    if (info) {
        info->queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
        info->latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        info->batch_requests = r->batch_requests;
    }
    const rsh_status st = r->status;
    s->reqs.erase(it);
    delete r;
    return st;
}

rsh_status rsh_server_stats(rsh_server* s, rsh_server_info* info) {
    if (!s || !info) return RSH_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(info, 0, sizeof(*info));
    for (SceneSlot* sc : s->scenes) {
        info->batches += sc->batches;
        info->requests += sc->requests_done;
        info->rays += sc->rays_done;
        info->busy_ms += sc->busy_ms;
    }
    info->pending = (uint32_t)s->pending.size();
    info->admitted_bytes = s->admitted_bytes;
    return RSH_OK;
}

void rsh_server_destroy(rsh_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (SceneSlot* sc : s->scenes) sc->cv.notify_all();
    }
    for (SceneSlot* sc : s->scenes)
        if (sc->worker.joinable()) sc->worker.join();
    for (SceneSlot* sc : s->scenes) {
        sc->rays.release(); sc->tmax.release(); sc->segs.release(); sc->out.release();
        if (sc->stage) (void)hipStreamDestroy(sc->stage);
        rsh_engine_destroy(sc->eng);
        delete sc;
    }
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
}

void* rsh_host_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void rsh_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"
