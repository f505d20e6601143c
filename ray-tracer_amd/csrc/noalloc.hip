// noalloc.hip -- the radix sort stage with no allocation at all.
//
// rsh_sort_pairs_noalloc sorts (key, value) pairs like rsh_sort_pairs, with
// the sort's scratch (ping buffers of keys and values, per-tile digit counts,
// digit totals) in a file-scope device array instead of library allocations.
// It is for callers that cannot allocate: code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc.  It calls no hipMalloc,
// hipFree or allocating routine, and it does not synchronise.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two host threads) race on it.  Nothing in it must start at
// zero (every digit count is written before it is read), so it is not cleared
// between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rsh.h"
#include "../../include/rsh_engine.h"
#include "rsh_internal.hpp"

namespace {

using namespace rsh;

// Pairs per call at most: a 1920 x 1080 frame's slots fit.
constexpr uint32_t NOALLOC_MAX_PAIRS = 1u << 21;
constexpr uint32_t NOALLOC_MAX_TILES = NOALLOC_MAX_PAIRS / TILE;
constexpr size_t PING_BYTES = (size_t)NOALLOC_MAX_PAIRS * 4;
constexpr size_t HIST_BYTES = (size_t)512 * NOALLOC_MAX_TILES * 4;
constexpr size_t TOTALS_BYTES = 512 * 4;
constexpr size_t SCRATCH_BYTES = 2 * PING_BYTES + HIST_BYTES + TOTALS_BYTES;   // 17 MiB + 2 KiB

static_assert(NOALLOC_MAX_PAIRS % TILE == 0, "whole tiles");

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// The array's device address per GPU (each GPU holds its own copy), looked up
// once per device with that device current.  The array's name is a host-side
// stand-in and never a kernel argument.
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
uint8_t* scratch_ptr[MAX_DEVICES] = {};

hipError_t device_scratch(uint8_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[dev] = static_cast<uint8_t*>(p);
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

// A ScanScratch over the array: its buffers are views, never reserved or freed.
ScanScratch scratch_views(uint8_t* base) {
    ScanScratch sc;
    sc.ping_k.p = base;                                   sc.ping_k.bytes = PING_BYTES;
    sc.ping_v.p = base + PING_BYTES;                      sc.ping_v.bytes = PING_BYTES;
    sc.hist.p = base + 2 * PING_BYTES;                    sc.hist.bytes = HIST_BYTES;
    sc.digit_totals.p = base + 2 * PING_BYTES + HIST_BYTES; sc.digit_totals.bytes = TOTALS_BYTES;
    return sc;
}

}  // namespace

extern "C" int rsh_sort_pairs_noalloc(const uint32_t* keys, const uint32_t* values, uint32_t n, uint32_t bits,
                                      uint32_t* out_keys, uint32_t* out_values, void* stream) {
    if (bits == 0 || bits > 32) return RSH_INVALID_ARG;
    if (n > NOALLOC_MAX_PAIRS) return RSH_CAPACITY;
    if (n == 0) return RSH_OK;
    if (!keys || !values || !out_keys || !out_values) return RSH_INVALID_ARG;
    uint8_t* base = nullptr;
    if (device_scratch(&base) != hipSuccess) return RSH_HIP_ERROR;
    ScanScratch sc = scratch_views(base);
    const hipError_t e = run_radix_sort(keys, values, n, bits, out_keys, out_values, sc,
                                        static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RSH_OK : RSH_HIP_ERROR;
}
