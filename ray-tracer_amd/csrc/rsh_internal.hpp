// rsh_internal.hpp -- host-side declarations shared by the library's
// translation units: kernel launchers and small device-memory utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>

#include "rsh_device.hpp"

namespace rsh {

// Tile geometry of the scan / compaction / sort kernels: 256 threads x 16
// items.  The single-block scan of tile totals bounds a pass to
// MAX_TILES * TILE items (67M).
constexpr uint32_t SCAN_BLOCK = 256;
constexpr uint32_t SCAN_ITEMS = 16;
constexpr uint32_t TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr uint32_t MAX_TILES = 16384;
// Exclusive scans up to this length run in one workgroup.
constexpr uint32_t SMALL_SCAN = 131072;

inline uint32_t tiles_for(uint64_t n) { return (uint32_t)((n + TILE - 1) / TILE); }

// Sizes of one ray pass kept in device memory (device-sized passes): the
// pack scan's live rays N and chunks C, and the hierarchy levels over N.
// Kernels of such a pass read them instead of taking host counts; their
// launches are sized by capacities (slots, or the recorded pass's sizes with
// a margin), and blocks beyond the device sizes exit at once.
struct PassSizes {
    uint32_t N, C;
    uint32_t count[4], offset[4];   // nodes per level, first node of each level
    uint32_t ntop;                  // count[depth - 1]
    uint32_t pad;
};

__host__ __device__ inline uint32_t tiles_of(uint32_t n) { return (n + TILE - 1) / TILE; }

// Process-wide count of device (re)allocations made through DevBuf.  A
// captured HIP graph holds raw pointers, so it stays valid only while this
// count is unchanged.
inline std::atomic<uint64_t>& alloc_epoch() {
    static std::atomic<uint64_t> n{0};
    return n;
}

// Growable device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t b) {
        if (b <= bytes) return hipSuccess;
        alloc_epoch()++;
        if (p) (void)hipFree(p);
        p = nullptr; bytes = 0;
        hipError_t e = hipMalloc(&p, b ? b : 1);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    void release() {
        if (p) { alloc_epoch()++; (void)hipFree(p); }
        p = nullptr; bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Scratch used by scans and the radix sort.
struct ScanScratch {
    DevBuf tile_counts;   // uint32[MAX_TILES + 1]
    DevBuf hist;          // uint32[512 * MAX_TILES] (9-bit radix digits)
    DevBuf digit_totals;  // uint32[512]
    DevBuf total;         // uint32[4]
    DevBuf ping_k, ping_v;
    hipError_t reserve(uint64_t n);
    void release();
};

struct PassParams;  // defined in the kernel translation units

// -------------------------------------------------------------- launchers --
void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
void launch_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s);

// Stable compaction of (keys, values) where flags == 0.  Writes the live count
// to *d_total (device).  Returns hipSuccess or the launch error.
hipError_t run_trim(const uint32_t* flags, const uint32_t* keys, const uint32_t* values, uint64_t n,
                    uint32_t* out_keys, uint32_t* out_values, uint32_t* d_total, ScanScratch& sc,
                    hipStream_t s);
// Inclusive scan of flags (the reference's scanArray by-product).
hipError_t run_inclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s);
// Exclusive scan of u32 values.
hipError_t run_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s);
// Run-length compression of equal consecutive keys: chunk bases, sizes, keys,
// values (= chunk index).  Chunk count -> *d_total.  d_n (optional): the key
// count in device memory, at most n.
hipError_t run_compress(const uint32_t* keys, uint64_t n, uint32_t* bases, uint32_t* sizes,
                        uint32_t* ckeys, uint32_t* cvalues, uint32_t* d_total, ScanScratch& sc,
                        hipStream_t s, const uint32_t* d_n = nullptr);
// Trimming and compression in one walk (the engine's path): flags one byte
// per slot (0: live); the live slots' values (values == nullptr: the slot
// index) in order -> tvalues, one chunk per run of equal live keys -> bases
// (first live position; bases[C] = N), ckeys, cvalues (chunk index).
// d_totals[0] = N, d_totals[1] = C.  bases holds n + 1 entries.
// clear64 (if set) receives `miss` in every slot, clear32 (if set) 0: the pass's
// per-slot results start cleared.  host (if set, device-visible coherent
// host memory) receives N and C in host[1], host[2] and then `seq` in host[0]
// as soon as they are known, ahead of the chunk writes.
hipError_t run_pack(const uint8_t* flags, const uint32_t* keys, const uint32_t* values, uint64_t n,
                    uint32_t* tvalues, uint32_t* bases, uint32_t* ckeys, uint32_t* cvalues, uint32_t* d_totals,
                    unsigned long long* clear64, unsigned long long miss, uint32_t* clear32, ScanScratch& sc,
                    hipStream_t s, uint32_t* host = nullptr, uint32_t seq = 0);
// Exclusive scan of the work-item counts ceil(cand[t] / chunk), t < ntop, with
// first[ntop] = their total (counts is scratch of ntop + 1 entries); max_cand
// bounds every cand entry (the scene's cluster count).  zero[0, zero_n) is
// cleared on the way (the item table, for the next pass); with npass, its
// exclusive scan (ntop + 1 entries) goes to off in the same launch.
hipError_t run_item_scan(const uint32_t* cand, uint32_t ntop, uint32_t chunk, uint32_t max_cand, uint32_t* counts,
                         uint32_t* first, ScanScratch& sc, hipStream_t s, unsigned long long* zero = nullptr,
                         uint32_t zero_n = 0, const uint32_t* npass = nullptr, uint32_t* off = nullptr,
                         const PassSizes* ps = nullptr);
// Decompression of run_pack's chunks in sorted order: out_values receives
// tvalues[bases[v] .. bases[v + 1]) for v = sorted_cvalues[0 .. C) in turn
// (N values in all).  soff (C + 1 entries) and sbase (C) are scratch.
hipError_t run_unpack(const uint32_t* bases, const uint32_t* sorted_cvalues, const uint32_t* tvalues, uint64_t C,
                      uint64_t N, uint32_t* soff, uint32_t* sbase, uint32_t* out_values, ScanScratch& sc,
                      hipStream_t s);
// Stable LSD radix sort of (keys, values) over the low `bits` bits; result
// in out_keys / out_values (inputs preserved).
hipError_t run_radix_sort(const uint32_t* keys, const uint32_t* values, uint64_t n, uint32_t bits,
                          uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc, hipStream_t s);
// The same two stages with the element counts in device memory: the sort of
// *d_n pairs (n_cap bounds it and sizes the launches; always a separate digit
// row scan), and the decompression of ps->C chunks into ps->N rays.
hipError_t run_radix_sort_dev(const uint32_t* keys, const uint32_t* values, uint64_t n_cap, const uint32_t* d_n,
                              uint32_t bits, uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc,
                              hipStream_t s);
hipError_t run_unpack_dev(const uint32_t* bases, const uint32_t* sorted_cvalues, const uint32_t* tvalues,
                          uint64_t cap, const PassSizes* ps, uint32_t* soff, uint32_t* sbase, uint32_t* out_values,
                          ScanScratch& sc, hipStream_t s);
// Level sizes of a pass from its pack totals (N, C) into ps, and (N, C) into
// rec[0..1] when rec is set (the frame's per-pass record).
void launch_pass_sizes(const uint32_t* totals, uint32_t leaf, uint32_t branch, uint32_t depth, PassSizes* ps,
                       uint32_t* rec, hipStream_t s);
// The same for a pass whose sizes the host knows (a frame's camera pass).
void launch_fixed_sizes(uint32_t N, uint32_t C, uint32_t leaf, uint32_t branch, uint32_t depth, PassSizes* ps,
                        uint32_t* rec, hipStream_t s);
// Level sizes of a hierarchy over min(*d_n, n_cap) rays into ps (C = 0).
void launch_ray_sizes(const uint32_t* d_n, uint32_t n_cap, uint32_t leaf, uint32_t branch, uint32_t depth,
                      PassSizes* ps, hipStream_t s);
// Expand sorted chunks back to rays: for chunk x in sorted order, rays
// trimmed[bases[sv[x]] .. + sizes[sv[x]]) in order.  `scan` receives the
// exclusive scan of the sorted chunk sizes, `skeleton` the sizes themselves.
// N == 0: the ray total is not known on the host and is read on the device
// (no synchronisation).  d_c (optional): the chunk total in device memory, at
// most C.
hipError_t run_decompress(const uint32_t* bases, const uint32_t* sizes, const uint32_t* sorted_cvalues,
                          const uint32_t* tkeys, const uint32_t* tvalues, uint32_t* skeleton,
                          uint32_t* scan, uint64_t C, uint64_t N, uint32_t* out_keys,
                          uint32_t* out_values, ScanScratch& sc, hipStream_t s, const uint32_t* d_c = nullptr);

}  // namespace rsh
