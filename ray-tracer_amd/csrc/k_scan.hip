// k_scan.hip -- scans, stream compaction, run-length compression, stable LSD
// radix sort and decompression for gfx950 (wave64).
//
// Replaces the CUB calls of the reference (projects/CUDA-Engine/source/
// RayTracer.cu:2302 InclusiveSum + CreateTrimmedRays :844-884 for trimming;
// :2361 InclusiveSum + CreateChunkFlags/Bases/Sizes :886-991 for compression;
// :2401 DeviceRadixSort::SortPairs for sorting; :2452 ExclusiveSum +
// CreateSortedRaySkeleton/CreateSortedRays :993-1059 for decompression).
//
// All scans are reduce-then-scan over tiles of 4096 items (256 threads x 16):
// a tile-total kernel, one single-block scan of the tile totals, and a tile
// kernel that rescans locally and writes.  No kernel waits on another
// workgroup, so there is no spinning.  The radix sort is the same shape per
// 8-bit digit: tile histograms (LDS counters per wave), a per-digit row scan,
// and a downsweep that ranks keys stably inside each wave with ballot-based
// digit matching, stages the tile in LDS in digit order and writes runs of
// equal digits contiguously.
#include "rsh_internal.hpp"

namespace rsh {

namespace {

constexpr uint32_t WAVES = SCAN_BLOCK / 64;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x >> 6; }

__device__ __forceinline__ uint32_t wave_inclusive_add(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread across a block of SCAN_BLOCK
// threads; *total receives the block sum.  `lds` holds WAVES entries.
__device__ __forceinline__ uint32_t block_exclusive_add(uint32_t v, uint32_t* lds, uint32_t* total) {
    uint32_t inc = wave_inclusive_add(v);
    if (lane_id() == 63) lds[wave_id()] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < WAVES; w++) {
        uint32_t t = lds[w];
        pre += (w < wave_id()) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Load 16 consecutive u32 of a tile (blocked: thread t owns [16t, 16t+16)).
__device__ __forceinline__ void load16(const uint32_t* p, uint64_t base, uint64_t n, uint32_t v[16], uint32_t fill) {
    if (base + 16 <= n) {
        const uint4* q = reinterpret_cast<const uint4*>(p + base);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 a = q[i];
            v[4 * i] = a.x; v[4 * i + 1] = a.y; v[4 * i + 2] = a.z; v[4 * i + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = (base + i < n) ? p[base + i] : fill;
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------- tile totals --
// mode 0: count flags == 0 (live rays); mode 1: sum of values;
// mode 2: count chunk heads (keys[x] != keys[x-1]); mode 3: sum of flags.
// d_n (optional): the element count in device memory, bounded by n (the
// capacity the grid covers); blocks past it exit.
__device__ __forceinline__ bool device_count(const uint32_t* d_n, uint64_t& n) {
    if (!d_n) return true;
    const uint64_t dn = *d_n;
    if (dn < n) n = dn;
    return (uint64_t)blockIdx.x * TILE < n;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_tile_total(const uint32_t* in, uint64_t n, uint32_t* tile_totals,
                                                    const uint32_t* d_n) {
    __shared__ uint32_t lds[WAVES];
    if (!device_count(d_n, n)) return;
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[16];
    load16(in, base, n, v, MODE == 0 ? 1u : 0u);
    uint32_t sum = 0;
    if (MODE == 2) {
        uint32_t prev = (base > 0 && base - 1 < n) ? in[base - 1] : 0u;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            uint64_t p = base + i;
            if (p < n) sum += (p == 0 || v[i] != prev) ? 1u : 0u;
            prev = v[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) sum += MODE == 0 ? (v[i] == 0u ? 1u : 0u) : v[i];
    }
    uint32_t total;
    (void)block_exclusive_add(sum, lds, &total);
    if (threadIdx.x == 0) tile_totals[blockIdx.x] = total;
}

// Single-block exclusive scan of up to MAX_TILES tile totals (in place);
// the grand total goes to *total and to tiles[ntiles].
__device__ __forceinline__ void scan_tiles_block(uint32_t* tiles, uint32_t ntiles, uint32_t* total, uint32_t* lds) {
    const uint32_t per = (ntiles + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t i = 0; i < per; i++) if (b + i < ntiles) s += tiles[b + i];
    uint32_t inc = wave_inclusive_add(s);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t k = 0; k < 16; k++) { uint32_t t = lds[k]; pre += (k < w) ? t : 0u; tot += t; }
    uint32_t run = pre + inc - s;
    for (uint32_t i = 0; i < per; i++) {
        if (b + i < ntiles) { uint32_t t = tiles[b + i]; tiles[b + i] = run; run += t; }
    }
    if (threadIdx.x == 0) { tiles[ntiles] = tot; if (total) *total = tot; }
}

// d_n (optional): the tiles of *d_n elements, at most ntiles.
__global__ void __launch_bounds__(1024) k_scan_tiles(uint32_t* tiles, uint32_t ntiles, uint32_t* total,
                                                     const uint32_t* d_n) {
    __shared__ uint32_t lds[16];
    if (d_n) ntiles = min(ntiles, tiles_of(*d_n));
    scan_tiles_block(tiles, ntiles, total, lds);
}

// The decompression's scan of tile totals (its own name in kernel traces);
// d_n: device-sized, the tiles of *d_n chunks.
__global__ void __launch_bounds__(1024) k_unpack_scan(uint32_t* tiles, uint32_t ntiles, uint32_t* total,
                                                      const uint32_t* d_n) {
    __shared__ uint32_t lds[16];
    if (d_n) ntiles = min(ntiles, tiles_of(*d_n));
    scan_tiles_block(tiles, ntiles, total, lds);
}

// ------------------------------------------------------------- trimming --
__global__ void __launch_bounds__(256) k_trim_scatter(const uint32_t* flags, const uint32_t* keys,
                                                      const uint32_t* values, uint64_t n,
                                                      const uint32_t* tile_offsets, uint32_t* ok,
                                                      uint32_t* ov) {
    __shared__ uint32_t lds[WAVES];
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t f[16];
    load16(flags, base, n, f, 1u);
    uint32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) cnt += f[i] == 0u ? 1u : 0u;
    uint32_t total;
    uint32_t pos = tile_offsets[blockIdx.x] + block_exclusive_add(cnt, lds, &total);
    if (cnt == 0) return;
    uint32_t k[16], v[16];
    load16(keys, base, n, k, 0u);
    if (values) load16(values, base, n, v, 0u);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (f[i] == 0u) { ok[pos] = k[i]; if (ov) ov[pos] = v[i]; pos++; }
    }
}

// Tile-local scan writer: mode 0 inclusive, mode 1 exclusive.
template <int MODE>
__global__ void __launch_bounds__(256) k_scan_write(const uint32_t* in, uint64_t n, const uint32_t* tile_offsets,
                                                    uint32_t* out, const uint32_t* d_n) {
    __shared__ uint32_t lds[WAVES];
    if (!device_count(d_n, n)) return;
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[16];
    load16(in, base, n, v, 0u);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) s += v[i];
    uint32_t total;
    uint32_t run = tile_offsets[blockIdx.x] + block_exclusive_add(s, lds, &total);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (base + i < n) {
            if (MODE == 0) { run += v[i]; out[base + i] = run; }
            else { out[base + i] = run; run += v[i]; }
        }
    }
}

// ----------------------------------------------------------- compression --
__global__ void __launch_bounds__(256) k_compress_scatter(const uint32_t* keys, uint64_t n,
                                                          const uint32_t* tile_offsets, uint32_t* bases,
                                                          uint32_t* ckeys, uint32_t* cvalues, const uint32_t* d_n) {
    __shared__ uint32_t lds[WAVES];
    if (!device_count(d_n, n)) return;
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t k[16];
    load16(keys, base, n, k, 0u);
    uint32_t prev = (base > 0 && base - 1 < n) ? keys[base - 1] : 0u;
    uint32_t head = 0, cnt = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint64_t p = base + i;
        bool h = p < n && (p == 0 || k[i] != prev);
        head |= (h ? 1u : 0u) << i;
        cnt += h ? 1u : 0u;
        prev = k[i];
    }
    uint32_t total;
    uint32_t c = tile_offsets[blockIdx.x] + block_exclusive_add(cnt, lds, &total);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (head & (1u << i)) { bases[c] = (uint32_t)(base + i); ckeys[c] = k[i]; cvalues[c] = c; c++; }
    }
}

__global__ void __launch_bounds__(256) k_chunk_sizes(const uint32_t* bases, const uint32_t* d_chunks, uint32_t n,
                                                     uint32_t* sizes, const uint32_t* d_n) {
    if (d_n) n = min(n, *d_n);
    const uint32_t C = *d_chunks;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < C; x += gridDim.x * blockDim.x)
        sizes[x] = (x + 1 < C ? bases[x + 1] : n) - bases[x];
}

// ------------------------------------------ fused trim + compress (engine) --
// One walk over the ray slots does the reference's trimming (stable compaction
// of the live slots, flags == 0) and compression (one chunk per run of equal
// keys among the live rays) together.  k_pack_tiles summarises each tile
// (live count, chunk heads after its first live key, first and last live key),
// k_pack_scan turns the summaries into offsets and the key carried into each
// tile, and k_pack_write writes the live slot values and the chunk heads
// (bases, keys, values = chunk index), staged in LDS so that every global
// write is coalesced.  bases[C] = N closes the last chunk, so chunk sizes are
// bases[v + 1] - bases[v] and are not stored.  The trimmed keys are not
// written: after compression only the chunk keys are used.

struct LastKey {
    uint32_t valid, key;
};

// Block-wide scan of "the last live key so far" in thread order: returns the
// value over the threads before this one, and the whole block's in *blk.
// lds holds blockDim.x / 64 entries.
__device__ __forceinline__ LastKey block_last_key(LastKey v, LastKey* lds, LastKey* blk) {
    const uint32_t lane = lane_id(), w = wave_id(), nw = blockDim.x >> 6;
    LastKey acc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t yv = __shfl_up(acc.valid, d, 64), yk = __shfl_up(acc.key, d, 64);
        if (lane >= (uint32_t)d && !acc.valid) { acc.valid = yv; acc.key = yk; }
    }
    LastKey ex;
    ex.valid = __shfl_up(acc.valid, 1, 64);
    ex.key = __shfl_up(acc.key, 1, 64);
    if (lane == 0) ex.valid = 0;
    if (lane == 63) lds[w] = acc;
    __syncthreads();
    LastKey carry{0u, 0u}, tot{0u, 0u};
    for (uint32_t k = 0; k < nw; k++) {
        const LastKey t = lds[k];
        if (t.valid) { if (k < w) carry = t; tot = t; }
    }
    __syncthreads();
    *blk = tot;
    return ex.valid ? ex : carry;
}

// Exclusive sum over a block of any multiple of 64 threads; lds holds
// blockDim.x / 64 entries.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* lds, uint32_t* total) {
    const uint32_t w = wave_id(), nw = blockDim.x >> 6;
    const uint32_t inc = wave_inclusive_add(v);
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t k = 0; k < nw; k++) {
        const uint32_t t = lds[k];
        pre += k < w ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Load the 16 entries of a thread's slots in groups of 4, skipping groups
// with no live slot (flags != 0 everywhere): sparse passes read little more
// than their flags.  Skipped entries read as 0 and are never used.
__device__ __forceinline__ void load16_live(const uint32_t* p, uint64_t base, uint64_t n, const uint32_t f[16],
                                            uint32_t v[16]) {
    if (base + 16 <= n) {
        const uint4* q = reinterpret_cast<const uint4*>(p + base);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 a = make_uint4(0u, 0u, 0u, 0u);
            if (f[4 * i] == 0u || f[4 * i + 1] == 0u || f[4 * i + 2] == 0u || f[4 * i + 3] == 0u) a = q[i];
            v[4 * i] = a.x; v[4 * i + 1] = a.y; v[4 * i + 2] = a.z; v[4 * i + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = (base + i < n && f[i] == 0u) ? p[base + i] : 0u;
    }
}

// The 16 byte flags from base, one uint4 load (1: dead past n).
__device__ __forceinline__ void load16_u8(const uint8_t* p, uint64_t base, uint64_t n, uint32_t v[16]) {
    if (base + 16 <= n) {
        const uint4 a = *reinterpret_cast<const uint4*>(p + base);
        const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = base + i < n ? (uint32_t)p[base + i] : 1u;
    }
}

// Exclusive scan of n <= SMALL_SCAN values in one workgroup of 1024 threads:
// wave w scans a contiguous run of `wper` values (a multiple of 256), 256 per
// step with lane l holding 4 consecutive ones, so every access is coalesced
// (one workgroup sits on one CU, whose address unit would take a cycle per
// distinct cache line of a scattered access).  Pass 1 sums each wave's run,
// the run offsets come from one block-wide sum, pass 2 scans and writes.
// out[n - 1] closes the scan when the last input is 0.  MODE 1 scans item
// counts ceil(in[i] / chunk) for i < n - 1 and 0 for the last entry (the
// work-list scan of the traversal), dividing by multiply-high when `fast`
// (every count + chunk - 1 within div_hi's bound).  Each value is read before
// it is written by the same lane, so in == out is allowed.  Returns the total.
template <int MODE>
__device__ uint32_t block_scan_run(const uint32_t* in, uint32_t n, uint32_t chunk, bool fast, uint32_t* out,
                                   uint32_t* red) {
    const uint32_t lane = lane_id(), w = wave_id(), nw = blockDim.x >> 6;
    const uint32_t wper = ((n + nw * 256u - 1u) / (nw * 256u)) * 256u;
    const uint32_t wb = w * wper;
    const uint32_t we = wb + wper < n ? wb + wper : n;
    const uint32_t mc = div_hi_magic(chunk);
    auto val = [&](uint32_t i) -> uint32_t {
        if (i >= we) return 0u;
        if (MODE == 1) {
            if (i + 1u >= n) return 0u;
            const uint32_t x = in[i] + chunk - 1u;
            return fast ? div_hi(x, chunk, mc) : x / chunk;
        }
        return in[i];
    };
    uint32_t sum = 0;
#pragma unroll 4
    for (uint32_t i0 = wb + 4u * lane; i0 < we; i0 += 256u)
        sum += val(i0) + val(i0 + 1u) + val(i0 + 2u) + val(i0 + 3u);
    uint32_t tot;
    const uint32_t wsum = wave_inclusive_add(sum);       // lane 63: the wave's run total
    uint32_t carry = block_excl_sum(lane == 63 ? wsum : 0u, red, &tot);
    carry = __shfl(carry, 63, 64);
    for (uint32_t s0 = wb; s0 < we; s0 += 256u) {
        const uint32_t i = s0 + 4u * lane;
        const uint32_t v0 = val(i), v1 = val(i + 1u), v2 = val(i + 2u), v3 = val(i + 3u);
        const uint32_t ls = v0 + v1 + v2 + v3;
        const uint32_t inc = wave_inclusive_add(ls);
        uint32_t run = carry + inc - ls;
        if (i < we) out[i] = run;
        run += v0;
        if (i + 1u < we) out[i + 1u] = run;
        run += v1;
        if (i + 2u < we) out[i + 2u] = run;
        run += v2;
        if (i + 3u < we) out[i + 3u] = run;
        carry += __shfl(inc, 63, 64);
    }
    return tot;
}

// One workgroup: the scan above of `in` (MODE), then, when in2 is given, a
// plain exclusive scan of in2 (n values) into out2; zero[0, zero_n) is
// cleared on the way.
// Threads of the single-workgroup item scan.  Smaller workgroups find room
// beside the other lane's traversal waves at once (a 1024-thread one waits
// for that traversal to drain), but then both lanes reach their traversals
// together and their set-up phases no longer overlap the other lane's
// traversal: 1080p atrium frame 5.55 ms with 1024 threads, 5.66 with 256,
// 5.89 with 64.
#ifndef RSH_ITEM_SCAN_THREADS
#define RSH_ITEM_SCAN_THREADS 1024
#endif
constexpr uint32_t ITEM_SCAN_THREADS = RSH_ITEM_SCAN_THREADS;

template <int MODE>
__global__ void __launch_bounds__(1024) k_scan_block(const uint32_t* in, uint32_t n, uint32_t chunk, bool fast,
                                                     uint32_t* out, uint32_t* total, unsigned long long* zero,
                                                     uint32_t zero_n, const uint32_t* in2, uint32_t* out2,
                                                     const PassSizes* ps) {
    __shared__ uint32_t red[16];
    if (ps) n = ps->ntop + 1;             // device-sized: the pass's top nodes + 1
    if (zero)
        for (uint32_t i = threadIdx.x; i < zero_n; i += blockDim.x) zero[i] = 0ull;
    const uint32_t tot = block_scan_run<MODE>(in, n, chunk, fast, out, red);
    if (threadIdx.x == 0 && total) *total = tot;
    if (in2) (void)block_scan_run<0>(in2, n, 1u, true, out2, red);
}

// A thread's 16 slots: live mask, heads after the thread's first live key,
// the index and key of that first live key, and the last live key.
struct PackThread {
    uint32_t live, heads, first_i;
    LastKey first, last;
};

__device__ __forceinline__ PackThread pack_thread(const uint32_t f[16], const uint32_t k[16], uint64_t base,
                                                  uint64_t n) {
    PackThread t{0u, 0u, 0u, {0u, 0u}, {0u, 0u}};
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (base + i < n && f[i] == 0u) {
            if (t.last.valid) {
                if (k[i] != t.last.key) t.heads |= 1u << i;
            } else {
                t.first.valid = 1u; t.first.key = k[i]; t.first_i = (uint32_t)i;
            }
            t.last.valid = 1u; t.last.key = k[i];
            t.live |= 1u << i;
        }
    }
    return t;
}

// meta[tile] = (live count, heads after the tile's first live key, first live key, last live key)
__global__ void __launch_bounds__(256) k_pack_tiles(const uint8_t* flags, const uint32_t* keys, uint64_t n,
                                                    uint4* meta) {
    __shared__ LastKey lk[WAVES];
    __shared__ uint32_t red[WAVES];
    __shared__ uint32_t s_first;
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t f[16], k[16];
    load16_u8(flags, base, n, f);
    load16_live(keys, base, n, f, k);
    const PackThread t = pack_thread(f, k, base, n);
    LastKey blk;
    const LastKey carry = block_last_key(t.last, lk, &blk);
    if (t.first.valid && !carry.valid) s_first = t.first.key;     // the tile's first live key
    const uint32_t h = (uint32_t)__popc(t.heads) + (t.first.valid && carry.valid && t.first.key != carry.key);
    uint32_t total;
    (void)block_excl_sum((h << 16) | (uint32_t)__popc(t.live), red, &total);
    if (threadIdx.x == 0) meta[blockIdx.x] = make_uint4(total & 0xFFFFu, total >> 16, blk.valid ? s_first : 0u, blk.key);
}

// Single block: meta[tile] becomes (live offset, chunk offset, carried key,
// carried key valid); totals[0] = N live rays, totals[1] = C chunks,
// bases[C] = N.
__global__ void __launch_bounds__(1024) k_pack_scan(uint4* meta, uint32_t ntiles, uint32_t* totals, uint32_t* bases,
                                                     uint32_t* host, uint32_t seq) {
    __shared__ LastKey lk[16];
    __shared__ uint32_t red[16];
    constexpr uint32_t PER = MAX_TILES / 1024;      // tiles per thread, held in registers
    const uint32_t b = threadIdx.x * PER;
    uint4 m[PER];
    uint32_t ls = 0;
    LastKey last{0u, 0u};
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
        m[i] = b + i < ntiles ? meta[b + i] : make_uint4(0u, 0u, 0u, 0u);
        ls += m[i].x;
        if (m[i].x) { last.valid = 1u; last.key = m[i].w; }
    }
    uint32_t ntot, ctot;
    const uint32_t lo0 = block_excl_sum(ls, red, &ntot);
    LastKey blk;
    const LastKey c0 = block_last_key(last, lk, &blk);
    uint32_t hs = 0;
    LastKey c = c0;
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
        if (m[i].x) {
            hs += m[i].y + ((!c.valid || m[i].z != c.key) ? 1u : 0u);
            c.valid = 1u; c.key = m[i].w;
        }
    }
    const uint32_t ho0 = block_excl_sum(hs, red, &ctot);
    uint32_t lo = lo0, ho = ho0;
    c = c0;
#pragma unroll
    for (uint32_t i = 0; i < PER; i++) {
        if (b + i < ntiles) meta[b + i] = make_uint4(lo, ho, c.key, c.valid);
        if (m[i].x) {
            ho += m[i].y + ((!c.valid || m[i].z != c.key) ? 1u : 0u);
            c.valid = 1u; c.key = m[i].w;
        }
        lo += m[i].x;
    }
    if (threadIdx.x == 0) {
        totals[0] = ntot; totals[1] = ctot; bases[ctot] = ntot;
        if (host) {
            // the sizes, then the sequence word the host polls for
            host[1] = ntot;
            host[2] = ctot;
            __threadfence_system();
            __hip_atomic_store(host, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

__global__ void __launch_bounds__(256) k_pack_write(const uint8_t* flags, const uint32_t* keys,
                                                    const uint32_t* values, uint64_t n, const uint4* meta,
                                                    uint32_t* tvalues, uint32_t* bases, uint32_t* ckeys,
                                                    uint32_t* cvalues, unsigned long long* clear64,
                                                    unsigned long long miss, uint32_t* clear32) {
    __shared__ uint32_t sa[TILE];
    __shared__ uint32_t sb[TILE];
    __shared__ LastKey lk[WAVES];
    __shared__ uint32_t red[WAVES];
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    const uint4 m = meta[blockIdx.x];
    uint32_t f[16], k[16], v[16];
    load16_u8(flags, base, n, f);
    load16_live(keys, base, n, f, k);
    if (values) {
        load16_live(values, base, n, f, v);
    } else {       // a live slot's value is its index
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = (uint32_t)(base + i);
    }
    // the pass's per-slot results start as misses (closest) or clear (shadow)
    if (clear64) {
        if (base + 16 <= n) {
            ulonglong2* q = reinterpret_cast<ulonglong2*>(clear64 + base);
#pragma unroll
            for (int i = 0; i < 8; i++) q[i] = make_ulonglong2(miss, miss);
        } else {
            for (int i = 0; i < 16; i++) if (base + i < n) clear64[base + i] = miss;
        }
    }
    if (clear32) {
        if (base + 16 <= n) {
            uint4* q = reinterpret_cast<uint4*>(clear32 + base);
#pragma unroll
            for (int i = 0; i < 4; i++) q[i] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            for (int i = 0; i < 16; i++) if (base + i < n) clear32[base + i] = 0u;
        }
    }
    PackThread t = pack_thread(f, k, base, n);
    LastKey blk;
    LastKey carry = block_last_key(t.last, lk, &blk);
    if (!carry.valid && m.w) { carry.valid = 1u; carry.key = m.z; }
    if (t.first.valid && (!carry.valid || t.first.key != carry.key)) t.heads |= 1u << t.first_i;
    const uint32_t nl = (uint32_t)__popc(t.live), nh = (uint32_t)__popc(t.heads);
    uint32_t total;
    const uint32_t ex = block_excl_sum((nh << 16) | nl, red, &total);
    const uint32_t tl = total & 0xFFFFu, th = total >> 16;
    uint32_t lp = ex & 0xFFFFu;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (t.live & (1u << i)) sa[lp++] = v[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tl; i += blockDim.x) tvalues[m.x + i] = sa[i];
    __syncthreads();
    lp = ex & 0xFFFFu;
    uint32_t hp = ex >> 16;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (t.live & (1u << i)) {
            if (t.heads & (1u << i)) { sa[hp] = m.x + lp; sb[hp] = k[i]; hp++; }
            lp++;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < th; i += blockDim.x) {
        bases[m.y + i] = sa[i];
        ckeys[m.y + i] = sb[i];
        cvalues[m.y + i] = m.y + i;
    }
}

// ------------------------------------------- decompression (engine) ------
// Sorted chunk x (value v = sv[x]) expands to the live rays
// [bases[v], bases[v + 1]) in order.  k_unpack_sizes sums each tile's chunk
// sizes; after the tile scan, k_unpack_offsets writes every sorted chunk's
// output offset and source base.  k_unpack_expand then gives each block a
// fixed range of TILE outputs (chunk sizes are very uneven, so output ranges
// per chunk tile would be too): two binary searches over the offsets find the
// chunks that cover the range, their offsets and bases are staged in LDS,
// and each output finds its chunk by a binary search there.
__global__ void __launch_bounds__(256) k_unpack_sizes(const uint32_t* sv, const uint32_t* bases, uint64_t C,
                                                      uint32_t* tile_totals, const PassSizes* ps) {
    __shared__ uint32_t red[WAVES];
    if (ps) {
        C = min(C, (uint64_t)ps->C);
        if (blockIdx.x >= tiles_of((uint32_t)C)) return;
    }
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[16];
    load16(sv, base, C, v, 0u);
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (base + i < C) sum += bases[v[i] + 1] - bases[v[i]];
    uint32_t total;
    (void)block_excl_sum(sum, red, &total);
    if (threadIdx.x == 0) tile_totals[blockIdx.x] = total;
}

// soff[x] = output offset of sorted chunk x, sbase[x] = its first live ray;
// soff[C] = N.
__global__ void __launch_bounds__(256) k_unpack_offsets(const uint32_t* sv, const uint32_t* bases, uint64_t C,
                                                        const uint32_t* tile_offsets, uint32_t* soff,
                                                        uint32_t* sbase, const PassSizes* ps) {
    __shared__ uint32_t red[WAVES];
    if (ps) {
        C = min(C, (uint64_t)ps->C);
        if (blockIdx.x >= tiles_of((uint32_t)C)) return;
    }
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[16], b0[16], sz[16];
    load16(sv, base, C, v, 0u);
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        b0[i] = 0u; sz[i] = 0u;
        if (base + i < C) { b0[i] = bases[v[i]]; sz[i] = bases[v[i] + 1] - b0[i]; }
        sum += sz[i];
    }
    uint32_t total;
    uint32_t o = tile_offsets[blockIdx.x] + block_excl_sum(sum, red, &total);
    if (base + 16 <= C) {
        uint4* qo = reinterpret_cast<uint4*>(soff + base);
        uint4* qb = reinterpret_cast<uint4*>(sbase + base);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 a, c;
            a.x = o; o += sz[4 * i]; a.y = o; o += sz[4 * i + 1]; a.z = o; o += sz[4 * i + 2]; a.w = o; o += sz[4 * i + 3];
            c.x = b0[4 * i]; c.y = b0[4 * i + 1]; c.z = b0[4 * i + 2]; c.w = b0[4 * i + 3];
            qo[i] = a; qb[i] = c;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if (base + i < C) { soff[base + i] = o; sbase[base + i] = b0[i]; }
            o += sz[i];
        }
    }
    if (base < C && base + 16 >= C) soff[C] = o;      // the thread holding the last chunk
}

// Largest x in [lo, hi) with a[x] <= p, given a[lo] <= p.
__device__ __forceinline__ uint32_t upper_index(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t p) {
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= p) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_unpack_expand(const uint32_t* soff, const uint32_t* sbase, uint32_t C,
                                                       uint32_t N, const uint32_t* tvalues, uint32_t* out_values,
                                                       const PassSizes* ps) {
    __shared__ uint32_t s_off[TILE + 1];
    __shared__ uint32_t s_base[TILE];
    __shared__ uint32_t s_x[2];
    if (ps) { C = min(C, ps->C); N = min(N, ps->N); }
    const uint32_t p0 = blockIdx.x * TILE;
    if (p0 >= N) return;                  // device-sized launches cover capacity
    const uint32_t p1 = p0 + TILE < N ? p0 + TILE : N;
    if (threadIdx.x < 2) s_x[threadIdx.x] = upper_index(soff, 0, C, threadIdx.x == 0 ? p0 : p1 - 1);
    __syncthreads();
    const uint32_t x0 = s_x[0], nx = s_x[1] - x0 + 1;     // chunks covering [p0, p1): at most TILE
    for (uint32_t i = threadIdx.x; i < nx; i += blockDim.x) {
        s_off[i] = soff[x0 + i];
        s_base[i] = sbase[x0 + i];
    }
    s_off[nx] = 0xFFFFFFFFu;                               // sentinel for the walk below
    __syncthreads();
    // thread t writes the 16 consecutive outputs q .. q + 15: one search, then
    // a walk (measured faster than lane-consecutive outputs, which need a
    // search each)
    const uint32_t q = p0 + threadIdx.x * SCAN_ITEMS;
    if (q >= p1) return;
    uint32_t x = upper_index(s_off, 0, nx, q);
    uint32_t nxt = s_off[x + 1];
    uint32_t o[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = q + i;
        while (p >= nxt) { x++; nxt = s_off[x + 1]; }
        o[i] = p < p1 ? tvalues[s_base[x] + (p - s_off[x])] : 0u;
    }
    if (q + 16 <= p1) {
        uint4* d = reinterpret_cast<uint4*>(out_values + q);
#pragma unroll
        for (int i = 0; i < 4; i++) d[i] = make_uint4(o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]);
    } else {
        for (int i = 0; i < 16; i++)
            if (q + i < p1) out_values[q + i] = o[i];
    }
}

// --------------------------------------------------------- radix sorting --
// RB-bit digits (RB = 8: 256 buckets, or 9: 512 buckets, which sorts a 27-bit
// key in three passes instead of four); DPT = digits per thread of a
// 256-thread block in the per-digit steps, each thread owning DPT
// consecutive digits so that a block scan keeps digit order.
template <uint32_t RB>
__global__ void __launch_bounds__(256) k_radix_upsweep(const uint32_t* keys, uint64_t n, uint32_t shift,
                                                       uint32_t* hist, uint32_t ntiles, const uint32_t* d_n) {
    constexpr uint32_t NB = 1u << RB, MASK = NB - 1u;
    if (d_n) {                            // device-sized: n and the tile count from memory
        n = min(n, (uint64_t)*d_n);
        ntiles = tiles_of((uint32_t)n);
        if (blockIdx.x >= ntiles) return;
    }
    __shared__ uint32_t cnt[WAVES][NB];
    for (uint32_t i = threadIdx.x; i < WAVES * NB; i += blockDim.x) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t w = wave_id(), lane = lane_id();
    const uint64_t tb = (uint64_t)blockIdx.x * TILE + (uint64_t)w * 1024;
    uint32_t k[16];      // all loads issued before the first use
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {
        const uint64_t p = tb + r * 64 + lane;
        k[r] = p < n ? keys[p] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < 16; r++)
        if (tb + r * 64 + lane < n) atomicAdd(&cnt[w][(k[r] >> shift) & MASK], 1u);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < NB; d += blockDim.x) {
        uint32_t s = 0;
#pragma unroll
        for (uint32_t q = 0; q < WAVES; q++) s += cnt[q][d];
        hist[(size_t)d * ntiles + blockIdx.x] = s;
    }
}

// One block per digit: exclusive scan of that digit's tile counts in place.
__global__ void __launch_bounds__(256) k_radix_rowscan(uint32_t* hist, uint32_t ntiles, uint32_t* digit_totals,
                                                       const uint32_t* d_n) {
    __shared__ uint32_t lds[WAVES];
    if (d_n) ntiles = min(ntiles, tiles_of(*d_n));
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    const uint32_t per = (ntiles + 255) / 256;
    const uint32_t b = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t i = 0; i < per; i++) if (b + i < ntiles) s += row[b + i];
    uint32_t total;
    uint32_t run = block_exclusive_add(s, lds, &total);
    for (uint32_t i = 0; i < per; i++) {
        if (b + i < ntiles) { uint32_t t = row[b + i]; row[b + i] = run; run += t; }
    }
    if (threadIdx.x == 0) digit_totals[blockIdx.x] = total;
}

// ROWSCAN: hist holds the upsweep's raw tile counts and each block sums its
// digit rows itself (for small tile counts, in place of k_radix_rowscan).
template <bool ROWSCAN, uint32_t RB>
__global__ void __launch_bounds__(256) k_radix_downsweep(const uint32_t* keys, const uint32_t* vals, uint64_t n,
                                                         uint32_t shift, const uint32_t* hist, uint32_t ntiles,
                                                         const uint32_t* digit_totals, uint32_t* okeys,
                                                         uint32_t* ovals, const uint32_t* d_n) {
    constexpr uint32_t NB = 1u << RB, MASK = NB - 1u, DPT = NB / 256u;
    if (d_n) {
        n = min(n, (uint64_t)*d_n);
        ntiles = tiles_of((uint32_t)n);
        if (blockIdx.x >= ntiles) return;
    }
    __shared__ uint32_t whist[WAVES][NB];
    __shared__ uint32_t tile_start[NB];
    __shared__ uint32_t gbase[NB];
    __shared__ uint32_t sk[TILE];
    __shared__ uint32_t sv[TILE];
    __shared__ uint32_t lds[WAVES];
    const uint32_t tid = threadIdx.x, w = wave_id(), lane = lane_id();
    for (uint32_t i = tid; i < WAVES * NB; i += blockDim.x) (&whist[0][0])[i] = 0;
    // global base of each digit for this tile
    {
        uint32_t dt[DPT], before[DPT], tsum = 0;
#pragma unroll
        for (uint32_t j = 0; j < DPT; j++) {
            const uint32_t d = tid * DPT + j;
            if (ROWSCAN) {
                const uint32_t* row = hist + (size_t)d * ntiles;
                uint32_t b = 0, t = 0;
                for (uint32_t i = 0; i < ntiles; i++) {
                    const uint32_t c = row[i];
                    b += i < blockIdx.x ? c : 0u;
                    t += c;
                }
                dt[j] = t; before[j] = b;
            } else {
                dt[j] = digit_totals[d];
                before[j] = hist[(size_t)d * ntiles + blockIdx.x];
            }
            tsum += dt[j];
        }
        uint32_t t;
        uint32_t pre = block_exclusive_add(tsum, lds, &t);
#pragma unroll
        for (uint32_t j = 0; j < DPT; j++) {
            gbase[tid * DPT + j] = pre + before[j];
            pre += dt[j];
        }
    }
    __syncthreads();
    const uint64_t tb = (uint64_t)blockIdx.x * TILE + (uint64_t)w * 1024;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t key[16], val[16], rank[16];
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {      // all loads issued before the ranking
        const uint64_t p = tb + r * 64 + lane;
        key[r] = p < n ? keys[p] : 0u;
        val[r] = p < n ? vals[p] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {
        uint64_t p = tb + r * 64 + lane;
        bool valid = p < n;
        uint32_t k = key[r];
        uint32_t d = (k >> shift) & MASK;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < RB; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cntp = (uint32_t)__popcll(peers);
        uint32_t basev = valid ? whist[w][d] : 0u;
        rank[r] = basev + before;
        bool leader = valid && before == 0;
        if (leader) whist[w][d] = basev + cntp;
    }
    __syncthreads();
    // per digit: prefix over waves and tile-local start
    {
        uint32_t run[DPT], tsum = 0;
#pragma unroll
        for (uint32_t j = 0; j < DPT; j++) {
            const uint32_t d = tid * DPT + j;
            uint32_t r = 0;
#pragma unroll
            for (uint32_t q = 0; q < WAVES; q++) { uint32_t c = whist[q][d]; whist[q][d] = r; r += c; }
            run[j] = r;
            tsum += r;
        }
        uint32_t t;
        uint32_t st = block_exclusive_add(tsum, lds, &t);
#pragma unroll
        for (uint32_t j = 0; j < DPT; j++) {
            tile_start[tid * DPT + j] = st;
            st += run[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {
        uint64_t p = tb + r * 64 + lane;
        if (p < n) {
            uint32_t d = (key[r] >> shift) & MASK;
            uint32_t lp = tile_start[d] + whist[w][d] + rank[r];
            sk[lp] = key[r];
            sv[lp] = val[r];
        }
    }
    __syncthreads();
    const uint64_t tile0 = (uint64_t)blockIdx.x * TILE;
    const uint32_t valid_n = (uint32_t)((n - tile0) < TILE ? (n - tile0) : TILE);
    for (uint32_t i = tid; i < valid_n; i += blockDim.x) {
        uint32_t k = sk[i];
        uint32_t d = (k >> shift) & MASK;
        uint32_t g = gbase[d] + (i - tile_start[d]);
        okeys[g] = k;
        ovals[g] = sv[i];
    }
}

// --------------------------------------------------------- decompression --
__global__ void __launch_bounds__(256) k_skeleton(const uint32_t* sizes, const uint32_t* sorted_cvalues, uint64_t C,
                                                  uint32_t* skeleton, const uint32_t* d_c) {
    if (d_c && *d_c < C) C = *d_c;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < C; x += (uint64_t)gridDim.x * blockDim.x)
        skeleton[x] = sizes[sorted_cvalues[x]];
}

// Output position p belongs to the sorted chunk x with scan[x] <= p < scan[x+1]
// (sizes are >= 1, so scan is strictly increasing).
__device__ __forceinline__ uint32_t find_chunk(const uint32_t* scan, uint32_t C, uint32_t p) {
    uint32_t lo = 0, hi = C;   // invariant: scan[lo] <= p, answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (scan[mid] <= p) lo = mid; else hi = mid;
    }
    return lo;
}

// d_c, d_n (optional): the chunk and ray totals in device memory, at most C
// and N.
__global__ void __launch_bounds__(256) k_expand(const uint32_t* bases, const uint32_t* sorted_cvalues,
                                                const uint32_t* scan, uint32_t C, uint32_t N, const uint32_t* tk,
                                                const uint32_t* tv, uint32_t* ok, uint32_t* ov, const uint32_t* d_c,
                                                const uint32_t* d_n) {
    if (d_c) C = min(C, *d_c);
    if (d_n) N = min(N, *d_n);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        uint32_t x = find_chunk(scan, C, p);
        uint32_t idx = bases[sorted_cvalues[x]] + (p - scan[x]);
        ok[p] = tk[idx];
        ov[p] = tv[idx];
    }
}

__global__ void __launch_bounds__(256) k_item_counts(const uint32_t* cand, uint32_t chunk, uint32_t ntop,
                                                     uint32_t* counts) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > ntop) return;
    counts[i] = i == ntop ? 0u : (uint32_t)(((uint64_t)cand[i] + chunk - 1) / chunk);
}

inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 1u << 20) {
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return (uint32_t)(g < cap ? g : cap);
}

}  // namespace

hipError_t ScanScratch::reserve(uint64_t n) {
    hipError_t e;
    if ((e = tile_counts.reserve((MAX_TILES + 1) * sizeof(uint32_t))) != hipSuccess) return e;
    uint32_t nt = tiles_for(n) + 1;
    if ((e = hist.reserve((size_t)512 * nt * sizeof(uint32_t))) != hipSuccess) return e;    // 9-bit digits
    if ((e = digit_totals.reserve(512 * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = total.reserve(4 * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = ping_k.reserve(n * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = ping_v.reserve(n * sizeof(uint32_t))) != hipSuccess) return e;
    return hipSuccess;
}

void ScanScratch::release() {
    tile_counts.release(); hist.release(); digit_totals.release(); total.release();
    ping_k.release(); ping_v.release();
}

void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n) k_fill_u32<<<grid_for(n, 256, 4096), 256, 0, s>>>(p, v, n);
}
void launch_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s) {
    if (n) k_fill_u64<<<grid_for(n, 256, 4096), 256, 0, s>>>(p, v, n);
}

hipError_t run_trim(const uint32_t* flags, const uint32_t* keys, const uint32_t* values, uint64_t n,
                    uint32_t* out_keys, uint32_t* out_values, uint32_t* d_total, ScanScratch& sc,
                    hipStream_t s) {
    uint32_t nt = tiles_for(n);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0) { launch_fill_u32(d_total, 0, 1, s); return hipGetLastError(); }
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_tile_total<0><<<nt, 256, 0, s>>>(flags, n, tiles, nullptr);
    k_scan_tiles<<<1, 1024, 0, s>>>(tiles, nt, d_total, nullptr);
    k_trim_scatter<<<nt, 256, 0, s>>>(flags, keys, values, n, tiles, out_keys, out_values);
    return hipGetLastError();
}

// d_n (optional): the element count in device memory, at most n (launches
// cover n; the total lands in sc.total either way).
static hipError_t run_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s,
                           bool inclusive, const uint32_t* d_n = nullptr) {
    uint32_t nt = tiles_for(n);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0) {
        if (d_n) launch_fill_u32(sc.total.as<uint32_t>(), 0u, 1, s);
        return hipGetLastError();
    }
    if (!inclusive && n <= SMALL_SCAN && !d_n) {     // one launch instead of three
        k_scan_block<0><<<1, 1024, 0, s>>>(in, (uint32_t)n, 1u, true, out, sc.total.as<uint32_t>(), nullptr, 0u,
                                            nullptr, nullptr, nullptr);
        return hipGetLastError();
    }
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_tile_total<1><<<nt, 256, 0, s>>>(in, n, tiles, d_n);
    k_scan_tiles<<<1, 1024, 0, s>>>(tiles, nt, sc.total.as<uint32_t>(), d_n);
    if (inclusive) k_scan_write<0><<<nt, 256, 0, s>>>(in, n, tiles, out, d_n);
    else k_scan_write<1><<<nt, 256, 0, s>>>(in, n, tiles, out, d_n);
    return hipGetLastError();
}

hipError_t run_inclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s) {
    return run_scan(in, out, n, sc, s, true);
}
hipError_t run_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s) {
    return run_scan(in, out, n, sc, s, false);
}

hipError_t run_item_scan(const uint32_t* cand, uint32_t ntop, uint32_t chunk, uint32_t max_cand, uint32_t* counts,
                         uint32_t* first, ScanScratch& sc, hipStream_t s, unsigned long long* zero,
                         uint32_t zero_n, const uint32_t* npass, uint32_t* off, const PassSizes* ps) {
    if (chunk == 0 || zero_n > 1024) return hipErrorInvalidValue;
    if ((uint64_t)ntop + 1 <= SMALL_SCAN) {
        const bool fast = div_hi_ok((uint64_t)max_cand + chunk - 1, chunk);
        k_scan_block<1><<<1, ITEM_SCAN_THREADS, 0, s>>>(cand, ntop + 1, chunk, fast, first, nullptr, zero, zero_n,
                                                        npass, off, ps);
        return hipGetLastError();
    }
    if (ps) return hipErrorInvalidValue;  // device-sized item scans are single-block (ntop < SMALL_SCAN)
    if (zero) launch_fill_u64(zero, 0ull, zero_n, s);
    if (npass) {
        hipError_t e = run_exclusive_scan(npass, off, (uint64_t)ntop + 1, sc, s);
        if (e != hipSuccess) return e;
    }
    k_item_counts<<<ntop / 256 + 1, 256, 0, s>>>(cand, chunk, ntop, counts);
    return run_scan(counts, first, (uint64_t)ntop + 1, sc, s, false);
}

hipError_t run_compress(const uint32_t* keys, uint64_t n, uint32_t* bases, uint32_t* sizes, uint32_t* ckeys,
                        uint32_t* cvalues, uint32_t* d_total, ScanScratch& sc, hipStream_t s, const uint32_t* d_n) {
    uint32_t nt = tiles_for(n);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0) { launch_fill_u32(d_total, 0, 1, s); return hipGetLastError(); }
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_tile_total<2><<<nt, 256, 0, s>>>(keys, n, tiles, d_n);
    k_scan_tiles<<<1, 1024, 0, s>>>(tiles, nt, d_total, d_n);
    k_compress_scatter<<<nt, 256, 0, s>>>(keys, n, tiles, bases, ckeys, cvalues, d_n);
    k_chunk_sizes<<<grid_for(n, 256, 8192), 256, 0, s>>>(bases, d_total, (uint32_t)n, sizes, d_n);
    return hipGetLastError();
}

hipError_t run_pack(const uint8_t* flags, const uint32_t* keys, const uint32_t* values, uint64_t n,
                    uint32_t* tvalues, uint32_t* bases, uint32_t* ckeys, uint32_t* cvalues, uint32_t* d_totals,
                    unsigned long long* clear64, unsigned long long miss, uint32_t* clear32, ScanScratch& sc,
                    hipStream_t s, uint32_t* host, uint32_t seq) {
    const uint32_t nt = tiles_for(n);
    if (nt > MAX_TILES || (nt == 0 && host)) return hipErrorInvalidValue;
    if (nt == 0) {
        launch_fill_u32(d_totals, 0, 2, s);
        launch_fill_u32(bases, 0, 1, s);
        return hipGetLastError();
    }
    uint4* meta = sc.hist.as<uint4>();
    k_pack_tiles<<<nt, 256, 0, s>>>(flags, keys, n, meta);
    k_pack_scan<<<1, 1024, 0, s>>>(meta, nt, d_totals, bases, host, seq);
    k_pack_write<<<nt, 256, 0, s>>>(flags, keys, values, n, meta, tvalues, bases, ckeys, cvalues, clear64, miss,
                                    clear32);
    return hipGetLastError();
}

hipError_t run_unpack(const uint32_t* bases, const uint32_t* sorted_cvalues, const uint32_t* tvalues, uint64_t C,
                      uint64_t N, uint32_t* soff, uint32_t* sbase, uint32_t* out_values, ScanScratch& sc,
                      hipStream_t s) {
    const uint32_t nt = tiles_for(C);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0 || N == 0) return hipSuccess;
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_unpack_sizes<<<nt, 256, 0, s>>>(sorted_cvalues, bases, C, tiles, nullptr);
    k_unpack_scan<<<1, 1024, 0, s>>>(tiles, nt, sc.total.as<uint32_t>(), nullptr);
    k_unpack_offsets<<<nt, 256, 0, s>>>(sorted_cvalues, bases, C, tiles, soff, sbase, nullptr);
    k_unpack_expand<<<tiles_for(N), 256, 0, s>>>(soff, sbase, (uint32_t)C, (uint32_t)N, tvalues, out_values, nullptr);
    return hipGetLastError();
}

hipError_t run_unpack_dev(const uint32_t* bases, const uint32_t* sorted_cvalues, const uint32_t* tvalues,
                          uint64_t cap, const PassSizes* ps, uint32_t* soff, uint32_t* sbase, uint32_t* out_values,
                          ScanScratch& sc, hipStream_t s) {
    const uint32_t nt = tiles_for(cap);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0) return hipSuccess;
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_unpack_sizes<<<nt, 256, 0, s>>>(sorted_cvalues, bases, cap, tiles, ps);
    k_unpack_scan<<<1, 1024, 0, s>>>(tiles, nt, sc.total.as<uint32_t>(), &ps->C);
    k_unpack_offsets<<<nt, 256, 0, s>>>(sorted_cvalues, bases, cap, tiles, soff, sbase, ps);
    k_unpack_expand<<<nt, 256, 0, s>>>(soff, sbase, (uint32_t)cap, (uint32_t)cap, tvalues, out_values, ps);
    return hipGetLastError();
}

// Hierarchy levels over q.N rays (as level_sizes on the host).
__device__ void fill_levels(PassSizes& q, uint32_t leaf, uint32_t branch, uint32_t depth) {
    uint32_t n = (q.N + leaf - 1) / leaf, off = 0;
    for (uint32_t k = 0; k < 4; k++) {
        q.count[k] = k < depth ? n : 0u;
        q.offset[k] = off;
        if (k < depth) { off += n; n = (n + branch - 1) / branch; }
    }
    q.ntop = q.count[depth - 1];
}

__global__ void k_pass_sizes(const uint32_t* totals, uint32_t leaf, uint32_t branch, uint32_t depth, PassSizes* ps,
                             uint32_t* rec) {
    PassSizes q{};
    q.N = totals[0];
    q.C = totals[1];
    fill_levels(q, leaf, branch, depth);
    *ps = q;
    if (rec) { rec[0] = q.N; rec[1] = q.C; }
}

__global__ void k_fixed_sizes(uint32_t N, uint32_t C, uint32_t leaf, uint32_t branch, uint32_t depth, PassSizes* ps,
                              uint32_t* rec) {
    PassSizes q{};
    q.N = N;
    q.C = C;
    fill_levels(q, leaf, branch, depth);
    *ps = q;
    if (rec) { rec[0] = N; rec[1] = C; }
}

__global__ void k_ray_sizes(const uint32_t* d_n, uint32_t n_cap, uint32_t leaf, uint32_t branch, uint32_t depth,
                            PassSizes* ps) {
    PassSizes q{};
    q.N = min(*d_n, n_cap);
    fill_levels(q, leaf, branch, depth);
    *ps = q;
}

void launch_pass_sizes(const uint32_t* totals, uint32_t leaf, uint32_t branch, uint32_t depth, PassSizes* ps,
                       uint32_t* rec, hipStream_t s) {
    k_pass_sizes<<<1, 1, 0, s>>>(totals, leaf, branch, depth, ps, rec);
}

void launch_fixed_sizes(uint32_t N, uint32_t C, uint32_t leaf, uint32_t branch, uint32_t depth, PassSizes* ps,
                        uint32_t* rec, hipStream_t s) {
    k_fixed_sizes<<<1, 1, 0, s>>>(N, C, leaf, branch, depth, ps, rec);
}

void launch_ray_sizes(const uint32_t* d_n, uint32_t n_cap, uint32_t leaf, uint32_t branch, uint32_t depth,
                      PassSizes* ps, hipStream_t s) {
    k_ray_sizes<<<1, 1, 0, s>>>(d_n, n_cap, leaf, branch, depth, ps);
}

// Up to this many tiles the downsweep sums the digit rows itself (saves a launch
// per digit; each block reads 2^RB x tiles counts).  Measured on the 1080p
// atrium frame with 9-bit digits (sort stage per frame): 160 tiles 0.54 ms,
// 64 0.46, 24 0.40, 8 0.38, 0 (always the separate row scan) 0.39.
constexpr uint32_t ROWSCAN_IN_DOWNSWEEP = 8;

// d_n: device-sized (n and nt are capacities; the kernels read the count)
template <uint32_t RB>
void radix_pass(const uint32_t* ik, const uint32_t* iv, uint64_t n, uint32_t shift, uint32_t nt, uint32_t* ok,
                uint32_t* ov, ScanScratch& sc, hipStream_t s, const uint32_t* d_n = nullptr) {
    k_radix_upsweep<RB><<<nt, 256, 0, s>>>(ik, n, shift, sc.hist.as<uint32_t>(), nt, d_n);
    if (nt <= ROWSCAN_IN_DOWNSWEEP && !d_n) {
        k_radix_downsweep<true, RB><<<nt, 256, 0, s>>>(ik, iv, n, shift, sc.hist.as<uint32_t>(), nt, nullptr, ok, ov,
                                                       nullptr);
    } else {
        k_radix_rowscan<<<1u << RB, 256, 0, s>>>(sc.hist.as<uint32_t>(), nt, sc.digit_totals.as<uint32_t>(), d_n);
        k_radix_downsweep<false, RB><<<nt, 256, 0, s>>>(ik, iv, n, shift, sc.hist.as<uint32_t>(), nt,
                                                        sc.digit_totals.as<uint32_t>(), ok, ov, d_n);
    }
}

// LSD radix sort of the low `bits` key bits: 8-bit digits, or 9-bit digits
// where they need fewer passes (17-18 and 25-27 bits).
hipError_t run_radix_sort(const uint32_t* keys, const uint32_t* values, uint64_t n, uint32_t bits,
                          uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint32_t nt = tiles_for(n);
    if (bits == 0) bits = 1;
    if (bits > 32) bits = 32;
    const uint32_t rb = (bits + 8) / 9 < (bits + 7) / 8 ? 9u : 8u;
    const uint32_t passes = (bits + rb - 1) / rb;
    // Ping-pong so that the final pass lands in out_*.
    uint32_t* bk[2] = {sc.ping_k.as<uint32_t>(), out_keys};
    uint32_t* bv[2] = {sc.ping_v.as<uint32_t>(), out_values};
    const uint32_t* ik = keys;
    const uint32_t* iv = values;
    // In place (out == in) with an odd pass count, the first pass would scatter
    // into the array its other blocks still read: start from a copy in the
    // ping buffers, which the first pass does not write.
    if ((passes & 1u) && (keys == out_keys || values == out_values || keys == out_values || values == out_keys)) {
        hipError_t e = hipMemcpyAsync(bk[0], keys, n * 4, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(bv[0], values, n * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        ik = bk[0];
        iv = bv[0];
    }
    for (uint32_t pass = 0; pass < passes; pass++) {
        uint32_t dst = ((passes - 1 - pass) & 1u) ? 0u : 1u;
        if (rb == 9) radix_pass<9>(ik, iv, n, 9 * pass, nt, bk[dst], bv[dst], sc, s);
        else radix_pass<8>(ik, iv, n, 8 * pass, nt, bk[dst], bv[dst], sc, s);
        ik = bk[dst];
        iv = bv[dst];
    }
    return hipGetLastError();
}

hipError_t run_radix_sort_dev(const uint32_t* keys, const uint32_t* values, uint64_t n_cap, const uint32_t* d_n,
                              uint32_t bits, uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc,
                              hipStream_t s) {
    if (n_cap == 0) return hipSuccess;
    const uint32_t nt = tiles_for(n_cap);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (bits == 0) bits = 1;
    if (bits > 32) bits = 32;
    const uint32_t rb = (bits + 8) / 9 < (bits + 7) / 8 ? 9u : 8u;
    const uint32_t passes = (bits + rb - 1) / rb;
    if (keys == out_keys || values == out_values || keys == out_values || values == out_keys)
        return hipErrorInvalidValue;
    uint32_t* bk[2] = {sc.ping_k.as<uint32_t>(), out_keys};
    uint32_t* bv[2] = {sc.ping_v.as<uint32_t>(), out_values};
    const uint32_t* ik = keys;
    const uint32_t* iv = values;
    for (uint32_t pass = 0; pass < passes; pass++) {
        uint32_t dst = ((passes - 1 - pass) & 1u) ? 0u : 1u;
        if (rb == 9) radix_pass<9>(ik, iv, n_cap, 9 * pass, nt, bk[dst], bv[dst], sc, s, d_n);
        else radix_pass<8>(ik, iv, n_cap, 8 * pass, nt, bk[dst], bv[dst], sc, s, d_n);
        ik = bk[dst];
        iv = bv[dst];
    }
    return hipGetLastError();
}

hipError_t run_decompress(const uint32_t* bases, const uint32_t* sizes, const uint32_t* sorted_cvalues,
                          const uint32_t* tkeys, const uint32_t* tvalues, uint32_t* skeleton, uint32_t* scan,
                          uint64_t C, uint64_t N, uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc,
                          hipStream_t s, const uint32_t* d_c) {
    if (C == 0) return hipSuccess;
    if (C > 0xFFFFFFFFull) return hipErrorInvalidValue;
    k_skeleton<<<grid_for(C, 256, 8192), 256, 0, s>>>(sizes, sorted_cvalues, C, skeleton, d_c);
    hipError_t e = run_scan(skeleton, scan, C, sc, s, false, d_c);
    if (e != hipSuccess) return e;
    // N == 0: the ray total is the scan's, read on the device; the grid
    // strides over it
    const uint32_t* d_n = N == 0 ? sc.total.as<uint32_t>() : nullptr;
    k_expand<<<N == 0 ? 2048u : grid_for(N, 256, 16384), 256, 0, s>>>(
        bases, sorted_cvalues, scan, (uint32_t)C, N == 0 ? 0xFFFFFFFFu : (uint32_t)N, tkeys, tvalues, out_keys,
        out_values, d_c, d_n);
    return hipGetLastError();
}

}  // namespace rsh
