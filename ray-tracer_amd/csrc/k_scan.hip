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
template <int MODE>
__global__ void __launch_bounds__(256) k_tile_total(const uint32_t* in, uint64_t n, uint32_t* tile_totals) {
    __shared__ uint32_t lds[WAVES];
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
__global__ void __launch_bounds__(1024) k_scan_tiles(uint32_t* tiles, uint32_t ntiles, uint32_t* total) {
    __shared__ uint32_t lds[16];
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
                                                    uint32_t* out) {
    __shared__ uint32_t lds[WAVES];
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
                                                          uint32_t* ckeys, uint32_t* cvalues) {
    __shared__ uint32_t lds[WAVES];
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
                                                     uint32_t* sizes) {
    const uint32_t C = *d_chunks;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < C; x += gridDim.x * blockDim.x)
        sizes[x] = (x + 1 < C ? bases[x + 1] : n) - bases[x];
}

// --------------------------------------------------------- radix sorting --
__global__ void __launch_bounds__(256) k_radix_upsweep(const uint32_t* keys, uint64_t n, uint32_t shift,
                                                       uint32_t* hist, uint32_t ntiles) {
    __shared__ uint32_t cnt[WAVES][256];
    for (uint32_t i = threadIdx.x; i < WAVES * 256; i += blockDim.x) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t w = wave_id(), lane = lane_id();
    const uint64_t tb = (uint64_t)blockIdx.x * TILE + (uint64_t)w * 1024;
#pragma unroll 4
    for (uint32_t r = 0; r < 16; r++) {
        uint64_t p = tb + r * 64 + lane;
        if (p < n) atomicAdd(&cnt[w][(keys[p] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    const uint32_t d = threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < WAVES; k++) s += cnt[k][d];
    hist[(size_t)d * ntiles + blockIdx.x] = s;
}

// One block per digit: exclusive scan of that digit's tile counts in place.
__global__ void __launch_bounds__(256) k_radix_rowscan(uint32_t* hist, uint32_t ntiles, uint32_t* digit_totals) {
    __shared__ uint32_t lds[WAVES];
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

__global__ void __launch_bounds__(256) k_radix_downsweep(const uint32_t* keys, const uint32_t* vals, uint64_t n,
                                                         uint32_t shift, const uint32_t* hist, uint32_t ntiles,
                                                         const uint32_t* digit_totals, uint32_t* okeys,
                                                         uint32_t* ovals) {
    __shared__ uint32_t whist[WAVES][256];
    __shared__ uint32_t tile_start[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t sk[TILE];
    __shared__ uint32_t sv[TILE];
    __shared__ uint32_t lds[WAVES];
    const uint32_t tid = threadIdx.x, w = wave_id(), lane = lane_id();
    for (uint32_t i = tid; i < WAVES * 256; i += blockDim.x) (&whist[0][0])[i] = 0;
    // global base of each digit for this tile
    {
        uint32_t t;
        uint32_t dt = digit_totals[tid];
        uint32_t pre = block_exclusive_add(dt, lds, &t);
        gbase[tid] = pre + hist[(size_t)tid * ntiles + blockIdx.x];
    }
    __syncthreads();
    const uint64_t tb = (uint64_t)blockIdx.x * TILE + (uint64_t)w * 1024;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t key[16], val[16], rank[16];
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {
        uint64_t p = tb + r * 64 + lane;
        bool valid = p < n;
        uint32_t k = valid ? keys[p] : 0u;
        key[r] = k;
        val[r] = valid ? vals[p] : 0u;
        uint32_t d = (k >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; b++) {
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
        const uint32_t d = tid;
        uint32_t run = 0;
#pragma unroll
        for (uint32_t k = 0; k < WAVES; k++) { uint32_t c = whist[k][d]; whist[k][d] = run; run += c; }
        uint32_t t;
        uint32_t st = block_exclusive_add(run, lds, &t);
        tile_start[d] = st;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < 16; r++) {
        uint64_t p = tb + r * 64 + lane;
        if (p < n) {
            uint32_t d = (key[r] >> shift) & 0xFFu;
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
        uint32_t d = (k >> shift) & 0xFFu;
        uint32_t g = gbase[d] + (i - tile_start[d]);
        okeys[g] = k;
        ovals[g] = sv[i];
    }
}

// --------------------------------------------------------- decompression --
__global__ void __launch_bounds__(256) k_skeleton(const uint32_t* sizes, const uint32_t* sorted_cvalues, uint64_t C,
                                                  uint32_t* skeleton) {
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

__global__ void __launch_bounds__(256) k_expand(const uint32_t* bases, const uint32_t* sorted_cvalues,
                                                const uint32_t* scan, uint32_t C, uint32_t N, const uint32_t* tk,
                                                const uint32_t* tv, uint32_t* ok, uint32_t* ov) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        uint32_t x = find_chunk(scan, C, p);
        uint32_t idx = bases[sorted_cvalues[x]] + (p - scan[x]);
        ok[p] = tk[idx];
        ov[p] = tv[idx];
    }
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
    if ((e = hist.reserve((size_t)256 * nt * sizeof(uint32_t))) != hipSuccess) return e;
    if ((e = digit_totals.reserve(256 * sizeof(uint32_t))) != hipSuccess) return e;
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
    k_tile_total<0><<<nt, 256, 0, s>>>(flags, n, tiles);
    k_scan_tiles<<<1, 1024, 0, s>>>(tiles, nt, d_total);
    k_trim_scatter<<<nt, 256, 0, s>>>(flags, keys, values, n, tiles, out_keys, out_values);
    return hipGetLastError();
}

static hipError_t run_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s,
                           bool inclusive) {
    uint32_t nt = tiles_for(n);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0) return hipSuccess;
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_tile_total<1><<<nt, 256, 0, s>>>(in, n, tiles);
    k_scan_tiles<<<1, 1024, 0, s>>>(tiles, nt, sc.total.as<uint32_t>());
    if (inclusive) k_scan_write<0><<<nt, 256, 0, s>>>(in, n, tiles, out);
    else k_scan_write<1><<<nt, 256, 0, s>>>(in, n, tiles, out);
    return hipGetLastError();
}

hipError_t run_inclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s) {
    return run_scan(in, out, n, sc, s, true);
}
hipError_t run_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanScratch& sc, hipStream_t s) {
    return run_scan(in, out, n, sc, s, false);
}

hipError_t run_compress(const uint32_t* keys, uint64_t n, uint32_t* bases, uint32_t* sizes, uint32_t* ckeys,
                        uint32_t* cvalues, uint32_t* d_total, ScanScratch& sc, hipStream_t s) {
    uint32_t nt = tiles_for(n);
    if (nt > MAX_TILES) return hipErrorInvalidValue;
    if (nt == 0) { launch_fill_u32(d_total, 0, 1, s); return hipGetLastError(); }
    uint32_t* tiles = sc.tile_counts.as<uint32_t>();
    k_tile_total<2><<<nt, 256, 0, s>>>(keys, n, tiles);
    k_scan_tiles<<<1, 1024, 0, s>>>(tiles, nt, d_total);
    k_compress_scatter<<<nt, 256, 0, s>>>(keys, n, tiles, bases, ckeys, cvalues);
    k_chunk_sizes<<<grid_for(n, 256, 8192), 256, 0, s>>>(bases, d_total, (uint32_t)n, sizes);
    return hipGetLastError();
}

hipError_t run_radix_sort(const uint32_t* keys, const uint32_t* values, uint64_t n, uint32_t bits,
                          uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint32_t nt = tiles_for(n);
    if (bits == 0) bits = 1;
    uint32_t passes = (bits + 7) / 8;
    // Ping-pong so that the final pass lands in out_*.
    uint32_t* bk[2] = {sc.ping_k.as<uint32_t>(), out_keys};
    uint32_t* bv[2] = {sc.ping_v.as<uint32_t>(), out_values};
    const uint32_t* ik = keys;
    const uint32_t* iv = values;
    for (uint32_t pass = 0; pass < passes; pass++) {
        uint32_t dst = ((passes - 1 - pass) & 1u) ? 0u : 1u;
        uint32_t shift = 8 * pass;
        k_radix_upsweep<<<nt, 256, 0, s>>>(ik, n, shift, sc.hist.as<uint32_t>(), nt);
        k_radix_rowscan<<<256, 256, 0, s>>>(sc.hist.as<uint32_t>(), nt, sc.digit_totals.as<uint32_t>());
        k_radix_downsweep<<<nt, 256, 0, s>>>(ik, iv, n, shift, sc.hist.as<uint32_t>(), nt,
                                             sc.digit_totals.as<uint32_t>(), bk[dst], bv[dst]);
        ik = bk[dst];
        iv = bv[dst];
    }
    return hipGetLastError();
}

hipError_t run_decompress(const uint32_t* bases, const uint32_t* sizes, const uint32_t* sorted_cvalues,
                          const uint32_t* tkeys, const uint32_t* tvalues, uint32_t* skeleton, uint32_t* scan,
                          uint64_t C, uint64_t N, uint32_t* out_keys, uint32_t* out_values, ScanScratch& sc,
                          hipStream_t s) {
    if (C == 0) return hipSuccess;
    k_skeleton<<<grid_for(C, 256, 8192), 256, 0, s>>>(sizes, sorted_cvalues, C, skeleton);
    hipError_t e = run_exclusive_scan(skeleton, scan, C, sc, s);
    if (e != hipSuccess) return e;
    if (N == 0) {   // ray total not known by the caller: read the scan total
        uint32_t t = 0;
        if ((e = hipMemcpyAsync(&t, sc.total.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        N = t;
        if (N == 0) return hipSuccess;
    }
    k_expand<<<grid_for(N, 256, 16384), 256, 0, s>>>(bases, sorted_cvalues, scan, (uint32_t)C, (uint32_t)N, tkeys,
                                                     tvalues, out_keys, out_values);
    return hipGetLastError();
}

}  // namespace rsh
