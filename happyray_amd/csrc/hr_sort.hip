// hr_sort.hip -- device-wide exclusive scan and stable LSD radix sort of
// (key, value) pairs for gfx950.
//
// These replace the reference's cub calls: ExclusiveScan
// (src/Utils/Scan.cu:39-58, cub::DeviceScan::ExclusiveSum) and RadixSort
// (src/Utils/Sort.cu:31-52, cub::DeviceRadixSort::SortPairs over all 32 key
// bits).  The sort here only visits the key bits that can be non-zero
// (ceil(log2(#cells)) instead of 32) and keeps stability, which together with
// pairs emitted in primitive order gives the canonical (cell, primitive)
// order.
//
// Sizes may live in device memory (n_dev) so that a whole build can run
// without host round trips; grids are sized from the capacity and tiles past
// n return at once.  Scans and sort passes are single launches with decoupled
// look-back: workgroups take tile ids in launch order, publish their tile's
// sum, and add up the published sums of earlier tiles.  A wait is bounded:
// after kSpinLimit polls the waiting workgroup recomputes the earlier tile's
// sum from the input itself, so no launch depends on co-residency.
#include "hr_device.h"

namespace hr {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds_waves,
                                                    uint32_t& total)
{
    const int lane = (int)__lane_id();
    const int w = (int)threadIdx.x / kWave;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == kWave - 1) lds_waves[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
    for (int i = 0; i < (int)blockDim.x / kWave; ++i) {
        const uint32_t c = lds_waves[i];
        if (i < w) off += c;
        total += c;
    }
    __syncthreads();
    return off + inc - v;
}

// ---------------------------------------------------------------------------
// single-pass scan (decoupled look-back)
// ---------------------------------------------------------------------------

// chain words: [0] tile counter, [1] epoch, [2..7] unused, then one 64-bit
// flag per tile: epoch << 34 | status << 32 | value, status 1 = the tile's
// sum, 2 = its inclusive prefix (0 = nothing published in this epoch)
constexpr uint32_t kChainHead = 8;
constexpr uint64_t kAggregate = 1, kPrefix = 2;
constexpr int kSpinLimit = 1 << 16;

__device__ __forceinline__ uint64_t flag_pack(uint32_t epoch, uint64_t status, uint32_t v)
{
    return ((uint64_t)(epoch & 0x3FFFFFFFu) << 34) | (status << 32) | v;
}

// the tile's sum recomputed from the input (look-back fallback)
__device__ uint32_t tile_sum(const uint32_t* in, uint32_t tile, uint32_t n)
{
    uint32_t s = 0;
    const uint32_t b = tile * kScanTile;
    const uint32_t e = b + kScanTile < n ? b + kScanTile : n;
    for (uint32_t i = b; i < e; ++i) s += in[i];
    return s;
}

__global__ __launch_bounds__(kBlock) void k_scan_1p(const uint32_t* __restrict__ in, uint32_t* out,
                                                    uint32_t out_stride, uint32_t n_cap,
                                                    const uint32_t* __restrict__ n_dev, uint32_t* total_dev,
                                                    ScanCheck check, uint32_t* chain)
{
    __shared__ uint32_t lds[kBlock / kWave];
    __shared__ uint32_t s_tile, s_epoch, s_excl;
    unsigned long long* flags = (unsigned long long*)(chain + kChainHead);
    if (threadIdx.x == 0) {
        // acquire: the epoch is read before the id is taken, so the workgroup
        // that takes the last id (and moves the epoch on) comes after it
        s_epoch = __hip_atomic_load(&chain[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        s_tile = atomicAdd(&chain[0], 1u);
    }
    __syncthreads();
    const uint32_t tile = s_tile, epoch = s_epoch;
    const uint32_t n = n_dev ? *n_dev : n_cap;
    const uint32_t base = tile * kScanTile;
    if (base < n) {
        constexpr int kItems = kScanTile / kBlock;
        const uint32_t first = base + threadIdx.x * kItems;
        uint32_t v[kItems];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint32_t i = first + k;
            v[k] = i < n ? in[i] : 0u;
            s += v[k];
        }
        uint32_t agg;
        const uint32_t ex = block_excl_scan(s, lds, agg);
        if (threadIdx.x < kWave) {
            // decoupled look-back by the first wave: lane i reads the flag of
            // tile (p_hi - i), so 64 predecessors are examined per round
            const int lane = (int)__lane_id();
            const uint32_t ep = epoch & 0x3FFFFFFFu;
            uint32_t excl = 0;
            if (tile == 0) {
                if (lane == 0)
                    __hip_atomic_store(&flags[0], flag_pack(epoch, kPrefix, agg), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (lane == 0)
                    __hip_atomic_store(&flags[tile], flag_pack(epoch, kAggregate, agg), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                int p_hi = (int)tile - 1;
                int spins = 0;
                while (true) {
                    const int p = p_hi - lane;
                    uint64_t st = 0;
                    uint32_t val = 0;
                    if (p >= 0) {
                        // the sum travels inside the flag word: relaxed atomics suffice
                        const uint64_t f = __hip_atomic_load(&flags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if ((uint32_t)(f >> 34) == ep) {
                            st = (f >> 32) & 3u;
                            val = (uint32_t)f;
                        }
                    }
                    if (p >= 0 && st == 0 && spins > kSpinLimit) {
                        // bounded wait: recompute the missing tile's sum from the input
                        st = kAggregate;
                        val = tile_sum(in, (uint32_t)p, n);
                    }
                    const uint64_t pre = __ballot(p >= 0 && st == kPrefix);
                    const uint64_t ready = __ballot(p < 0 || st != 0);
                    // lanes up to the nearest prefix (or the whole window) must have published
                    const int last = pre ? (int)__builtin_ctzll(pre) : kWave - 1;
                    const uint64_t need = last == kWave - 1 ? ~0ull : ((1ull << (last + 1)) - 1ull);
                    if ((ready & need) == need) {
                        uint32_t v = (lane <= last && p >= 0) ? val : 0u;
#pragma unroll
                        for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
                        excl += v;
                        if (pre || p_hi - (kWave - 1) <= 0) break;
                        p_hi -= kWave;
                        spins = 0;
                        continue;
                    }
                    ++spins;
                    __builtin_amdgcn_s_sleep(1);
                }
                if (lane == 0)
                    __hip_atomic_store(&flags[tile], flag_pack(epoch, kPrefix, excl + agg), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            if (lane == 0) {
                s_excl = excl;
                // the tile holding the last element knows the total
                if ((n - 1) / kScanTile == tile) {
                    const uint32_t total = excl + agg;
                    if (total_dev) *total_dev = total;
                    if (check.st) {
                        if (total > check.cap) atomicOr(&check.st->overflow, check.bit);
                        *check.eff = check.st->overflow ? 0u : total;
                    }
                }
            }
        }
        __syncthreads();
        uint32_t run = s_excl + ex;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint32_t i = first + k;
            if (i < n) out[(size_t)i * out_stride] = run;
            run += v[k];
        }
    } else if (tile == 0 && threadIdx.x == 0) {
        // nothing to scan
        if (total_dev) *total_dev = 0u;
        if (check.st) *check.eff = 0u;
    }
    if (tile == gridDim.x - 1 && threadIdx.x == 0) {
        // every workgroup has taken its id: ready the chain for the next scan
        __hip_atomic_store(&chain[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&chain[1], epoch + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_fill(uint32_t* __restrict__ p, uint32_t value, size_t n_cap,
                       const uint32_t* __restrict__ n_dev, uint32_t n_mul)
{
    const size_t n = n_dev ? (size_t)(*n_dev) * n_mul : n_cap;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------

constexpr int kRadix = 256;
constexpr int kSortItems = kSortTile / kBlock;   // 8
constexpr int kSortWaves = kBlock / kWave;       // 4

// Chunked LSD radix sort.  A pass runs over a fixed grid of G workgroups;
// workgroup b owns a contiguous chunk of whole 2048-key tiles.
//  1. k_sort_count: each workgroup counts the digits of its chunk (one LDS
//     histogram per wave; counts do not depend on the order of the adds)
//     into counts[digit * G + b];
//  2. k_sort_rowscan: one workgroup per digit scans that digit's row (each
//     chunk's first slot among the digit's keys) and writes the digit total;
//  3. k_sort_scatter: each workgroup adds the totals of the smaller digits to
//     its row entries, then walks its tiles in order, ranks the keys
//     of a tile by digit (stable), stages the tile in LDS by digit and writes
//     each digit's run at the chunk's running offset.
// Every chunk's tiles keep their order, so the sort is stable.

constexpr int kSortGrid = 1024;                  // workgroups per pass (at most)

// rank of each of the wave's keys among the equal digits of the wave (lanes
// in order), and the wave's counts per digit accumulated into h[digit]
__device__ __forceinline__ uint32_t wave_digit_rank(uint32_t d, bool valid, uint32_t* h)
{
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        match &= bit ? bal : ~bal;
    }
    const uint32_t below = (uint32_t)__popcll(match & __lanemask_lt());
    const uint32_t cnt = h[d];
    wave_sync();
    if (valid && below == 0u) h[d] = cnt + (uint32_t)__popcll(match);
    wave_sync();
    return cnt + below;
}

__device__ __forceinline__ void chunk_tiles(uint32_t n, uint32_t b, uint32_t g, uint32_t& t0, uint32_t& t1)
{
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const uint32_t per = (tiles + g - 1) / g;
    t0 = b * per;
    t1 = t0 + per < tiles ? t0 + per : tiles;
    if (t0 > t1) t0 = t1;
}

__global__ __launch_bounds__(kBlock) void k_sort_count(const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ n_dev, int shift,
                                                       uint32_t* __restrict__ counts)
{
    __shared__ uint32_t h[kSortWaves][kRadix];
    const int w = (int)threadIdx.x / kWave;
    const int lane = (int)__lane_id();
#pragma unroll
    for (int i = 0; i < kSortWaves; ++i) h[i][threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = *n_dev;
    uint32_t t0, t1;
    chunk_tiles(n, blockIdx.x, gridDim.x, t0, t1);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t base = t * kSortTile + (uint32_t)w * (kWave * kSortItems) + lane;
        uint32_t kr[kSortItems];
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const uint32_t i = base + (uint32_t)k * kWave;
            kr[k] = i < n ? keys[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const uint32_t i = base + (uint32_t)k * kWave;
            if (i < n) atomicAdd(&h[w][(kr[k] >> shift) & (kRadix - 1)], 1u);
        }
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < kSortWaves; ++i) c += h[i][threadIdx.x];
    // digit-major: the exclusive scan of this array is each chunk's first
    // output position per digit
    counts[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = c;
}

// one workgroup per digit d: the exclusive scan of counts[d * G .. d * G + G)
// in place (each chunk's first slot among the digit's keys) and the digit's
// total; rows are independent, so there is no chain between workgroups
__global__ __launch_bounds__(kBlock) void k_sort_rowscan(uint32_t* __restrict__ counts, uint32_t g,
                                                         uint32_t* __restrict__ totals)
{
    __shared__ uint32_t lds_scan[kSortWaves];
    constexpr uint32_t kPer = (uint32_t)kSortGrid / kBlock;
    uint32_t* row = counts + (size_t)blockIdx.x * g;
    uint32_t v[kPer];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = threadIdx.x * kPer + k;
        v[k] = i < g ? row[i] : 0u;
        s += v[k];
    }
    uint32_t all;
    uint32_t run = block_excl_scan(s, lds_scan, all);
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = threadIdx.x * kPer + k;
        if (i < g) row[i] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = all;
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(
    const uint32_t* __restrict__ keys_src, const uint32_t* __restrict__ vals_src,
    uint32_t* __restrict__ keys_dst, uint32_t* __restrict__ vals_dst,
    const uint32_t* __restrict__ n_dev, int shift, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ totals)
{
    __shared__ uint32_t skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    __shared__ uint32_t whist[kSortWaves][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t offs[kRadix];
    __shared__ uint32_t lds_scan[kSortWaves];

    const uint32_t n = *n_dev;
    uint32_t t0, t1;
    chunk_tiles(n, blockIdx.x, gridDim.x, t0, t1);
    if (t0 >= t1) return;
    const int w = (int)threadIdx.x / kWave;
    const int lane = (int)__lane_id();
    {
        // the digit's first slot: the keys of all smaller digits
        uint32_t all;
        const uint32_t base = block_excl_scan(totals[threadIdx.x], lds_scan, all);
        offs[threadIdx.x] = base + counts[(size_t)threadIdx.x * gridDim.x + blockIdx.x];
    }
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t base = t * kSortTile;
#pragma unroll
        for (int i = 0; i < kSortWaves; ++i) whist[i][threadIdx.x] = 0u;
        __syncthreads();
        uint32_t kr[kSortItems], vr[kSortItems], rank[kSortItems];
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const uint32_t i = base + (uint32_t)w * (kWave * kSortItems) + (uint32_t)k * kWave + lane;
            const bool valid = i < n;
            kr[k] = valid ? keys_src[i] : 0u;
            vr[k] = valid ? vals_src[i] : 0u;
            rank[k] = wave_digit_rank((kr[k] >> shift) & (kRadix - 1), valid, whist[w]);
        }
        __syncthreads();
        {
            // per digit: prefix over the waves and the tile's digit starts
            const uint32_t d = threadIdx.x;
            uint32_t tot = 0;
#pragma unroll
            for (int i = 0; i < kSortWaves; ++i) {
                const uint32_t c = whist[i][d];
                whist[i][d] = tot;
                tot += c;
            }
            uint32_t all;
            dstart[d] = block_excl_scan(tot, lds_scan, all);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const uint32_t i = base + (uint32_t)w * (kWave * kSortItems) + (uint32_t)k * kWave + lane;
            if (i < n) {
                const uint32_t d = (kr[k] >> shift) & (kRadix - 1);
                const uint32_t pos = dstart[d] + whist[w][d] + rank[k];
                skeys[pos] = kr[k];
                svals[pos] = vr[k];
            }
        }
        __syncthreads();
        const uint32_t tile_n = (n - base) < kSortTile ? (n - base) : kSortTile;
        for (uint32_t i = threadIdx.x; i < tile_n; i += kBlock) {
            const uint32_t key = skeys[i];
            const uint32_t d = (key >> shift) & (kRadix - 1);
            const uint32_t dst = offs[d] + (i - dstart[d]);
            if (dst < n) {   // always true when the counts agree; keeps a bad count in bounds
                keys_dst[dst] = key;
                vals_dst[dst] = svals[i];
            }
        }
        __syncthreads();
        // advance the chunk's offsets by this tile's digit counts
        {
            const uint32_t d = threadIdx.x;
            const uint32_t next = d + 1 < kRadix ? dstart[d + 1] : tile_n;
            offs[d] += next - dstart[d];
        }
        __syncthreads();
    }
}

inline uint32_t cdiv(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

size_t scan_chain_words(uint32_t n_cap) { return kChainHead + 2 * (size_t)cdiv(n_cap > 0 ? n_cap : 1, kScanTile); }

void scan_exclusive_1p(const uint32_t* in, uint32_t* out, uint32_t out_stride, uint32_t n_cap,
                       const uint32_t* n_dev, uint32_t* total_dev, const ScanCheck* check, ScanTemp& tmp,
                       hipStream_t s)
{
    const uint32_t nb = cdiv(n_cap > 0 ? n_cap : 1, kScanTile);
    const ScanCheck c = check ? *check : ScanCheck{ nullptr, 0u, 0u, nullptr };
    hipLaunchKernelGGL(k_scan_1p, dim3(nb), dim3(kBlock), 0, s, in, out, out_stride, n_cap, n_dev, total_dev, c,
                       tmp.chain);
}

void fill_u32(uint32_t* p, uint32_t value, size_t n_cap, const uint32_t* n_dev, uint32_t n_mul,
              hipStream_t s)
{
    size_t blocks = (n_cap + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n_cap, n_dev,
                       n_mul);
}

static uint32_t sort_grid(uint32_t n_cap)
{
    const uint32_t tiles = cdiv(n_cap > 0 ? n_cap : 1, kSortTile);
    return tiles < (uint32_t)kSortGrid ? tiles : (uint32_t)kSortGrid;
}

size_t sort_hist_words(uint32_t n_cap) { return 2 * (size_t)kRadix * sort_grid(n_cap); }

size_t sort_chain_words(uint32_t n_cap) { return scan_chain_words((uint32_t)kRadix * sort_grid(n_cap)); }

void radix_sort_pairs(uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      uint32_t n_cap, const uint32_t* n_dev, int bits, SortTemp& tmp, hipStream_t s)
{
    const uint32_t g = sort_grid(n_cap);
    const uint32_t m = (uint32_t)kRadix * g;
    uint32_t* counts = tmp.hist;
    uint32_t* totals = tmp.hist + m;     // kRadix words of the second half
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    if (passes > 4) passes = 4;
    uint32_t* ks = keys_in;
    uint32_t* vs = vals_in;
    for (int p = 0; p < passes; ++p) {
        uint32_t* kd;
        uint32_t* vd;
        if (p == passes - 1) { kd = keys_out; vd = vals_out; }
        else if ((p & 1) == 0) { kd = tmp.keys_tmp; vd = tmp.vals_tmp; }
        else { kd = keys_in; vd = vals_in; }
        hipLaunchKernelGGL(k_sort_count, dim3(g), dim3(kBlock), 0, s, ks, n_dev, 8 * p, counts);
        hipLaunchKernelGGL(k_sort_rowscan, dim3(kRadix), dim3(kBlock), 0, s, counts, g, totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(g), dim3(kBlock), 0, s, ks, vs, kd, vd, n_dev, 8 * p, counts,
                           totals);
        ks = kd;
        vs = vd;
    }
}

}  // namespace hr
