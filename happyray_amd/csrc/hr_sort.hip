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
        if (threadIdx.x == 0) {
            uint32_t excl = 0;
            if (tile == 0) {
                __hip_atomic_store(&flags[0], flag_pack(epoch, kPrefix, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(&flags[tile], flag_pack(epoch, kAggregate, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                int p = (int)tile - 1;
                int spins = 0;
                while (p >= 0) {
                    // the sum travels inside the flag word: relaxed atomics suffice
                    const uint64_t f = __hip_atomic_load(&flags[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint32_t fe = (uint32_t)(f >> 34);
                    const uint64_t st = (f >> 32) & 3u;
                    if (fe == (epoch & 0x3FFFFFFFu) && st == kPrefix) {
                        excl += (uint32_t)f;
                        break;
                    }
                    if (fe == (epoch & 0x3FFFFFFFu) && st == kAggregate) {
                        excl += (uint32_t)f;
                        --p;
                        spins = 0;
                        continue;
                    }
                    if (++spins > kSpinLimit) {
                        excl += tile_sum(in, (uint32_t)p, n);
                        --p;
                        spins = 0;
                        continue;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __hip_atomic_store(&flags[tile], flag_pack(epoch, kPrefix, excl + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
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
constexpr int kSortItems = kSortTile / kBlock;   // 16
constexpr int kSortWaves = kBlock / kWave;       // 4
constexpr int kLookback = 16;                    // tiles read per look-back round trip

// Onesweep-style sort: one upsweep reads the keys once and counts the digits
// of every pass; each pass is then one scatter kernel whose tiles (ids in
// launch order) publish their per-digit counts and find the counts of all
// earlier tiles by decoupled look-back, so a pass needs no histogram kernel
// and no scan of per-tile histograms.

// global digit counts of every pass: hist[pass * 256 + digit] (zeroed before)
__global__ __launch_bounds__(kBlock) void k_sort_upsweep(const uint32_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ n_dev, int passes,
                                                         uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[4][kSortWaves][kRadix];
    const int w = (int)threadIdx.x / kWave;
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int i = 0; i < kSortWaves; ++i) h[p][i][threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p][w][(k >> (8 * p)) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < kSortWaves; ++i) c += h[p][i][threadIdx.x];
        if (c) atomicAdd(&hist[p * kRadix + threadIdx.x], c);
    }
}

// count of digit d in the keys of tile p (look-back fallback)
__device__ uint32_t tile_digit_count(const uint32_t* keys, uint32_t p, uint32_t n, int shift, uint32_t d)
{
    uint32_t c = 0;
    const uint32_t b = p * kSortTile;
    const uint32_t e = b + kSortTile < n ? b + kSortTile : n;
    for (uint32_t i = b; i < e; ++i) c += ((keys[i] >> shift) & (kRadix - 1)) == d ? 1u : 0u;
    return c;
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(
    const uint32_t* __restrict__ keys_src, const uint32_t* __restrict__ vals_src,
    uint32_t* __restrict__ keys_dst, uint32_t* __restrict__ vals_dst,
    const uint32_t* __restrict__ n_dev, int shift, const uint32_t* __restrict__ digit_hist,
    uint32_t* chain)
{
    __shared__ uint32_t skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    __shared__ uint32_t whist[kSortWaves][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t lds_scan[kSortWaves];
    __shared__ uint32_t s_tile, s_epoch;

    unsigned long long* flags = (unsigned long long*)(chain + kChainHead);
    if (threadIdx.x == 0) {
        // acquire: the epoch is read before the id is taken, so the workgroup
        // that takes the last id (and moves the epoch on) comes after it
        s_epoch = __hip_atomic_load(&chain[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        s_tile = atomicAdd(&chain[0], 1u);
    }
    __syncthreads();
    const uint32_t tile = s_tile, epoch = s_epoch;
    const uint32_t n = *n_dev;
    const uint32_t base = tile * kSortTile;
    if (base < n) {
        const int w = (int)threadIdx.x / kWave;
        const int lane = (int)__lane_id();
        const uint64_t lt = __lanemask_lt();
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
            const uint32_t d = (kr[k] >> shift) & (kRadix - 1);
            uint64_t match = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t bal = __ballot(bit);
                match &= bit ? bal : ~bal;
            }
            const uint32_t below = (uint32_t)__popcll(match & lt);
            const uint32_t cnt = whist[w][d];
            wave_sync();
            if (valid && below == 0u) whist[w][d] = cnt + (uint32_t)__popcll(match);
            wave_sync();
            rank[k] = cnt + below;
        }
        __syncthreads();
        {
            // per digit: prefix over waves, tile count, its look-back over the
            // earlier tiles, and the global start of the digit
            const uint32_t d = threadIdx.x;
            uint32_t tot = 0;
#pragma unroll
            for (int i = 0; i < kSortWaves; ++i) {
                const uint32_t c = whist[i][d];
                whist[i][d] = tot;
                tot += c;
            }
            const uint32_t mask = epoch & 0x3FFFFFFFu;
            uint32_t excl = 0;
            if (tile == 0) {
                __hip_atomic_store(&flags[d], flag_pack(epoch, kPrefix, tot), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(&flags[(size_t)tile * kRadix + d], flag_pack(epoch, kAggregate, tot),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // Look back kLookback tiles per round trip: the flags of a
                // window are loaded together, then consumed from the nearest
                // tile until a prefix (done) or an unpublished tile (wait).
                // The count travels inside the flag word, so relaxed atomics
                // suffice.
                int p = (int)tile - 1;
                int spins = 0;
                while (p >= 0) {
                    uint64_t f[kLookback];
#pragma unroll
                    for (int k = 0; k < kLookback; ++k)
                        f[k] = p - k >= 0 ? __hip_atomic_load(&flags[(size_t)(p - k) * kRadix + d], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                          : 0ull;
                    int used = 0;
                    bool done = false;
#pragma unroll
                    for (int k = 0; k < kLookback; ++k) {
                        if (done || used < k || p - k < 0) continue;
                        const uint32_t fe = (uint32_t)(f[k] >> 34);
                        const uint64_t st = (f[k] >> 32) & 3u;
                        if (fe == mask && st != 0u) {
                            excl += (uint32_t)f[k];
                            used = k + 1;
                            done = st == kPrefix;
                        }
                    }
                    if (done) break;
                    p -= used;
                    if (used > 0) {
                        spins = 0;
                        continue;
                    }
                    if (++spins > kSpinLimit) {
                        excl += tile_digit_count(keys_src, (uint32_t)p, n, shift, d);
                        --p;
                        spins = 0;
                        continue;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __hip_atomic_store(&flags[(size_t)tile * kRadix + d], flag_pack(epoch, kPrefix, excl + tot),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            uint32_t all;
            dstart[d] = block_excl_scan(tot, lds_scan, all);
            uint32_t gall;
            const uint32_t dbase = block_excl_scan(digit_hist[d], lds_scan, gall);
            gbase[d] = dbase + excl;
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
            const uint32_t dst = gbase[d] + (i - dstart[d]);
            if (dst < n) {   // always true when the counts agree; keeps a bad count in bounds
                keys_dst[dst] = key;
                vals_dst[dst] = svals[i];
            }
        }
    }
    if (tile == gridDim.x - 1 && threadIdx.x == 0) {
        __hip_atomic_store(&chain[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&chain[1], epoch + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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

size_t sort_hist_words(uint32_t n_cap) { return 4 * (size_t)kRadix; }

size_t sort_chain_words(uint32_t n_cap)
{
    return kChainHead + 2 * (size_t)kRadix * cdiv(n_cap > 0 ? n_cap : 1, kSortTile);
}

void radix_sort_pairs(uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      uint32_t n_cap, const uint32_t* n_dev, int bits, SortTemp& tmp, hipStream_t s)
{
    const uint32_t tiles = cdiv(n_cap > 0 ? n_cap : 1, kSortTile);
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    if (passes > 4) passes = 4;
    fill_u32(tmp.hist, 0u, 4 * (size_t)kRadix, nullptr, 1, s);
    uint32_t ublocks = tiles < 1024 ? tiles : 1024;
    hipLaunchKernelGGL(k_sort_upsweep, dim3(ublocks), dim3(kBlock), 0, s, keys_in, n_dev, passes, tmp.hist);
    uint32_t* ks = keys_in;
    uint32_t* vs = vals_in;
    for (int p = 0; p < passes; ++p) {
        uint32_t* kd;
        uint32_t* vd;
        if (p == passes - 1) { kd = keys_out; vd = vals_out; }
        else if ((p & 1) == 0) { kd = tmp.keys_tmp; vd = tmp.vals_tmp; }
        else { kd = keys_in; vd = vals_in; }
        hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kBlock), 0, s, ks, vs, kd, vd, n_dev, 8 * p,
                           tmp.hist + p * kRadix, tmp.chain);
        ks = kd;
        vs = vd;
    }
}

}  // namespace hr
