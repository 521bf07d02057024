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
// n return at once.  No kernel here waits on another workgroup.
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

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                        uint32_t n_cap,
                                                        const uint32_t* __restrict__ n_dev,
                                                        uint32_t* __restrict__ block_sums)
{
    __shared__ uint32_t lds[kBlock / kWave];
    const uint32_t n = n_dev ? *n_dev : n_cap;
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    if (base < n) {
#pragma unroll
        for (int k = 0; k < (int)(kScanTile / kBlock); ++k) {
            const uint32_t i = base + k * kBlock + threadIdx.x;
            if (i < n) s += in[i];
        }
    }
    s = wave_sum(s);
    if (__lane_id() == 0) lds[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kBlock / kWave; ++i) t += lds[i];
        block_sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_scan_spine(uint32_t* __restrict__ block_sums,
                                                     uint32_t n_cap,
                                                     const uint32_t* __restrict__ n_dev,
                                                     uint32_t* __restrict__ total_dev)
{
    __shared__ uint32_t lds[1024 / kWave];
    const uint32_t n = n_dev ? *n_dev : n_cap;
    const uint32_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? block_sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, lds, tot);
        if (i < nb) block_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_dev) *total_dev = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t* in, uint32_t* out,
                                                      uint32_t out_stride, uint32_t n_cap,
                                                      const uint32_t* __restrict__ n_dev,
                                                      const uint32_t* __restrict__ block_sums)
{
    __shared__ uint32_t lds[kBlock / kWave];
    const uint32_t n = n_dev ? *n_dev : n_cap;
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n) return;
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
    uint32_t tot;
    uint32_t run = block_excl_scan(s, lds, tot) + block_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t i = first + k;
        if (i < n) out[(size_t)i * out_stride] = run;
        run += v[k];
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

__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ n_dev,
                                                       int shift, uint32_t num_tiles,
                                                       uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[kSortWaves][kRadix];
    const int w = (int)threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < kSortWaves; ++i) h[i][threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t n = *n_dev;
    const uint32_t base = blockIdx.x * kSortTile;
    if (base < n) {
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const uint32_t i = base + k * kBlock + threadIdx.x;
            if (i < n) atomicAdd(&h[w][(keys[i] >> shift) & (kRadix - 1)], 1u);
        }
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < kSortWaves; ++i) c += h[i][threadIdx.x];
    hist[(size_t)threadIdx.x * num_tiles + blockIdx.x] = c;
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const uint32_t* __restrict__ keys_src, const uint32_t* __restrict__ vals_src,
    uint32_t* __restrict__ keys_dst, uint32_t* __restrict__ vals_dst,
    const uint32_t* __restrict__ n_dev, int shift, uint32_t num_tiles,
    const uint32_t* __restrict__ hist_scanned)
{
    __shared__ uint32_t skeys[kSortTile];
    __shared__ uint32_t svals[kSortTile];
    __shared__ uint32_t whist[kSortWaves][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t lds_scan[kSortWaves];

    const uint32_t n = *n_dev;
    const uint32_t base = blockIdx.x * kSortTile;
    if (base >= n) return;
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
    // per digit: exclusive prefix over waves, tile total, then tile-local
    // digit starts by a block scan
    {
        const int d = (int)threadIdx.x;
        uint32_t tot = 0;
#pragma unroll
        for (int i = 0; i < kSortWaves; ++i) {
            const uint32_t c = whist[i][d];
            whist[i][d] = tot;
            tot += c;
        }
        uint32_t all;
        dstart[d] = block_excl_scan(tot, lds_scan, all);
        gbase[d] = hist_scanned[(size_t)d * num_tiles + blockIdx.x];
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
        keys_dst[dst] = key;
        vals_dst[dst] = svals[i];
    }
}

inline uint32_t cdiv(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

size_t scan_temp_words(uint32_t n_cap) { return (size_t)cdiv(n_cap, kScanTile) + 2; }

void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n_cap, const uint32_t* n_dev,
                    uint32_t* total_dev, ScanTemp& tmp, hipStream_t s)
{
    scan_exclusive_strided(in, out, 1u, n_cap, n_dev, total_dev, tmp, s);
}

void scan_exclusive_strided(const uint32_t* in, uint32_t* out, uint32_t out_stride, uint32_t n_cap,
                            const uint32_t* n_dev, uint32_t* total_dev, ScanTemp& tmp,
                            hipStream_t s)
{
    const uint32_t nb = cdiv(n_cap > 0 ? n_cap : 1, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n_cap, n_dev,
                       tmp.block_sums);
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(1024), 0, s, tmp.block_sums, n_cap, n_dev,
                       total_dev);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kBlock), 0, s, in, out, out_stride, n_cap, n_dev,
                       tmp.block_sums);
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

size_t sort_hist_words(uint32_t n_cap) { return (size_t)kRadix * cdiv(n_cap > 0 ? n_cap : 1, kSortTile); }

void radix_sort_pairs(uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      uint32_t n_cap, const uint32_t* n_dev, int bits, SortTemp& tmp, hipStream_t s)
{
    const uint32_t tiles = cdiv(n_cap > 0 ? n_cap : 1, kSortTile);
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    uint32_t* ks = keys_in;
    uint32_t* vs = vals_in;
    for (int p = 0; p < passes; ++p) {
        uint32_t* kd;
        uint32_t* vd;
        if (p == passes - 1) { kd = keys_out; vd = vals_out; }
        else if ((p & 1) == 0) { kd = tmp.keys_tmp; vd = tmp.vals_tmp; }
        else { kd = keys_in; vd = vals_in; }
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlock), 0, s, ks, n_dev, shift, tiles,
                           tmp.hist);
        scan_exclusive(tmp.hist, tmp.hist, (uint32_t)kRadix * tiles, nullptr, nullptr, tmp.scan, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlock), 0, s, ks, vs, kd, vd, n_dev,
                           shift, tiles, tmp.hist);
        ks = kd;
        vs = vd;
    }
}

}  // namespace hr
