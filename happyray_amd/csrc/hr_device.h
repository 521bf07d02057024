// hr_device.h -- device-resident structures and wave64 helpers shared by the
// build, sort and trace kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hr_math.h"

namespace hr {

constexpr int kWave = 64;

// Grid parameters as the builder leaves them in device memory
// (UniformGrid / TwoLevelGrid, src/RT/Structure/UniformGrid.h:37-138,
// src/RT/Structure/TwoLevelGrid.h:163-230).  Written by the setup kernel so
// that no host round trip is needed between bounds and resolution.
struct GridParams {
    float bmin[3], bmax[3];
    int res[3];
    float cs[3], csr[3];        // cell size, reciprocal cell size
    uint32_t num_cells;
    uint32_t fallback;          // 1 when the 32^3 fallback was taken
    uint32_t pad[2];
};

// Build counters and capacity flags.
struct BuildStatus {
    uint32_t bounds_enc[6];     // order-preserving encodings: min xyz, max xyz
    uint32_t num_cells;         // required sizes (valid up to the first overflow)
    uint32_t num_pairs;         // uniform pairs / two-level top pairs
    uint32_t num_leaves;
    uint32_t num_leaf_pairs;
    uint32_t overflow;          // kOverflow* bits
    uint32_t n_pairs_eff;       // sizes the later kernels use: 0 after an overflow
    uint32_t n_leaves_eff;
    uint32_t n_leaf_pairs_eff;
    uint32_t pad[2];
};

// The per-frame selection of a captured frame (HIP graph): key frames, the
// interpolation coefficient and the image id, read by the kernels from device
// memory so that one recorded graph serves every frame.  seq numbers the
// context's captured frames.  skip is set on the device after the frame's
// rebuild when that rebuild, or the rebuild of an earlier frame since the host
// last reset the halt word, outgrew its buffers: the frame's render kernels
// then return at entry.
struct FrameSel {
    uint32_t k1, k2;
    float coeff;
    int image;
    uint32_t skip;
    uint32_t seq;
};

constexpr uint32_t kOverflowCells = 1u;
constexpr uint32_t kOverflowPairs = 2u;
constexpr uint32_t kOverflowLeaves = 4u;
constexpr uint32_t kOverflowLeafPairs = 8u;

// order-preserving float <-> uint encoding for atomic min / max
__device__ __forceinline__ uint32_t enc_f(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
HR_HD float dec_f(uint32_t e)
{
    const uint32_t b = (e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e;
    return __builtin_bit_cast(float, b);
}

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------

// orders LDS traffic between the lanes of one wave (no workgroup barrier)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = (int)__lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, kWave));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
    return v;
}

// largest o in [0, 64) with excl[o] <= j (excl[0] == 0 <= j)
__device__ __forceinline__ int wave_owner(const uint32_t* excl, uint32_t j)
{
    int o = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
        if (excl[o + step] <= j) o += step;
    return o;
}

// triangle record: nine packed floats (v0, v1, v2) per triangle, in a float4
// buffer of ceil(9 nt / 4) elements
__device__ __forceinline__ void load_tri(const float4* __restrict__ rec, uint32_t p, f3& v0, f3& v1,
                                         f3& v2)
{
    const float* f = reinterpret_cast<const float*>(rec) + 9 * (size_t)p;
    v0 = mk(f[0], f[1], f[2]);
    v1 = mk(f[3], f[4], f[5]);
    v2 = mk(f[6], f[7], f[8]);
}

// BBoxExtractor<Triangle> (src/RT/Primitive/BBox.hpp:96-110)
__device__ __forceinline__ void tri_bounds(f3 v0, f3 v1, f3 v2, f3& lo, f3& hi)
{
    lo = rep(kFltMax);
    hi = rep(-kFltMax);
    lo = vmin(lo, v0); hi = vmax(hi, v0);
    lo = vmin(lo, v1); hi = vmax(hi, v1);
    lo = vmin(lo, v2); hi = vmax(hi, v2);
}

// clamped cell range of a box (countPairs, UniformGridBuildKernels.h:59-69);
// an empty range is kept empty instead of going negative.
__device__ __forceinline__ void box_cells(f3 lo, f3 hi, f3 origin, f3 rcp, int rx, int ry, int rz,
                                          int mn[3], int mx[3])
{
    const f3 a = (lo - origin) * rcp;
    const f3 b = (hi - origin) * rcp + rep(1.f);
    const int a0 = (int)a.x, a1 = (int)a.y, a2 = (int)a.z;
    const int b0 = (int)b.x, b1 = (int)b.y, b2 = (int)b.z;
    mn[0] = a0 > 0 ? a0 : 0;
    mn[1] = a1 > 0 ? a1 : 0;
    mn[2] = a2 > 0 ? a2 : 0;
    mx[0] = rx < b0 ? rx : b0;
    mx[1] = ry < b1 ? ry : b1;
    mx[2] = rz < b2 ? rz : b2;
    mx[0] = mx[0] < mn[0] ? mn[0] : mx[0];
    mx[1] = mx[1] < mn[1] ? mn[1] : mx[1];
    mx[2] = mx[2] < mn[2] ? mn[2] : mx[2];
}

// float -> int conversion that saturates (NaN -> the lower bound): defined for
// every input on both the GPU and the host
HR_HD int to_int_sat(float x)
{
    return (int)fminf(fmaxf(x, -2147483520.f), 2147483520.f);
}

// cell range of a box in an object grid of the instanced hierarchy
// (countPairsMultiUniformGrid, TLGridHierarchySortBuilder.cu:131-147): the
// first cell is clamped to [0, res - 1] and the end to [1, res], so a triangle
// that has moved outside its object's grid still lands in a border cell
__device__ __forceinline__ void object_box_cells(f3 lo, f3 hi, f3 origin, f3 rcp, int rx, int ry, int rz,
                                                 int mn[3], int mx[3])
{
    const f3 a = (lo - origin) * rcp;
    const f3 b = (hi - origin) * rcp + rep(1.f);
    const int r[3] = { rx, ry, rz };
    const float av[3] = { a.x, a.y, a.z }, bv[3] = { b.x, b.y, b.z };
    for (int i = 0; i < 3; ++i) {
        int lo_i = to_int_sat(av[i]);
        lo_i = lo_i > 0 ? lo_i : 0;
        mn[i] = lo_i < r[i] - 1 ? lo_i : r[i] - 1;
        int hi_i = to_int_sat(bv[i]);
        hi_i = hi_i < r[i] ? hi_i : r[i];
        hi_i = hi_i > 1 ? hi_i : 1;
        mx[i] = hi_i < mn[i] ? mn[i] : hi_i;
    }
}

// ---------------------------------------------------------------------------
// host launch helpers (hr_build.hip, hr_sort.hip, hr_trace.hip)
// ---------------------------------------------------------------------------

struct ScanTemp {
    uint32_t* chain;        // scan_chain_words(n_cap), zeroed when allocated
    size_t chain_cap;
};

// A capacity check fused into a single-pass scan: the workgroup that knows the
// total sets st->overflow |= bit when total > cap, and *eff = overflow ? 0 : total.
struct ScanCheck {
    BuildStatus* st;
    uint32_t cap;
    uint32_t bit;
    uint32_t* eff;
};

constexpr uint32_t kScanTile = 2048;     // 256 threads x 8
constexpr uint32_t kSortTile = 2048;     // 256 threads x 8

// Exclusive scan of n uint32 (n_cap, or *n_dev when n_dev != nullptr) into
// out[i * out_stride], with decoupled look-back: one launch; tiles take
// ids in launch order, publish their sum, then their inclusive prefix, and a
// tile reads its predecessors' published values.  A predecessor that has not
// published after a bounded wait has its sum recomputed from `in`, so every
// workgroup finishes.  `out` must not alias `in`.  *total_dev and the check
// (both may be null) are written by the tile holding the last element.
void scan_exclusive_1p(const uint32_t* in, uint32_t* out, uint32_t out_stride, uint32_t n_cap,
                       const uint32_t* n_dev, uint32_t* total_dev, const ScanCheck* check, ScanTemp& tmp,
                       hipStream_t s);
size_t scan_chain_words(uint32_t n_cap);

// stable LSD radix sort of (key, value) pairs on the low `bits` key bits
// (at most 32).  n = *n_dev (<= n_cap).  keys_in / vals_in are overwritten
// when the number of passes exceeds 2.  The result lands in keys_out /
// vals_out.  Each pass is a digit count per chunk of tiles, a scan of the
// counts and an in-order scatter per chunk.
struct SortTemp {
    uint32_t* keys_tmp;
    uint32_t* vals_tmp;
    uint32_t* hist;         // sort_hist_words: per-chunk digit counts, then their offsets
    uint32_t* chain;        // sort_chain_words, zeroed when allocated: the offsets scan's flags
};
size_t sort_hist_words(uint32_t n_cap);
size_t sort_chain_words(uint32_t n_cap);
void radix_sort_pairs(uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      uint32_t n_cap, const uint32_t* n_dev, int bits, SortTemp& tmp, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t value, size_t n_cap, const uint32_t* n_dev, uint32_t n_mul,
              hipStream_t s);

}  // namespace hr
