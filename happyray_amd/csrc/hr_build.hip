// hr_build.hip -- key-frame interpolation, bounds and the uniform /
// two-level grid builds for gfx950.
//
// Reference call chain replaced (host-synchronous there, device-resident
// here):
//   ObjUploader::uploadObjFrameVertexData   src/Application/WFObjectUploader.h:48-119
//   UGridSortBuilder::init / build          src/RT/Algorithm/UGridSortBuilder.h:64-407
//   TLGridSortBuilder::init / build         src/RT/Algorithm/TLGridSortBuilder.h:66-452
//   countPairs, writeKeysAndValues<exact>   src/RT/Algorithm/UniformGridBuildKernels.h:40-252,
//                                           src/RT/Algorithm/UniformGridBuildKernels.cu:8-383
//   prepareCellRanges                       src/RT/Algorithm/UniformGridBuildKernels.h:384-487
//   countLeafLevelCells, prepareTopLevelCellRanges, countLeafLevelPairs,
//   writeLeafLevelKeysAndValues, prepareLeafCellRanges
//                                           src/RT/Algorithm/TLGridBuildKernels.h:45-834
//
// Design (MI355X):
//  * the bounds reduction and the resolution are computed on the device and
//    left in a GridParams block, so a build has no host round trip; sizes
//    that depend on the data stay in device memory and every buffer is
//    sized by a capacity that the host grows when a BuildStatus overflow bit
//    comes back (then the build is repeated);
//  * pairs are produced by wave64-cooperative enumeration: the 64 lanes of a
//    wave expand the cell ranges of 64 consecutive primitives into one
//    candidate list and test 64 candidates per step, so one large triangle
//    does not serialise a lane.  Output slots come from a ballot prefix and a
//    scan of per-wave counts, so pairs are written in primitive order without
//    atomics and the stable radix sort yields (cell, primitive) order;
//  * exact insertion tests candidates in the count pass and the write pass
//    and writes only the pairs that pass (the reference writes failing pairs
//    with key = #cells and sorts them to the end);
//  * triangles are re-laid once per build as three float4 per triangle
//    (vertex xyz), the record every later pass and the traversal read.
#include "hr_build.h"
#include "hr_hier.h"

namespace hr {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ bool stop(const BuildStatus* st) { return st->overflow != 0u; }

// ---------------------------------------------------------------------------
// reset, interpolation, bounds, resolution
// ---------------------------------------------------------------------------

__global__ void k_reset(BuildStatus* st)
{
    if (threadIdx.x == 0) {
        for (int i = 0; i < 3; ++i) { st->bounds_enc[i] = 0xFFFFFFFFu; st->bounds_enc[3 + i] = 0u; }
        st->num_cells = st->num_pairs = st->num_leaves = st->num_leaf_pairs = 0u;
        st->overflow = 0u;
        st->n_pairs_eff = st->n_leaves_eff = st->n_leaf_pairs_eff = 0u;
        st->pad[0] = st->pad[1] = 0u;
    }
}

// v = v1 * (1 - c) + v2 * c (WFObjectUploader.h:82-88) and min / max of all
// vertices (WFObjectUploader.h:90-95)
// With a frame selection fs (captured frames) the key frames and the
// coefficient come from device memory: blocks fs->k1 and fs->k2 of keys.
__global__ __launch_bounds__(kBlock) void k_lerp_bounds(
    const float* __restrict__ k1, const float* __restrict__ k2, float c, const FrameSel* __restrict__ fs,
    const float* __restrict__ keys, uint32_t nv, float* __restrict__ ox, float* __restrict__ oy,
    float* __restrict__ oz, uint32_t* __restrict__ part)
{
    __shared__ float red[6][kWavesPerBlock];
    if (fs) {
        k1 = keys + (size_t)fs->k1 * 3 * nv;
        k2 = keys + (size_t)fs->k2 * 3 * nv;
        c = fs->coeff;
    }
    const float* __restrict__ x1 = k1;
    const float* __restrict__ y1 = k1 + nv;
    const float* __restrict__ z1 = k1 + 2 * (size_t)nv;
    const float* __restrict__ x2 = k2;
    const float* __restrict__ y2 = k2 + nv;
    const float* __restrict__ z2 = k2 + 2 * (size_t)nv;
    f3 lo = rep(kFltMax), hi = rep(-kFltMax);
    const float ic = 1.f - c;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        const float x = x1[i] * ic + x2[i] * c;
        const float y = y1[i] * ic + y2[i] * c;
        const float z = z1[i] * ic + z2[i] * c;
        ox[i] = x; oy[i] = y; oz[i] = z;
        lo = vmin(mk(x, y, z), lo);
        hi = vmax(mk(x, y, z), hi);
    }
    lo.x = wave_min(lo.x); lo.y = wave_min(lo.y); lo.z = wave_min(lo.z);
    hi.x = wave_max(hi.x); hi.y = wave_max(hi.y); hi.z = wave_max(hi.z);
    const int w = (int)threadIdx.x / kWave;
    if (__lane_id() == 0) {
        red[0][w] = lo.x; red[1][w] = lo.y; red[2][w] = lo.z;
        red[3][w] = hi.x; red[4][w] = hi.y; red[5][w] = hi.z;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = red[threadIdx.x][0];
        for (int i = 1; i < kWavesPerBlock; ++i)
            v = threadIdx.x < 3 ? fminf(v, red[threadIdx.x][i]) : fmaxf(v, red[threadIdx.x][i]);
        // one partial per block: same-address device atomics from every block
        // serialise at the memory side, so the blocks' bounds are combined by
        // k_bounds_final instead
        part[threadIdx.x * gridDim.x + blockIdx.x] = enc_f(v);
    }
}

// the scene bounds from the lerp's per-block partials (one workgroup)
__global__ __launch_bounds__(kBlock) void k_bounds_final(const uint32_t* __restrict__ part, uint32_t nparts,
                                                         BuildStatus* __restrict__ st)
{
    __shared__ uint32_t red[6][kWavesPerBlock];
    uint32_t v[6] = { 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u };
    for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x)
        for (int k = 0; k < 6; ++k) {
            const uint32_t e = part[k * nparts + b];
            v[k] = k < 3 ? (e < v[k] ? e : v[k]) : (e > v[k] ? e : v[k]);
        }
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) {
            const uint32_t o = __shfl_xor(v[k], d, kWave);
            v[k] = k < 3 ? (o < v[k] ? o : v[k]) : (o > v[k] ? o : v[k]);
        }
    const int w = (int)threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int k = 0; k < 6; ++k) red[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = (int)threadIdx.x;
        uint32_t r = red[k][0];
        for (int i = 1; i < kWavesPerBlock; ++i) r = k < 3 ? (red[k][i] < r ? red[k][i] : r) : (red[k][i] > r ? red[k][i] : r);
        st->bounds_enc[k] = r;
    }
}

// Bounds padding (WFObjectUploader.h:112-114) and resolution from density
// (UGridSortBuilder.h:78-92 / TLGridSortBuilder.h:81-98, checkResolution
// UGridMemoryManager.cu:338-346); cell size and its reciprocal as the memory
// managers compute them (UGridMemoryManager.h:107-115).  Fills every field
// of gp but num_cells and returns the cell count.
__device__ uint64_t grid_from_bounds(const uint32_t* bounds_enc, uint32_t nprims, float lambda, GridParams* gp)
{
    f3 lo = mk(dec_f(bounds_enc[0]), dec_f(bounds_enc[1]), dec_f(bounds_enc[2]));
    f3 hi = mk(dec_f(bounds_enc[3]), dec_f(bounds_enc[4]), dec_f(bounds_enc[5]));
    const float diag = len(hi - lo);
    lo = lo - rep(0.001f * diag);
    hi = hi + rep(0.001f * diag);
    const f3 d = hi - lo;
    const float volume = d.x * d.y * d.z;
    const float magic = spec_powf(lambda * (float)nprims / volume, 0.3333333f);
    const f3 dm = d * magic;
    const float dv[3] = { dm.x, dm.y, dm.z };
    int res[3];
    bool bad = false;
    for (int i = 0; i < 3; ++i) {
        if (!(dv[i] >= 1.f && dv[i] < 2147483520.f)) bad = true;
        else res[i] = (int)dv[i];
    }
    gp->fallback = bad ? 1u : 0u;
    if (bad) res[0] = res[1] = res[2] = 32;
    const f3 rf = mk((float)res[0], (float)res[1], (float)res[2]);
    const f3 cs = d / rf;
    const f3 csr = rf / d;
    gp->bmin[0] = lo.x; gp->bmin[1] = lo.y; gp->bmin[2] = lo.z;
    gp->bmax[0] = hi.x; gp->bmax[1] = hi.y; gp->bmax[2] = hi.z;
    gp->res[0] = res[0]; gp->res[1] = res[1]; gp->res[2] = res[2];
    gp->cs[0] = cs.x; gp->cs[1] = cs.y; gp->cs[2] = cs.z;
    gp->csr[0] = csr.x; gp->csr[1] = csr.y; gp->csr[2] = csr.z;
    return (uint64_t)res[0] * (uint64_t)res[1] * (uint64_t)res[2];
}

__global__ void k_grid_setup(GridParams* __restrict__ gp, BuildStatus* __restrict__ st,
                             uint32_t nprims, float lambda, uint64_t cells_cap)
{
    if (threadIdx.x != 0) return;
    const uint64_t nc = grid_from_bounds(st->bounds_enc, nprims, lambda, gp);
    if (nc > cells_cap || nc > 0xFFFFFFF0ull) {
        st->overflow |= kOverflowCells;
        st->num_cells = nc > 0xFFFFFFF0ull ? 0xFFFFFFFFu : (uint32_t)nc;
        gp->num_cells = 0u;
    } else {
        st->num_cells = (uint32_t)nc;
        gp->num_cells = (uint32_t)nc;
    }
}

// cells[0 .. *n) = (0, 0), grid-stride over the launch: a count kernel clears
// the cells (or leaves) that the ranges kernel fills after the sort
__device__ __forceinline__ void zero_cells(uint2* cells, const uint32_t* n)
{
    if (!cells) return;
    const uint32_t m = *n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        cells[i] = make_uint2(0u, 0u);
}



// ---------------------------------------------------------------------------
// exact triangle / cell test (UniformGridBuildKernels.cu:8-48, :305-361)
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool axis_sep(float a, float b, float fa, float fb, float p0a, float p0b,
                                         float p1a, float p1b, float ha, float hb)
{
    const float q0 = a * p0a + b * p0b;
    const float q1 = a * p1a + b * p1b;
    const float mn = fminf(q0, q1);
    const float mx = fmaxf(q0, q1);
    const float rad = fa * ha + fb * hb;
    return !(mn > rad + kEps || mx + kEps < -rad);
}

__device__ __forceinline__ bool tri_cell_overlap(f3 t0, f3 t1, f3 t2, f3 n, f3 c, f3 h)
{
    const f3 a = t0 - c, b = t1 - c, d = t2 - c;
    const f3 e0 = b - a, e1 = d - b, e2 = a - d;
    f3 lo;
    lo.x = (n.x > 0.f) ? -h.x : h.x;
    lo.y = (n.y > 0.f) ? -h.y : h.y;
    lo.z = (n.z > 0.f) ? -h.z : h.z;
    f3 up = -lo;
    up = up - a;
    lo = lo - a;
    bool ok = dot(n, lo) <= 0.f && dot(n, up) > 0.f;
    f3 f = vabs(e0);
    ok = ok && axis_sep(e0.z, -e0.y, f.z, f.y, a.y, a.z, d.y, d.z, h.y, h.z);
    ok = ok && axis_sep(-e0.z, e0.x, f.z, f.x, a.x, a.z, d.x, d.z, h.x, h.z);
    ok = ok && axis_sep(e0.y, -e0.x, f.y, f.x, b.x, b.y, d.x, d.y, h.x, h.y);
    f = vabs(e1);
    ok = ok && axis_sep(e1.z, -e1.y, f.z, f.y, a.y, a.z, d.y, d.z, h.y, h.z);
    ok = ok && axis_sep(-e1.z, e1.x, f.z, f.x, a.x, a.z, d.x, d.z, h.x, h.z);
    ok = ok && axis_sep(e1.y, -e1.x, f.y, f.x, a.x, a.y, b.x, b.y, h.x, h.y);
    f = vabs(e2);
    ok = ok && axis_sep(e2.z, -e2.y, f.z, f.y, a.y, a.z, b.y, b.z, h.y, h.z);
    ok = ok && axis_sep(-e2.z, e2.x, f.z, f.x, a.x, a.z, b.x, b.z, h.x, h.z);
    ok = ok && axis_sep(e2.y, -e2.x, f.y, f.x, b.x, b.y, d.x, d.y, h.x, h.y);
    return ok;
}

// cell-centre coordinate of index i along one axis, accumulated as the
// reference's nested loops do (start - size, then + size per step)
__device__ __forceinline__ float centre_along(float first, float size, int steps)
{
    float c = first - size;
    for (int i = 0; i <= steps; ++i) c += size;
    return c;
}

// ---------------------------------------------------------------------------
// top level (uniform grid / two-level top): triangle records and counts
// ---------------------------------------------------------------------------

struct WaveItems {
    int lo[3][kWave];
    int dx[kWave], dy[kWave];
    uint32_t excl[kWave];
    uint32_t id[kWave];
    // exact insertion
    float v[9][kWave];
    float n[3][kWave];
    float mcc[3][kWave];
    // leaf level
    uint32_t lbeg[kWave];
    uint32_t r0[kWave], r1[kWave];
    // scenes of a batch: per-owner grid width, z stride and first cell
    uint32_t gx[kWave], gz[kWave], cbase[kWave];
};

// scene of a global cell index: the last s with cell_base[s] <= c (cell
// counts are positive, so the bases increase strictly)
__device__ __forceinline__ uint32_t scene_of(const uint32_t* base, uint32_t ns, uint32_t c)
{
    uint32_t lo = 0, hi = ns;          // base[lo] <= c < base[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

// kMulti: the triangles of a batch of scenes; each lane takes the grid of its
// triangle's scene and writes global cell keys (cell_base of the scene + key)
template <bool kExact, bool kWrite, bool kMulti = false>
__global__ __launch_bounds__(kBlock) void k_top_pairs(
    const uint32_t* __restrict__ idx, const float* __restrict__ vx, const float* __restrict__ vy,
    const float* __restrict__ vz, uint32_t nt, const GridParams* __restrict__ gp,
    const BuildStatus* __restrict__ st, float4* __restrict__ rec, uint32_t* __restrict__ wave_counts,
    const uint32_t* __restrict__ wave_offsets, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
    uint32_t waves_cap, SceneTable T = SceneTable{}, uint2* __restrict__ zero = nullptr,
    const uint32_t* __restrict__ zero_n = nullptr)
{
    static_assert(!(kExact && kMulti), "batches use bounding-box insertion");
    __shared__ WaveItems sh[kWavesPerBlock];
    if (stop(st)) return;
    zero_cells(zero, zero_n);
    const int wl = (int)threadIdx.x / kWave;
    const int lane = (int)__lane_id();
    const uint32_t wg = blockIdx.x * kWavesPerBlock + (uint32_t)wl;
    const uint32_t p = wg * kWave + (uint32_t)lane;
    const bool valid = p < nt;
    if (wg * kWave >= nt) {         // whole wave idle
        if (!kWrite && lane == 0 && wg < waves_cap) wave_counts[wg] = 0u;
        return;
    }

    uint32_t cbase = 0;
    if (kMulti) {
        const uint32_t sc = T.tri_scene[valid ? p : nt - 1u];
        gp = T.gps + sc;
        cbase = T.cell_base[sc];
    }
    const f3 bmin = mk(gp->bmin[0], gp->bmin[1], gp->bmin[2]);
    const f3 cs = mk(gp->cs[0], gp->cs[1], gp->cs[2]);
    const f3 csr = mk(gp->csr[0], gp->csr[1], gp->csr[2]);
    const int rx = gp->res[0], ry = gp->res[1], rz = gp->res[2];

    f3 v0 = rep(0.f), v1 = rep(0.f), v2 = rep(0.f);
    uint32_t cnt = 0;
    int mn[3] = { 0, 0, 0 }, mx[3] = { 0, 0, 0 };
    if (valid) {
        if (!kWrite) {
            const uint32_t i0 = idx[3 * (size_t)p], i1 = idx[3 * (size_t)p + 1], i2 = idx[3 * (size_t)p + 2];
            v0 = mk(vx[i0], vy[i0], vz[i0]);
            v1 = mk(vx[i1], vy[i1], vz[i1]);
            v2 = mk(vx[i2], vy[i2], vz[i2]);
            float* f = reinterpret_cast<float*>(rec) + 9 * (size_t)p;
            f[0] = v0.x; f[1] = v0.y; f[2] = v0.z;
            f[3] = v1.x; f[4] = v1.y; f[5] = v1.z;
            f[6] = v2.x; f[7] = v2.y; f[8] = v2.z;
        } else {
            load_tri(rec, p, v0, v1, v2);
        }
        f3 lo, hi;
        tri_bounds(v0, v1, v2, lo, hi);
        if (kMulti && T.object_grids) object_box_cells(lo, hi, bmin, csr, rx, ry, rz, mn, mx);
        else box_cells(lo, hi, bmin, csr, rx, ry, rz, mn, mx);
        cnt = (uint32_t)((mx[0] - mn[0]) * (mx[1] - mn[1]) * (mx[2] - mn[2]));
    }

    const uint32_t kcell_z = (uint32_t)rx * (uint32_t)ry;
    if (!kExact && !kWrite) {
        const uint32_t s = wave_sum(cnt);
        if (lane == 0) wave_counts[wg] = s;
        return;
    }

    WaveItems& W = sh[wl];
    const uint32_t inc = wave_incl_scan(cnt);
    const uint32_t total = __shfl(inc, kWave - 1, kWave);
    W.excl[lane] = inc - cnt;
    W.lo[0][lane] = mn[0]; W.lo[1][lane] = mn[1]; W.lo[2][lane] = mn[2];
    W.dx[lane] = mx[0] - mn[0];
    W.dy[lane] = mx[1] - mn[1];
    W.id[lane] = p;
    if (kMulti) {
        W.gx[lane] = (uint32_t)rx;
        W.gz[lane] = kcell_z;
        W.cbase[lane] = cbase;
    }
    if (kExact) {
        W.v[0][lane] = v0.x; W.v[1][lane] = v0.y; W.v[2][lane] = v0.z;
        W.v[3][lane] = v1.x; W.v[4][lane] = v1.y; W.v[5][lane] = v1.z;
        W.v[6][lane] = v2.x; W.v[7][lane] = v2.y; W.v[8][lane] = v2.z;
        const f3 n = normalize(cross(v1 - v0, v2 - v0));
        W.n[0][lane] = n.x; W.n[1][lane] = n.y; W.n[2][lane] = n.z;
        const f3 half = cs * 0.505f;
        f3 m = mk((float)mn[0], (float)mn[1], (float)mn[2]);
        m = m * cs;
        m = m + bmin + half;
        W.mcc[0][lane] = m.x; W.mcc[1][lane] = m.y; W.mcc[2][lane] = m.z;
    }
    wave_sync();

    const f3 half = cs * 0.505f;
    const uint64_t lt = __lanemask_lt();
    uint32_t run = kWrite ? wave_offsets[wg] : 0u;
    for (uint32_t j0 = 0; j0 < total; j0 += kWave) {
        const uint32_t j = j0 + (uint32_t)lane;
        bool ok = false;
        uint32_t key = 0, id = 0;
        if (j < total) {
            const int o = wave_owner(W.excl, j);
            const uint32_t k = j - W.excl[o];
            const uint32_t dxo = (uint32_t)W.dx[o], dyo = (uint32_t)W.dy[o];
            const int ix = (int)(k % dxo), t = (int)(k / dxo);
            const int iy = t % (int)dyo, iz = t / (int)dyo;
            const int x = W.lo[0][o] + ix, y = W.lo[1][o] + iy, z = W.lo[2][o] + iz;
            id = W.id[o];
            if (kMulti) key = W.cbase[o] + (uint32_t)x + (uint32_t)y * W.gx[o] + (uint32_t)z * W.gz[o];
            else key = (uint32_t)x + (uint32_t)y * (uint32_t)rx + (uint32_t)z * kcell_z;
            if (kExact) {
                const f3 c = mk(centre_along(W.mcc[0][o], cs.x, ix), centre_along(W.mcc[1][o], cs.y, iy),
                                centre_along(W.mcc[2][o], cs.z, iz));
                const f3 a = mk(W.v[0][o], W.v[1][o], W.v[2][o]);
                const f3 b = mk(W.v[3][o], W.v[4][o], W.v[5][o]);
                const f3 d = mk(W.v[6][o], W.v[7][o], W.v[8][o]);
                ok = tri_cell_overlap(a, b, d, mk(W.n[0][o], W.n[1][o], W.n[2][o]), c, half);
            } else {
                ok = true;
            }
        }
        const uint64_t bal = __ballot(ok);
        if (kWrite && ok) {
            const uint32_t pos = run + (uint32_t)__popcll(bal & lt);
            keys[pos] = key;
            vals[pos] = id;
        }
        run += (uint32_t)__popcll(bal);
    }
    if (!kWrite && lane == 0) wave_counts[wg] = run;
}

// prepareCellRanges (UniformGridBuildKernels.h:384-487) on sorted keys:
// begin at the first pair of a key, end one past its last pair.  Keys are
// always < #cells here (no invalid-key tail).
// Cell ranges of sorted keys.  The build's last launch also seals the grid
// (seal_gp != nullptr): after an overflow the grid is unusable, and a zero
// resolution makes every ray miss at its entry test, so kernels that run after
// the build without a host check (captured frames) read nothing out of bounds.
__global__ void k_ranges(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ n_dev,
                         uint2* __restrict__ cells, const BuildStatus* __restrict__ st,
                         GridParams* __restrict__ seal_gp = nullptr)
{
    if (seal_gp && blockIdx.x == 0 && threadIdx.x == 0 && st->overflow != 0u) {
        seal_gp->res[0] = seal_gp->res[1] = seal_gp->res[2] = 0;
        seal_gp->num_cells = 0u;
    }
    if (stop(st)) return;
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        if (i == 0 || keys[i - 1] != k) cells[k].x = i;
        if (i + 1 == n || keys[i + 1] != k) cells[k].y = i + 1;
    }
}

// ---------------------------------------------------------------------------
// leaf level
// ---------------------------------------------------------------------------

// countLeafLevelCells<false> (TLGridBuildKernels.h:57-114): leaf resolution
// per top cell and the packed TwoLevelGridCell (TwoLevelGrid.h:38-161)
template <bool kMulti = false>
__global__ void k_leaf_res(const GridParams* __restrict__ gp, uint2* __restrict__ cells,
                           uint32_t* __restrict__ counts, float lambda_leaf,
                           const BuildStatus* __restrict__ st, SceneTable T = SceneTable{})
{
    if (stop(st)) return;
    const uint32_t nc = gp->num_cells;
    f3 cs = mk(gp->cs[0], gp->cs[1], gp->cs[2]);
    float vol = cs.x * cs.y * cs.z;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if (kMulti) {
            const uint32_t sc = scene_of(T.cell_base, T.ns, c);
            const GridParams* g = T.gps + sc;
            cs = mk(g->cs[0], g->cs[1], g->cs[2]);
            vol = cs.x * cs.y * cs.z;
            lambda_leaf = T.lambda_leaf[sc];
        }
        const uint2 r = cells[c];
        const uint32_t n = r.y - r.x;
        const float magic = spec_powf(lambda_leaf * (float)n / vol, 0.3333333f);
        const f3 res = vmin(rep(255.f), vmax(rep(1.f), cs * magic));
        const bool leaf = n <= 16u;
        const uint32_t lx = leaf ? 1u : (uint32_t)res.x;
        const uint32_t ly = leaf ? 1u : (uint32_t)res.y;
        const uint32_t lz = leaf ? 1u : (uint32_t)res.z;
        counts[c] = lx * ly * lz;
        uint32_t y = leaf ? 0x10000000u : 0u;
        y |= 0x40000000u;
        if (n == 0u) y &= 0x3FFFFFFFu;
        y |= (lx & 0xFFu) | ((ly & 0xFFu) << 8) | ((lz & 0xFFu) << 16);
        cells[c].y = y;
    }
}

// countLeafLevelPairs / writeLeafLevelKeysAndValues
// (TLGridBuildKernels.h:304-409, :536-663) over the sorted top pairs
template <bool kWrite, bool kMulti = false>
__global__ __launch_bounds__(kBlock) void k_leaf_pairs(
    const uint32_t* __restrict__ top_keys, const uint32_t* __restrict__ top_vals,
    const uint32_t* __restrict__ n_top, const uint2* __restrict__ cells,
    const GridParams* __restrict__ gp, const BuildStatus* __restrict__ st,
    const float4* __restrict__ rec, uint32_t* __restrict__ wave_counts,
    const uint32_t* __restrict__ wave_offsets, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
    uint32_t waves_cap, SceneTable T = SceneTable{}, uint2* __restrict__ zero = nullptr,
    const uint32_t* __restrict__ zero_n = nullptr)
{
    __shared__ WaveItems sh[kWavesPerBlock];
    if (stop(st)) return;
    zero_cells(zero, zero_n);
    const uint32_t n = *n_top;
    const int wl = (int)threadIdx.x / kWave;
    const int lane = (int)__lane_id();
    const uint32_t wg = blockIdx.x * kWavesPerBlock + (uint32_t)wl;
    if (wg * kWave >= n) {
        if (!kWrite && lane == 0 && wg < waves_cap) wave_counts[wg] = 0u;
        return;
    }
    const uint32_t i = wg * kWave + (uint32_t)lane;
    const bool valid = i < n;

    uint32_t cbase = 0;
    if (kMulti) {
        const uint32_t sc = T.tri_scene[top_vals[valid ? i : n - 1u]];
        gp = T.gps + sc;
        cbase = T.cell_base[sc];
    }
    const f3 bmin = mk(gp->bmin[0], gp->bmin[1], gp->bmin[2]);
    const f3 cs = mk(gp->cs[0], gp->cs[1], gp->cs[2]);
    const uint32_t gx = (uint32_t)gp->res[0], gy = (uint32_t)gp->res[1];

    uint32_t cnt = 0, prim = 0, lb = 0, r0 = 1, r1 = 1;
    int mn[3] = { 0, 0, 0 }, mx[3] = { 0, 0, 0 };
    if (valid) {
        const uint32_t gkey = top_keys[i];
        const uint32_t key = gkey - cbase;
        prim = top_vals[i];
        f3 v0, v1, v2, lo, hi;
        load_tri(rec, prim, v0, v1, v2);
        tri_bounds(v0, v1, v2, lo, hi);
        const uint32_t tx = key % gx;
        const uint32_t ty = (key % (gx * gy)) / gx;
        const uint32_t tz = key / (gx * gy);
        const uint2 cell = cells[gkey];
        lb = cell.x;
        const int c0 = (int)(cell.y & 0xFFu), c1 = (int)((cell.y >> 8) & 0xFFu),
                  c2 = (int)((cell.y >> 16) & 0xFFu);
        r0 = (uint32_t)c0;
        r1 = (uint32_t)c1;
        const f3 cres = mk((float)c0, (float)c1, (float)c2);
        f3 org;
        org.x = (float)tx * cs.x + bmin.x;
        org.y = (float)ty * cs.y + bmin.y;
        org.z = (float)tz * cs.z + bmin.z;
        const f3 sub_rcp = cres / cs;
        box_cells(lo, hi, org, sub_rcp, c0, c1, c2, mn, mx);
        cnt = (uint32_t)((mx[0] - mn[0]) * (mx[1] - mn[1]) * (mx[2] - mn[2]));
    }
    if (!kWrite) {
        const uint32_t s = wave_sum(cnt);
        if (lane == 0) wave_counts[wg] = s;
        return;
    }

    WaveItems& W = sh[wl];
    const uint32_t inc = wave_incl_scan(cnt);
    const uint32_t total = __shfl(inc, kWave - 1, kWave);
    W.excl[lane] = inc - cnt;
    W.lo[0][lane] = mn[0]; W.lo[1][lane] = mn[1]; W.lo[2][lane] = mn[2];
    W.dx[lane] = mx[0] - mn[0];
    W.dy[lane] = mx[1] - mn[1];
    W.id[lane] = prim;
    W.lbeg[lane] = lb;
    W.r0[lane] = r0;
    W.r1[lane] = r1;
    wave_sync();

    const uint64_t lt = __lanemask_lt();
    uint32_t run = wave_offsets[wg];
    for (uint32_t j0 = 0; j0 < total; j0 += kWave) {
        const uint32_t j = j0 + (uint32_t)lane;
        const bool ok = j < total;
        uint32_t key = 0, id = 0;
        if (ok) {
            const int o = wave_owner(W.excl, j);
            const uint32_t k = j - W.excl[o];
            const uint32_t dxo = (uint32_t)W.dx[o], dyo = (uint32_t)W.dy[o];
            const uint32_t ix = k % dxo, t = k / dxo;
            const uint32_t iy = t % dyo, iz = t / dyo;
            const uint32_t x = (uint32_t)W.lo[0][o] + ix, y = (uint32_t)W.lo[1][o] + iy,
                           z = (uint32_t)W.lo[2][o] + iz;
            key = x + y * W.r0[o] + z * W.r0[o] * W.r1[o] + W.lbeg[o];
            id = W.id[o];
        }
        const uint64_t bal = __ballot(ok);
        if (ok) {
            const uint32_t pos = run + (uint32_t)__popcll(bal & lt);
            keys[pos] = key;
            vals[pos] = id;
        }
        run += (uint32_t)__popcll(bal);
    }
}

// ---------------------------------------------------------------------------
// batches of scenes
// ---------------------------------------------------------------------------

__global__ void k_batch_reset(SceneTable T)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < T.ns; s += gridDim.x * blockDim.x)
        for (int i = 0; i < 3; ++i) {
            T.bounds_enc[6 * s + i] = 0xFFFFFFFFu;
            T.bounds_enc[6 * s + 3 + i] = 0u;
        }
}

// per-scene bounds (the reduction of k_lerp_bounds, over static vertices):
// a wave whose 64 vertices share a scene reduces them and issues one atomic
// per bound; a wave across a scene boundary updates per lane
__global__ __launch_bounds__(kBlock) void k_batch_bounds(const float* __restrict__ vx,
                                                         const float* __restrict__ vy,
                                                         const float* __restrict__ vz, uint32_t nv,
                                                         SceneTable T)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~(uint32_t)(kWave - 1); base < nv;
         base += stride) {
        const uint32_t i = base + __lane_id();
        const bool valid = i < nv;
        f3 v = rep(0.f);
        if (valid) v = mk(vx[i], vy[i], vz[i]);
        const uint32_t last = (base + kWave - 1 < nv ? base + kWave - 1 : nv - 1);
        const uint32_t s0 = scene_of(T.vert_base, T.ns, base);
        const uint32_t s1 = scene_of(T.vert_base, T.ns, last);
        if (s0 == s1) {
            f3 lo = valid ? v : rep(kFltMax), hi = valid ? v : rep(-kFltMax);
            lo.x = wave_min(lo.x); lo.y = wave_min(lo.y); lo.z = wave_min(lo.z);
            hi.x = wave_max(hi.x); hi.y = wave_max(hi.y); hi.z = wave_max(hi.z);
            if (__lane_id() == 0) {
                uint32_t* b = T.bounds_enc + 6 * (size_t)s0;
                atomicMin(b + 0, enc_f(lo.x)); atomicMin(b + 1, enc_f(lo.y)); atomicMin(b + 2, enc_f(lo.z));
                atomicMax(b + 3, enc_f(hi.x)); atomicMax(b + 4, enc_f(hi.y)); atomicMax(b + 5, enc_f(hi.z));
            }
        } else if (valid) {
            uint32_t* b = T.bounds_enc + 6 * (size_t)scene_of(T.vert_base, T.ns, i);
            atomicMin(b + 0, enc_f(v.x)); atomicMin(b + 1, enc_f(v.y)); atomicMin(b + 2, enc_f(v.z));
            atomicMax(b + 3, enc_f(v.x)); atomicMax(b + 4, enc_f(v.y)); atomicMax(b + 5, enc_f(v.z));
        }
    }
}

// per-scene resolution and cell count
__global__ void k_batch_setup(SceneTable T, BuildStatus* __restrict__ st)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < T.ns; s += gridDim.x * blockDim.x) {
        const uint32_t nprims = T.tri_base[s + 1] - T.tri_base[s];
        GridParams* g = T.gps + s;
        const uint64_t nc = grid_from_bounds(T.bounds_enc + 6 * (size_t)s, nprims, T.lambda_top[s], g);
        if (nc > 0xFFFFFFF0ull) {
            atomicOr(&st->overflow, kOverflowCells);
            T.cell_counts[s] = 0u;
            g->num_cells = 0u;
        } else {
            T.cell_counts[s] = (uint32_t)nc;
            g->num_cells = (uint32_t)nc;
        }
    }
}

// total cells after the scan of the per-scene counts
__global__ void k_batch_total(SceneTable T, GridParams* agg, BuildStatus* st, uint64_t cells_cap)
{
    if (threadIdx.x != 0) return;
    const uint32_t total = st->num_cells;
    T.cell_base[T.ns] = total;
    if ((uint64_t)total > cells_cap || (st->overflow & kOverflowCells)) {
        st->overflow |= kOverflowCells;
        agg->num_cells = 0u;
    } else {
        agg->num_cells = total;
    }
}

// first position of each scene in sorted keys: lower bound of the scene's
// first key (top level: its first cell; leaf level: its first leaf)
template <bool kLeaf>
__global__ void k_batch_bases(SceneTable T, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ n_dev,
                              const uint2* __restrict__ cells, const uint32_t* __restrict__ num_leaves,
                              const BuildStatus* __restrict__ st)
{
    if (stop(st)) return;
    const uint32_t n = *n_dev;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s <= T.ns; s += gridDim.x * blockDim.x) {
        uint32_t first;
        if (kLeaf) {
            first = s < T.ns ? cells[T.cell_base[s]].x : *num_leaves;
            T.leaf_base[s] = first;
        } else {
            first = T.cell_base[s];
        }
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (keys[mid] < first) lo = mid + 1;
            else hi = mid;
        }
        if (kLeaf) T.pair_base[s] = s < T.ns ? lo : n;
        else T.top_base[s] = s < T.ns ? lo : n;
    }
}

__global__ void k_batch_seal(SceneTable T, const BuildStatus* __restrict__ st)
{
    if (st->overflow == 0u) return;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < T.ns; s += gridDim.x * blockDim.x) {
        T.gps[s].res[0] = T.gps[s].res[1] = T.gps[s].res[2] = 0;
        T.gps[s].num_cells = 0u;
    }
}

inline uint32_t cdiv(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

inline int bits_for(uint64_t n)
{
    int b = 1;
    while (b < 32 && (1ull << b) < n) ++b;
    return b;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

void launch_lerp_bounds(const float* k1, const float* k2, float c, uint32_t nv, float* out,
                        BuildStatus* st, uint32_t* part, hipStream_t s, const FrameSel* fs, const float* keys)
{
    uint32_t blocks = cdiv(nv > 0 ? nv : 1, kBlock);
    if (blocks > kBoundsParts) blocks = kBoundsParts;
    hipLaunchKernelGGL(k_lerp_bounds, dim3(blocks), dim3(kBlock), 0, s, k1, k2, c, fs, keys, nv, out, out + nv,
                       out + 2 * (size_t)nv, part);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlock), 0, s, part, blocks, st);
}

void launch_reset(BuildStatus* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, st);
}

void launch_grid_setup(GridParams* gp, BuildStatus* st, uint32_t nprims, float lambda,
                       uint64_t cells_cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, s, gp, st, nprims, lambda, cells_cap);
}

void run_build(const BuildArgs& a, BuildBuffers& b, hipStream_t s)
{
    const uint32_t nt = a.nt;
    const uint32_t waves = cdiv(nt > 0 ? nt : 1, kWave);
    const uint32_t blocks = cdiv(waves, kWavesPerBlock);
    BuildStatus* st = b.status;
    const bool exact = a.kind == 1;
    const bool two = a.kind == 2;

    launch_grid_setup(b.params, st, nt, a.lambda_top, b.cells_cap, s);

    // ---- top level / uniform: count (and clear the cells), scan, write ----
    if (exact)
        hipLaunchKernelGGL((k_top_pairs<true, false>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx,
                           a.vy, a.vz, nt, b.params, st, b.rec, b.wave_counts, nullptr, nullptr, nullptr, b.waves_cap,
                           SceneTable{}, b.cells, &b.params->num_cells);
    else
        hipLaunchKernelGGL((k_top_pairs<false, false>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx,
                           a.vy, a.vz, nt, b.params, st, b.rec, b.wave_counts, nullptr, nullptr, nullptr, b.waves_cap,
                           SceneTable{}, b.cells, &b.params->num_cells);
    const ScanCheck pairs_check{ st, b.pairs_cap, kOverflowPairs, &st->n_pairs_eff };
    scan_exclusive_1p(b.wave_counts, b.wave_offsets, 1, waves, nullptr, &st->num_pairs, &pairs_check, b.scan, s);
    if (exact)
        hipLaunchKernelGGL((k_top_pairs<true, true>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx,
                           a.vy, a.vz, nt, b.params, st, b.rec, nullptr, b.wave_offsets, b.keys, b.vals, b.waves_cap);
    else
        hipLaunchKernelGGL((k_top_pairs<false, true>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx,
                           a.vy, a.vz, nt, b.params, st, b.rec, nullptr, b.wave_offsets, b.keys, b.vals, b.waves_cap);

    // ---- sort by cell key, then cell ranges ----
    const int cell_bits = bits_for(b.cells_cap);
    uint32_t* out_vals = two ? b.top_vals_sorted : b.prims;
    radix_sort_pairs(b.keys, b.vals, b.keys_sorted, out_vals, b.pairs_cap, &st->n_pairs_eff, cell_bits,
                     b.sort, s);
    const uint32_t rblocks = cdiv(b.pairs_cap > 0 ? b.pairs_cap : 1, 256) > 4096 ? 4096
                                 : cdiv(b.pairs_cap > 0 ? b.pairs_cap : 1, 256);
    hipLaunchKernelGGL(k_ranges, dim3(rblocks), dim3(256), 0, s, b.keys_sorted, &st->n_pairs_eff, b.cells, st,
                       two ? nullptr : b.params);
    if (!two) return;

    // ---- leaf resolution, leaf begin offsets ----
    const uint32_t cblocks = cdiv(b.cells_cap, 256) > 4096 ? 4096 : cdiv(b.cells_cap, 256);
    hipLaunchKernelGGL(k_leaf_res, dim3(cblocks), dim3(256), 0, s, b.params, b.cells, b.leaf_counts,
                       a.lambda_leaf, st);
    const ScanCheck leaves_check{ st, b.leaves_cap, kOverflowLeaves, &st->n_leaves_eff };
    scan_exclusive_1p(b.leaf_counts, (uint32_t*)b.cells, 2, (uint32_t)b.cells_cap, &b.params->num_cells,
                      &st->num_leaves, &leaves_check, b.scan, s);

    // ---- leaf pairs: count, scan, write ----
    const uint32_t lwaves = cdiv(b.pairs_cap > 0 ? b.pairs_cap : 1, kWave);
    const uint32_t lblocks = cdiv(lwaves, kWavesPerBlock);
    hipLaunchKernelGGL((k_leaf_pairs<false>), dim3(lblocks), dim3(kBlock), 0, s, b.keys_sorted,
                       b.top_vals_sorted, &st->n_pairs_eff, b.cells, b.params, st, b.rec, b.wave_counts,
                       nullptr, nullptr, nullptr, b.waves_cap, SceneTable{}, b.leaves, &st->n_leaves_eff);
    const ScanCheck lp_check{ st, b.leaf_pairs_cap, kOverflowLeafPairs, &st->n_leaf_pairs_eff };
    scan_exclusive_1p(b.wave_counts, b.wave_offsets, 1, lwaves, nullptr, &st->num_leaf_pairs, &lp_check, b.scan,
                      s);
    hipLaunchKernelGGL((k_leaf_pairs<true>), dim3(lblocks), dim3(kBlock), 0, s, b.keys_sorted,
                       b.top_vals_sorted, &st->n_pairs_eff, b.cells, b.params, st, b.rec, nullptr,
                       b.wave_offsets, b.keys, b.vals, b.waves_cap);

    // ---- leaf sort and ranges ----
    const int leaf_bits = bits_for(b.leaves_cap);
    radix_sort_pairs(b.keys, b.vals, b.keys_sorted, b.prims, b.leaf_pairs_cap, &st->n_leaf_pairs_eff,
                     leaf_bits, b.sort, s);
    const uint32_t lrblocks = cdiv(b.leaf_pairs_cap > 0 ? b.leaf_pairs_cap : 1, 256) > 4096 ? 4096
                                  : cdiv(b.leaf_pairs_cap > 0 ? b.leaf_pairs_cap : 1, 256);
    hipLaunchKernelGGL(k_ranges, dim3(lrblocks), dim3(256), 0, s, b.keys_sorted, &st->n_leaf_pairs_eff,
                       b.leaves, st, b.params);
}

}  // namespace hr

namespace hr {

void run_build_batch(const BuildArgs& a, BuildBuffers& b, const SceneTable& T, uint32_t nv_total, hipStream_t s)
{
    const uint32_t nt = a.nt;
    const uint32_t waves = cdiv(nt > 0 ? nt : 1, kWave);
    const uint32_t blocks = cdiv(waves, kWavesPerBlock);
    BuildStatus* st = b.status;
    const uint32_t sblocks = cdiv(T.ns + 1, 256);

    launch_reset(st, s);
    hipLaunchKernelGGL(k_batch_reset, dim3(sblocks), dim3(256), 0, s, T);
    uint32_t vblocks = cdiv(nv_total > 0 ? nv_total : 1, kBlock);
    if (vblocks > 2048) vblocks = 2048;
    hipLaunchKernelGGL(k_batch_bounds, dim3(vblocks), dim3(kBlock), 0, s, a.vx, a.vy, a.vz, nv_total, T);
    hipLaunchKernelGGL(k_batch_setup, dim3(sblocks), dim3(256), 0, s, T, st);
    scan_exclusive_1p(T.cell_counts, T.cell_base, 1, T.ns, nullptr, &st->num_cells, nullptr, b.scan, s);
    hipLaunchKernelGGL(k_batch_total, dim3(1), dim3(64), 0, s, T, b.params, st, b.cells_cap);

    // ---- top level: count, scan, write ----
    hipLaunchKernelGGL((k_top_pairs<false, false, true>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx, a.vy,
                       a.vz, nt, b.params, st, b.rec, b.wave_counts, nullptr, nullptr, nullptr, b.waves_cap, T);
    const ScanCheck pairs_check{ st, b.pairs_cap, kOverflowPairs, &st->n_pairs_eff };
    scan_exclusive_1p(b.wave_counts, b.wave_offsets, 1, waves, nullptr, &st->num_pairs, &pairs_check, b.scan, s);
    hipLaunchKernelGGL((k_top_pairs<false, true, true>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx, a.vy,
                       a.vz, nt, b.params, st, b.rec, nullptr, b.wave_offsets, b.keys, b.vals, b.waves_cap, T);

    // ---- sort by global cell key, cell ranges ----
    radix_sort_pairs(b.keys, b.vals, b.keys_sorted, b.top_vals_sorted, b.pairs_cap, &st->n_pairs_eff,
                     bits_for(b.cells_cap), b.sort, s);
    hipLaunchKernelGGL((k_batch_bases<false>), dim3(sblocks), dim3(256), 0, s, T, b.keys_sorted, &st->n_pairs_eff,
                       b.cells, nullptr, st);
    fill_u32((uint32_t*)b.cells, 0u, 2 * (size_t)b.cells_cap, &b.params->num_cells, 2, s);
    const uint32_t rblocks = cdiv(b.pairs_cap > 0 ? b.pairs_cap : 1, 256) > 4096 ? 4096
                                 : cdiv(b.pairs_cap > 0 ? b.pairs_cap : 1, 256);
    hipLaunchKernelGGL(k_ranges, dim3(rblocks), dim3(256), 0, s, b.keys_sorted, &st->n_pairs_eff, b.cells, st);

    // ---- leaf resolution, leaf begin offsets ----
    const uint32_t cblocks = cdiv(b.cells_cap, 256) > 4096 ? 4096 : cdiv(b.cells_cap, 256);
    hipLaunchKernelGGL((k_leaf_res<true>), dim3(cblocks), dim3(256), 0, s, b.params, b.cells, b.leaf_counts, 0.f,
                       st, T);
    const ScanCheck leaves_check{ st, b.leaves_cap, kOverflowLeaves, &st->n_leaves_eff };
    scan_exclusive_1p(b.leaf_counts, (uint32_t*)b.cells, 2, (uint32_t)b.cells_cap, &b.params->num_cells,
                      &st->num_leaves, &leaves_check, b.scan, s);

    // ---- leaf pairs ----
    const uint32_t lwaves = cdiv(b.pairs_cap > 0 ? b.pairs_cap : 1, kWave);
    const uint32_t lblocks = cdiv(lwaves, kWavesPerBlock);
    hipLaunchKernelGGL((k_leaf_pairs<false, true>), dim3(lblocks), dim3(kBlock), 0, s, b.keys_sorted,
                       b.top_vals_sorted, &st->n_pairs_eff, b.cells, b.params, st, b.rec, b.wave_counts, nullptr,
                       nullptr, nullptr, b.waves_cap, T);
    const ScanCheck lp_check{ st, b.leaf_pairs_cap, kOverflowLeafPairs, &st->n_leaf_pairs_eff };
    scan_exclusive_1p(b.wave_counts, b.wave_offsets, 1, lwaves, nullptr, &st->num_leaf_pairs, &lp_check, b.scan,
                      s);
    hipLaunchKernelGGL((k_leaf_pairs<true, true>), dim3(lblocks), dim3(kBlock), 0, s, b.keys_sorted,
                       b.top_vals_sorted, &st->n_pairs_eff, b.cells, b.params, st, b.rec, nullptr, b.wave_offsets,
                       b.keys, b.vals, b.waves_cap, T);

    // ---- leaf sort and ranges ----
    radix_sort_pairs(b.keys, b.vals, b.keys_sorted, b.prims, b.leaf_pairs_cap, &st->n_leaf_pairs_eff,
                     bits_for(b.leaves_cap), b.sort, s);
    hipLaunchKernelGGL((k_batch_bases<true>), dim3(sblocks), dim3(256), 0, s, T, b.keys_sorted,
                       &st->n_leaf_pairs_eff, b.cells, &st->n_leaves_eff, st);
    fill_u32((uint32_t*)b.leaves, 0u, 2 * (size_t)b.leaves_cap, &st->n_leaves_eff, 2, s);
    const uint32_t lrblocks = cdiv(b.leaf_pairs_cap > 0 ? b.leaf_pairs_cap : 1, 256) > 4096 ? 4096
                                  : cdiv(b.leaf_pairs_cap > 0 ? b.leaf_pairs_cap : 1, 256);
    hipLaunchKernelGGL(k_ranges, dim3(lrblocks), dim3(256), 0, s, b.keys_sorted, &st->n_leaf_pairs_eff, b.leaves,
                       st);
    hipLaunchKernelGGL(k_batch_seal, dim3(sblocks), dim3(256), 0, s, T, st);
}

}  // namespace hr

// ---------------------------------------------------------------------------
// instanced hierarchy: object grids, then a top grid over instance boxes
// ---------------------------------------------------------------------------
//
// The object grids reuse the pair kernels of batched builds (one "scene" per
// object, bounding-box insertion with the object-grid clamp, global cell key =
// first cell of the object + local cell); the instance boxes are then
// inserted into the top grid.  Both levels go through the stable radix sort,
// so every cell lists its ids in ascending order.

namespace hr {

namespace {

// resolution along one axis from diagonal * magic: values below 1 (and NaN)
// give 1; -1 marks a resolution too large for an int
__device__ __forceinline__ int hier_res(float r)
{
    if (r >= 2147483520.f) return -1;
    return r >= 1.f ? (int)r : 1;
}

__device__ void hier_grid(f3 lo, f3 hi, float lambda, float count, GridParams* g, uint64_t* ncells)
{
    const f3 d = hi - lo;
    const float volume = d.x * d.y * d.z;
    const float magic = spec_powf(lambda * count / volume, 0.3333333f);
    const f3 r = d * magic;
    int res[3] = { hier_res(r.x), hier_res(r.y), hier_res(r.z) };
    const bool bad = res[0] < 0 || res[1] < 0 || res[2] < 0;
    if (bad) res[0] = res[1] = res[2] = 1;
    const f3 rf = mk((float)res[0], (float)res[1], (float)res[2]);
    const f3 cs = d / rf, csr = rf / d;
    g->bmin[0] = lo.x; g->bmin[1] = lo.y; g->bmin[2] = lo.z;
    g->bmax[0] = hi.x; g->bmax[1] = hi.y; g->bmax[2] = hi.z;
    g->res[0] = res[0]; g->res[1] = res[1]; g->res[2] = res[2];
    g->cs[0] = cs.x; g->cs[1] = cs.y; g->cs[2] = cs.z;
    g->csr[0] = csr.x; g->csr[1] = csr.y; g->csr[2] = csr.z;
    g->fallback = 0u;
    g->pad[0] = g->pad[1] = 0u;
    *ncells = bad ? ~0ull : (uint64_t)res[0] * (uint64_t)res[1] * (uint64_t)res[2];
}

// top grid: the scene bounds padded by 0.001 |diagonal| (WFObjectUploader.h:
// 112-114), resolution max(1, diagonal * (lambda * #instances / volume)^0.3333333)
// (TLGridHierarchySortBuilder.cu:275-291), cell size diagonal / resolution
// (TLGridHierarchyMemoryManager.h:111-119)
__global__ void k_hier_top_setup(GridParams* __restrict__ gp, BuildStatus* __restrict__ st, uint32_t ninst,
                                 float lambda, uint64_t cells_cap)
{
    if (threadIdx.x != 0) return;
    f3 lo = mk(dec_f(st->bounds_enc[0]), dec_f(st->bounds_enc[1]), dec_f(st->bounds_enc[2]));
    f3 hi = mk(dec_f(st->bounds_enc[3]), dec_f(st->bounds_enc[4]), dec_f(st->bounds_enc[5]));
    const float diag = len(hi - lo);
    lo = lo - rep(0.001f * diag);
    hi = hi + rep(0.001f * diag);
    uint64_t nc;
    hier_grid(lo, hi, lambda, (float)ninst, gp, &nc);
    if (nc > cells_cap || nc > 0xFFFFFFF0ull) {
        st->overflow |= kOverflowCells;
        st->num_cells = nc > 0xFFFFFFF0ull ? 0xFFFFFFFFu : (uint32_t)nc;
        gp->num_cells = 0u;
    } else {
        st->num_cells = (uint32_t)nc;
        gp->num_cells = (uint32_t)nc;
    }
}

__global__ void k_obj_reset(uint32_t* __restrict__ enc, uint32_t nobj)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nobj) return;
    for (int k = 0; k < 3; ++k) {
        enc[6 * (size_t)i + k] = 0xFFFFFFFFu;
        enc[6 * (size_t)i + 3 + k] = 0u;
    }
}

__device__ __forceinline__ void flush_bounds(uint32_t* enc, uint32_t ob, f3 lo, f3 hi)
{
    uint32_t* b = enc + 6 * (size_t)ob;
    atomicMin(b + 0, enc_f(lo.x)); atomicMin(b + 1, enc_f(lo.y)); atomicMin(b + 2, enc_f(lo.z));
    atomicMax(b + 3, enc_f(hi.x)); atomicMax(b + 4, enc_f(hi.y)); atomicMax(b + 5, enc_f(hi.z));
}

// bounds of each object's triangles on the first key frame of the current
// frame (RTEngine.cu:270-289).  Each block takes a contiguous run of
// per_block triangles; a thread keeps running bounds while its triangles stay
// in one object (objects are contiguous triangle ranges) and flushes them at
// a change; a block whose triangles all share one object reduces its bounds
// into its partial, which k_obj_bounds_final combines per object (same-address
// device atomics from many blocks serialise at the memory side).
__global__ __launch_bounds__(kBlock) void k_obj_bounds(const uint32_t* __restrict__ idx,
                                                       const float* __restrict__ kx, const float* __restrict__ ky,
                                                       const float* __restrict__ kz, uint32_t nt,
                                                       const uint32_t* __restrict__ tri_obj,
                                                       uint32_t* __restrict__ enc, uint32_t per_block,
                                                       uint32_t* __restrict__ part)
{
    __shared__ float red[6][kWavesPerBlock];
    __shared__ uint32_t obj0;
    __shared__ int uniform;
    const uint32_t none = 0xFFFFFFFFu;
    const uint32_t b0 = blockIdx.x * per_block;
    const uint32_t b1 = nt - b0 < per_block ? nt : b0 + per_block;
    uint32_t cur = none;
    f3 lo = rep(kFltMax), hi = rep(-kFltMax);
    for (uint32_t t = b0 + threadIdx.x; t < b1; t += kBlock) {
        const uint32_t ob = tri_obj[t];
        if (ob != cur) {
            if (cur != none) flush_bounds(enc, cur, lo, hi);
            cur = ob;
            lo = rep(kFltMax);
            hi = rep(-kFltMax);
        }
        for (int c = 0; c < 3; ++c) {
            const uint32_t v = idx[3 * (size_t)t + c];
            const f3 p = mk(kx[v], ky[v], kz[v]);
            lo = vmin(lo, p);
            hi = vmax(hi, p);
        }
    }
    if (threadIdx.x == 0) {
        obj0 = cur;          // thread 0 has triangles whenever the block has any
        uniform = 1;
    }
    __syncthreads();
    if (cur != none && cur != obj0) uniform = 0;
    __syncthreads();
    if (obj0 == none) return;
    if (!uniform) {
        if (threadIdx.x == 0) part[7 * (size_t)blockIdx.x] = none;
        if (cur != none) flush_bounds(enc, cur, lo, hi);
        return;
    }
    const float v[6] = { wave_min(lo.x), wave_min(lo.y), wave_min(lo.z), wave_max(hi.x), wave_max(hi.y),
                         wave_max(hi.z) };
    const int w = (int)threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int k = 0; k < 6; ++k) red[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        f3 l = mk(red[0][0], red[1][0], red[2][0]), h = mk(red[3][0], red[4][0], red[5][0]);
        for (int i = 1; i < kWavesPerBlock; ++i) {
            l = vmin(l, mk(red[0][i], red[1][i], red[2][i]));
            h = vmax(h, mk(red[3][i], red[4][i], red[5][i]));
        }
        uint32_t* q = part + 7 * (size_t)blockIdx.x;
        q[0] = obj0;
        q[1] = enc_f(l.x); q[2] = enc_f(l.y); q[3] = enc_f(l.z);
        q[4] = enc_f(h.x); q[5] = enc_f(h.y); q[6] = enc_f(h.z);
    }
}

// one wave per object: the partials of the blocks inside its triangle range
__global__ __launch_bounds__(kWave) void k_obj_bounds_final(const uint32_t* __restrict__ obj_first, uint32_t nobj,
                                                             uint32_t per_block, const uint32_t* __restrict__ part,
                                                             uint32_t* __restrict__ enc)
{
    const uint32_t k = blockIdx.x;
    if (k >= nobj) return;
    const uint32_t t0 = obj_first[k], t1 = obj_first[k + 1];
    if (t0 >= t1) return;
    const uint32_t b0 = t0 / per_block, b1 = (t1 - 1) / per_block;
    uint32_t v[6] = { 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u };
    bool any = false;
    for (uint32_t b = b0 + __lane_id(); b <= b1; b += kWave) {
        const uint32_t* q = part + 7 * (size_t)b;
        if (q[0] != k) continue;
        any = true;
        for (int i = 0; i < 3; ++i) {
            v[i] = q[1 + i] < v[i] ? q[1 + i] : v[i];
            v[3 + i] = q[4 + i] > v[3 + i] ? q[4 + i] : v[3 + i];
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) {
            const uint32_t o = __shfl_xor(v[i], d, kWave);
            v[i] = i < 3 ? (o < v[i] ? o : v[i]) : (o > v[i] ? o : v[i]);
        }
    if (__ballot(any) != 0ull && __lane_id() == 0) {
        uint32_t* e = enc + 6 * (size_t)k;
        atomicMin(e + 0, v[0]); atomicMin(e + 1, v[1]); atomicMin(e + 2, v[2]);
        atomicMax(e + 3, v[3]); atomicMax(e + 4, v[4]); atomicMax(e + 5, v[5]);
    }
}

// object grid: bounds padded by 10 EPS (RTEngine.cu:291-296; an object without
// triangles keeps the inverted FLT_MAX box), resolution from the leaf density
// and the object's triangle count (RTEngine.cu:312-326, countGridCellsMultiUniformGrid
// TLGridHierarchySortBuilder.cu:29-68)
__global__ void k_obj_setup(const uint32_t* __restrict__ enc, const uint32_t* __restrict__ obj_first,
                            uint32_t nobj, float lambda, GridParams* __restrict__ gps,
                            uint32_t* __restrict__ cell_counts, BuildStatus* __restrict__ st)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nobj) return;
    const uint32_t count = obj_first[k + 1] - obj_first[k];
    f3 lo = rep(kFltMax), hi = rep(-kFltMax);
    if (count > 0) {
        const uint32_t* e = enc + 6 * (size_t)k;
        lo = mk(dec_f(e[0]), dec_f(e[1]), dec_f(e[2]));
        hi = mk(dec_f(e[3]), dec_f(e[4]), dec_f(e[5]));
    }
    const float pad = 10.f * kEps;
    lo = mk(lo.x - pad, lo.y - pad, lo.z - pad);
    hi = mk(hi.x + pad, hi.y + pad, hi.z + pad);
    uint64_t nc;
    hier_grid(lo, hi, lambda, (float)count, gps + k, &nc);
    if (nc > 0x0FFFFFFFull) {
        // one object's grid alone beyond 2^28 cells: no capacity growth helps
        atomicOr(&st->overflow, kOverflowCells);
        st->num_cells = 0xFFFFFFFFu;
        nc = 0;
    }
    gps[k].num_cells = (uint32_t)nc;
    cell_counts[k] = stop(st) ? 0u : (uint32_t)nc;
}

// countPairs over instance boxes (UniformGridBuildKernels.h:40-105): one lane
// per instance
__global__ void k_inst_count(const InstanceDev* __restrict__ inst, uint32_t ninst, const GridParams* __restrict__ gp,
                             const BuildStatus* __restrict__ st, uint32_t* __restrict__ counts,
                             const GridParams* __restrict__ obj_gp, const uint32_t* __restrict__ obj_base,
                             InstGrid* __restrict__ inst_grid)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    if (stop(st)) {
        counts[i] = 0u;
        return;
    }
    const InstanceDev I = inst[i];
    {
        // the instance's object grid, for the traversal
        const GridParams& O = obj_gp[I.object];
        InstGrid q;
        for (int k = 0; k < 3; ++k) {
            q.bmin[k] = O.bmin[k];
            q.bmax[k] = O.bmax[k];
            q.cs[k] = O.cs[k];
            q.csr[k] = O.csr[k];
            q.res[k] = O.res[k];
        }
        q.base = obj_base[I.object];
        inst_grid[i] = q;
    }
    int mn[3], mx[3];
    box_cells(mk(I.bmin[0], I.bmin[1], I.bmin[2]), mk(I.bmax[0], I.bmax[1], I.bmax[2]),
              mk(gp->bmin[0], gp->bmin[1], gp->bmin[2]), mk(gp->csr[0], gp->csr[1], gp->csr[2]), gp->res[0],
              gp->res[1], gp->res[2], mn, mx);
    counts[i] = (uint32_t)((mx[0] - mn[0]) * (mx[1] - mn[1]) * (mx[2] - mn[2]));
}

// writeKeysAndValues over instance boxes (UniformGridBuildKernels.h:180-252),
// one lane per pair: pair j belongs to the last instance whose first pair is
// <= j and is its (j - first)-th cell in x, then y, then z order -- the same
// sequence a serial loop per instance writes, without one lane walking a box
// that covers the whole grid
__global__ void k_inst_write(const InstanceDev* __restrict__ inst, uint32_t ninst, const GridParams* __restrict__ gp,
                             const BuildStatus* __restrict__ st, const uint32_t* __restrict__ offsets,
                             const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals)
{
    if (stop(st)) return;
    const uint32_t n = *n_dev;
    const int rx = gp->res[0], ry = gp->res[1], rz = gp->res[2];
    const f3 org = mk(gp->bmin[0], gp->bmin[1], gp->bmin[2]);
    const f3 csr = mk(gp->csr[0], gp->csr[1], gp->csr[2]);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ninst;      // offsets[lo] <= j < offsets[hi] (offsets[ninst] = n)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offsets[mid] <= j) lo = mid;
            else hi = mid;
        }
        const InstanceDev I = inst[lo];
        int mn[3], mx[3];
        box_cells(mk(I.bmin[0], I.bmin[1], I.bmin[2]), mk(I.bmax[0], I.bmax[1], I.bmax[2]), org, csr, rx, ry, rz,
                  mn, mx);
        const uint32_t k = j - offsets[lo];
        const uint32_t dx = (uint32_t)(mx[0] - mn[0]), dy = (uint32_t)(mx[1] - mn[1]);
        const uint32_t q = k / dx;
        const uint32_t x = (uint32_t)mn[0] + k % dx, y = (uint32_t)mn[1] + q % dy, z = (uint32_t)mn[2] + q / dy;
        keys[j] = x + y * (uint32_t)rx + z * (uint32_t)rx * (uint32_t)ry;
        vals[j] = lo;
    }
}

inline uint32_t range_blocks(uint32_t cap)
{
    const uint32_t b = cdiv(cap > 0 ? cap : 1, 256);
    return b > 4096 ? 4096 : b;
}

}  // namespace

void run_build_hier(const HierArgs& a, HierBuffers& b, hipStream_t s)
{
    BuildStatus* st = b.status;
    const uint32_t nt = a.nt;
    const uint32_t nob = a.nobj;
    const uint32_t oblocks = cdiv(nob > 0 ? nob : 1, 256);

    // ---- grids: top, then one per object (counts scanned into first cells) ----
    hipLaunchKernelGGL(k_hier_top_setup, dim3(1), dim3(64), 0, s, b.top, st, a.ninst, a.lambda_top, b.cells_cap);
    hipLaunchKernelGGL(k_obj_reset, dim3(oblocks), dim3(256), 0, s, b.obj_bounds_enc, nob);
    if (nt > 0) {
        uint32_t tblocks = cdiv(nt, 256);   // one triangle per thread (0.197 -> 0.162 ms of rebuild against two)
        const uint32_t cap = (a.bounds_blocks >= 1 && a.bounds_blocks <= kObjBoundsBlocks) ? a.bounds_blocks
                                                                                        : kObjBoundsBlocks;
        if (tblocks > cap) tblocks = cap;
        const uint32_t per_block = cdiv(cdiv(nt, tblocks), kBlock) * kBlock;
        hipLaunchKernelGGL(k_obj_bounds, dim3(cdiv(nt, per_block)), dim3(kBlock), 0, s, a.idx, a.kx, a.ky, a.kz, nt,
                           a.tri_obj, b.obj_bounds_enc, per_block, b.obj_part);
        hipLaunchKernelGGL(k_obj_bounds_final, dim3(nob), dim3(kWave), 0, s, a.obj_first, nob, per_block, b.obj_part,
                           b.obj_bounds_enc);
    }
    hipLaunchKernelGGL(k_obj_setup, dim3(oblocks), dim3(256), 0, s, b.obj_bounds_enc, a.obj_first, nob,
                       a.lambda_leaf, b.obj_gp, b.obj_cell_counts, st);
    const ScanCheck cells_check{ st, b.leaves_cap, kOverflowLeaves, &st->n_leaves_eff };
    scan_exclusive_1p(b.obj_cell_counts, b.obj_cell_base, 1, nob, nullptr, &st->num_leaves, &cells_check, b.scan, s);

    // ---- object references: count (and clear the object cells), scan, write, sort, ranges ----
    SceneTable T{};
    T.ns = nob;
    T.tri_scene = a.tri_obj;
    T.gps = b.obj_gp;
    T.cell_base = b.obj_cell_base;
    T.object_grids = 1;
    const uint32_t waves = cdiv(nt > 0 ? nt : 1, kWave);
    const uint32_t blocks = cdiv(waves, kWavesPerBlock);
    hipLaunchKernelGGL((k_top_pairs<false, false, true>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx, a.vy, a.vz,
                       nt, b.top, st, b.rec, b.wave_counts, nullptr, nullptr, nullptr, b.waves_cap, T, b.obj_cells,
                       &st->n_leaves_eff);
    const ScanCheck refs_check{ st, b.leaf_pairs_cap, kOverflowLeafPairs, &st->n_leaf_pairs_eff };
    scan_exclusive_1p(b.wave_counts, b.wave_offsets, 1, waves, nullptr, &st->num_leaf_pairs, &refs_check, b.scan, s);
    hipLaunchKernelGGL((k_top_pairs<false, true, true>), dim3(blocks), dim3(kBlock), 0, s, a.idx, a.vx, a.vy, a.vz,
                       nt, b.top, st, b.rec, nullptr, b.wave_offsets, b.keys, b.vals, b.waves_cap, T);
    radix_sort_pairs(b.keys, b.vals, b.keys_sorted, b.prims, b.leaf_pairs_cap, &st->n_leaf_pairs_eff,
                     bits_for(b.leaves_cap), b.sort, s);
    hipLaunchKernelGGL(k_ranges, dim3(range_blocks(b.leaf_pairs_cap)), dim3(256), 0, s, b.keys_sorted,
                       &st->n_leaf_pairs_eff, b.obj_cells, st, nullptr);

    // ---- top references: count, scan, write, sort, clear the top cells, ranges (and seal) ----
    const uint32_t iblocks = cdiv(a.ninst > 0 ? a.ninst : 1, 256);
    hipLaunchKernelGGL(k_inst_count, dim3(iblocks), dim3(256), 0, s, a.inst, a.ninst, b.top, st, b.inst_counts,
                       b.obj_gp, b.obj_cell_base, b.inst_grid);
    const ScanCheck inst_check{ st, b.pairs_cap, kOverflowPairs, &st->n_pairs_eff };
    scan_exclusive_1p(b.inst_counts, b.inst_offsets, 1, a.ninst, nullptr, &st->num_pairs, &inst_check, b.scan, s);
    hipLaunchKernelGGL(k_inst_write, dim3(range_blocks(b.pairs_cap)), dim3(256), 0, s, a.inst, a.ninst, b.top, st,
                       b.inst_offsets, &st->n_pairs_eff, b.keys, b.vals);
    radix_sort_pairs(b.keys, b.vals, b.keys_sorted, b.inst_refs, b.pairs_cap, &st->n_pairs_eff,
                     bits_for(b.cells_cap), b.sort, s);
    fill_u32((uint32_t*)b.top_cells, 0u, 2 * (size_t)b.cells_cap, &b.top->num_cells, 2, s);
    hipLaunchKernelGGL(k_ranges, dim3(range_blocks(b.pairs_cap)), dim3(256), 0, s, b.keys_sorted, &st->n_pairs_eff,
                       b.top_cells, st, b.top);
}

}  // namespace hr
