// hr_flat.h -- the flat per-lane traversal state (one unit of work per step:
// a triangle test or a DDA advance) shared by the persistent kernels
// (hr_persist.hip) and the wave-cooperative trace of the path tracer
// (hr_trace.hip), and the wave-batched candidate tests both use.
//
// The operations on each ray, and their order, are those of the reference
// traversal (src/RT/Algorithm/TLGridTraverser.h:132-350, UGridTraverser.h:
// 44-177, RayTriangleIntersector.h:52-90), so every kernel built on them
// returns the one-ray-per-lane kernels' results bit for bit.
#pragma once

#include "hr_traverse.h"

namespace hr {
namespace trav {

constexpr float kInf = __builtin_huge_valf();

struct RayState {
    f3 org, dir, rcp;
    float t, tterm;
    uint32_t best;
    int cx, cy, cz;
    float tx, ty, tz, tentry;
    int lx, ly, lz;
    float ux, uy, uz;
    float ddx, ddy, ddz;
    uint32_t cell;      // packed word of the current top cell (two-level)
    uint32_t lbeg;      // its first leaf
    uint32_t it, end;   // candidate range still to test
    uint32_t pn;        // the primitive id at it, loaded one test ahead
    bool in_leaf;
};

__device__ __forceinline__ uint32_t cell_index(const GridRegs& G, int cx, int cy, int cz)
{
    return (uint32_t)cx + (uint32_t)cy * (uint32_t)G.rx + (uint32_t)cz * ((uint32_t)G.rx * (uint32_t)G.ry);
}

__device__ __forceinline__ void leaf_load(const TraceScene& S, RayState& r, Work& wk)
{
    ++wk.steps;
    const uint32_t r0 = r.cell & 0xFFu, r1 = (r.cell >> 8) & 0xFFu;
    const uint2 leaf = S.leaves[r.lbeg + (uint32_t)r.lx + r0 * (uint32_t)r.ly + r0 * r1 * (uint32_t)r.lz];
    r.it = leaf.x;
    r.end = leaf.y;
    if (r.it < r.end) r.pn = S.prims[r.it];
}

// visit the current top cell: non-empty cells start their leaf DDA at
// tentry (TLGridTraverser.h:147-197) and load the first leaf range
__device__ __forceinline__ void top_visit(const GridRegs& G, const TraceScene& S, RayState& r, f3 org, Work& wk)
{
    ++wk.steps;
    const uint2 cell = S.cells[cell_index(G, r.cx, r.cy, r.cz)];
    r.it = r.end = 0u;
    r.in_leaf = false;
    if (!(cell.y & 0x40000000u)) return;
    const int r0 = (int)(cell.y & 0xFFu), r1 = (int)((cell.y >> 8) & 0xFFu), r2 = (int)((cell.y >> 16) & 0xFFu);
    const f3 ep = org + rep(r.tentry) / r.rcp;
    const f3 cres = mk((float)r0, (float)r1, (float)r2);
    const f3 srcp = G.csr * cres;
    f3 mb;
    mb.x = (float)r.cx * G.cs.x + G.bmin.x;
    mb.y = (float)r.cy * G.cs.y + G.bmin.y;
    mb.z = (float)r.cz * G.cs.z + G.bmin.z;
    f3 cf;
    cf.x = floorf((ep - mb).x * srcp.x);
    cf.y = floorf((ep - mb).y * srcp.y);
    cf.z = floorf((ep - mb).z * srcp.z);
    const float scx = 1.f / srcp.x, scy = 1.f / srcp.y, scz = 1.f / srcp.z;
    cf = vmin(cres - rep(1.f), vmax(cf, rep(0.f)));
    r.ux = ((cf.x + (r.rcp.x > 0.f ? 1.f : 0.f)) * scx + mb.x - org.x) * r.rcp.x;
    r.uy = ((cf.y + (r.rcp.y > 0.f ? 1.f : 0.f)) * scy + mb.y - org.y) * r.rcp.y;
    r.uz = ((cf.z + (r.rcp.z > 0.f ? 1.f : 0.f)) * scz + mb.z - org.z) * r.rcp.z;
    r.lx = (int)cf.x;
    r.ly = (int)cf.y;
    r.lz = (int)cf.z;
    r.ddx = scx * fabsf(r.rcp.x);
    r.ddy = scy * fabsf(r.rcp.y);
    r.ddz = scz * fabsf(r.rcp.z);
    r.cell = cell.y;
    r.lbeg = cell.x;
    r.in_leaf = true;
    leaf_load(S, r, wk);
}

__device__ __forceinline__ void top_visit(const GridRegs& G, const TraceScene& S, RayState& r, Work& wk)
{
    top_visit(G, S, r, r.org, wk);
}

__device__ __forceinline__ void uniform_load(const GridRegs& G, const TraceScene& S, RayState& r, Work& wk)
{
    ++wk.steps;
    const uint2 cell = S.cells[cell_index(G, r.cx, r.cy, r.cz)];
    r.it = cell.x;
    r.end = cell.y;
    if (r.it < r.end) r.pn = S.prims[r.it];
}

// MollerTrumboreIntersectionTest (RayTriangleIntersector.h:52-90): the hit
// distance, or +inf when the triangle is missed (barycentrics outside, or
// not beyond the intersection epsilon); a NaN distance compares false too
__device__ __forceinline__ float mt_dist(f3 org, f3 dir, f3 v0, f3 v1, f3 v2)
{
    const f3 e1 = v1 - v0;
    const f3 e2 = v2 - v0;
    const f3 pv = cross(dir, e2);
    const float det_rcp = 1.f / dot(e1, pv);
    f3 tv = org - v0;
    const float alpha = det_rcp * dot(tv, pv);
    tv = cross(tv, e1);
    const float beta = det_rcp * dot(tv, dir);
    const float dist = det_rcp * dot(e2, tv);
    return (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > kIsectEps) ? dist : kInf;
}

// one ray/triangle test; true when an any-hit ray is finished.  kMixed: the
// ray's kind is its own -- any hit when tterm > 0 (it stops once t < tterm),
// closest hit when tterm = 0 -- for kernels that trace both kinds
template <bool kAny, bool kMixed = false>
__device__ __forceinline__ bool test_one(const TraceScene& S, RayState& r, Work& wk)
{
    ++wk.tests;
    const uint32_t p = r.pn;
    ++r.it;
    // the next candidate's id is fetched while this triangle is loaded and tested
    if (r.it < r.end) r.pn = S.prims[r.it];
    f3 v0, v1, v2;
    load_tri(S.rec, p, v0, v1, v2);
    const float dist = mt_dist(r.org, r.dir, v0, v1, v2);
    if (dist < r.t) {
        r.t = dist;
        r.best = p;
        return kMixed ? r.t < r.tterm : kAny;
    }
    return false;
}

// the step after a cell's candidates are exhausted; false when the ray ends
template <bool kAny, bool kMixed = false>
__device__ __forceinline__ bool advance_uniform(const GridRegs& G, const TraceScene& S, RayState& r, Work& wk)
{
    const int d = min_dim(r.tx, r.ty, r.tz);
    if ((kAny || kMixed) && r.t < r.tterm) return false;
    bool flag;
    if (d == 0) {
        flag = r.t > r.tx;
        r.cx += r.rcp.x > 0.f ? 1 : -1;
        r.tx += G.cs.x * fabsf(r.rcp.x);
        flag = flag && r.cx < G.rx && r.cx > -1;
    } else if (d == 1) {
        flag = r.t > r.ty;
        r.cy += r.rcp.y > 0.f ? 1 : -1;
        r.ty += G.cs.y * fabsf(r.rcp.y);
        flag = flag && r.cy < G.ry && r.cy > -1;
    } else {
        flag = r.t > r.tz;
        r.cz += r.rcp.z > 0.f ? 1 : -1;
        r.tz += G.cs.z * fabsf(r.rcp.z);
        flag = flag && r.cz < G.rz && r.cz > -1;
    }
    if (!flag) return false;
    uniform_load(G, S, r, wk);
    return true;
}

// the leaf part of a two-level advance (TLGridTraverser.h:209-286): 0 = the
// ray has ended, 1 = it moved to the next leaf of its top cell (loaded),
// 2 = it left the top cell's leaves (the top-level step comes next)
template <bool kAny, bool kMixed = false>
__device__ __forceinline__ int leaf_step(const TraceScene& S, RayState& r, Work& wk)
{
    if ((kAny || kMixed) && r.t < r.tterm) return 0;
    const int d = min_dim(r.ux, r.uy, r.uz);
    const int r0 = (int)(r.cell & 0xFFu), r1 = (int)((r.cell >> 8) & 0xFFu), r2 = (int)((r.cell >> 16) & 0xFFu);
    bool second;
    if (d == 0) {
        if (r.t < r.ux) return 0;
        r.lx += r.rcp.x > 0.f ? 1 : -1;
        r.ux += r.ddx;
        second = r.lx != r0 && r.lx != -1;
    } else if (d == 1) {
        if (r.t < r.uy) return 0;
        r.ly += r.rcp.y > 0.f ? 1 : -1;
        r.uy += r.ddy;
        second = r.ly != r1 && r.ly != -1;
    } else {
        if (r.t < r.uz) return 0;
        r.lz += r.rcp.z > 0.f ? 1 : -1;
        r.uz += r.ddz;
        second = r.lz != r2 && r.lz != -1;
    }
    if (second) {
        leaf_load(S, r, wk);
        return 1;
    }
    r.in_leaf = false;
    return 2;
}

// the top-level step (TLGridTraverser.h:292-347) and the visit of the next
// top cell; false when the ray leaves the grid or its hit lies before the cell
__device__ __forceinline__ bool top_step(const GridRegs& G, const TraceScene& S, RayState& r, f3 org, Work& wk)
{
    const int d = min_dim(r.tx, r.ty, r.tz);
    bool flag;
    if (d == 0) {
        flag = r.t > r.tx;
        r.tentry = r.tx;
        r.cx += r.rcp.x > 0.f ? 1 : -1;
        r.tx += G.cs.x * fabsf(r.rcp.x);
        flag = flag && r.cx != G.rx && r.cx != -1;
    } else if (d == 1) {
        flag = r.t > r.ty;
        r.tentry = r.ty;
        r.cy += r.rcp.y > 0.f ? 1 : -1;
        r.ty += G.cs.y * fabsf(r.rcp.y);
        flag = flag && r.cy != G.ry && r.cy != -1;
    } else {
        flag = r.t > r.tz;
        r.tentry = r.tz;
        r.cz += r.rcp.z > 0.f ? 1 : -1;
        r.tz += G.cs.z * fabsf(r.rcp.z);
        flag = flag && r.cz != G.rz && r.cz != -1;
    }
    if (!flag) return false;
    top_visit(G, S, r, org, wk);
    return true;
}

__device__ __forceinline__ bool top_step(const GridRegs& G, const TraceScene& S, RayState& r, Work& wk)
{
    return top_step(G, S, r, r.org, wk);
}

template <bool kAny, bool kMixed = false>
__device__ __forceinline__ bool advance_two_level(const GridRegs& G, const TraceScene& S, RayState& r,
                                                  Work& wk)
{
    if (r.in_leaf) {
        const int k = leaf_step<kAny, kMixed>(S, r, wk);
        if (k != 2) return k == 1;
    }
    return top_step(G, S, r, wk);
}

// set up the DDA of a generated ray; false when it does not enter the grid
template <bool kTwo, bool kAny, bool kMixed = false>
__device__ __forceinline__ bool ray_begin(const GridRegs& G, const TraceScene& S, RayState& r, Work& wk)
{
    r.best = S.nt;
    r.it = r.end = 0u;
    r.in_leaf = false;
    bool flag = !((kAny || kMixed) && r.t < r.tterm);
    flag = dda_setup<!kTwo>(G, r.org, r.rcp, r.tentry, r.cx, r.cy, r.cz, r.tx, r.ty, r.tz) && flag;
    if (!flag) return false;
    if (kTwo) top_visit(G, S, r, r.org, wk);
    else uniform_load(G, S, r, wk);
    return true;
}

__device__ __forceinline__ void set_ray(RayState& r, f3 org, f3 d)
{
    r.org = org;
    r.rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    r.dir = mk(1.f / r.rcp.x, 1.f / r.rcp.y, 1.f / r.rcp.z);
}

// Wave-batched candidate tests.  In k_persist a lane tests its own candidates
// one per iteration, so the test code runs for the lanes that happen to have
// one -- about a third of the wave.  Here each lane hands up to kK candidates
// of its current cell or leaf to the wave per iteration; the wave tests all of
// them at full width (item k of the wave's list on lane k mod 64), each item
// reading its ray from LDS, and every hit is folded into its ray's 64-bit key
// in LDS by an atomic minimum:
//   closest hit: key = (distance bits, candidate index) -- the smallest
//     distance, the first candidate in leaf order among equal distances,
//     which is what the sequential test with `dist < t` keeps;
//   any hit: key = (candidate index, distance bits) -- the first hit in
//     leaf order, where the sequential test stops.
// Candidates are consumed in leaf order, kK at a time, so every ray ends with
// the t and primitive of the one-ray-per-lane traversal.  The ray origins and
// directions live in LDS only (the lane keeps the DDA state), which pays for
// the registers of the batched test.  An any-hit ray may test up to kK - 1
// candidates past its first hit; the work statistics count the tests of the
// sequential traversal (up to and including that hit), as the reference's
// traversal would make them.
//
// A lane also takes up to `steps_max` grid steps per iteration when the leaves
// or cells it steps into are empty, so that more lanes reach candidates by the
// wave's test batch.  With kK = 12 the workgroup uses 24 KB of LDS, and six
// waves per SIMD (80 VGPRs per lane, a few values spilled) beat five waves
// without spills and 16 candidates per lane (AO 2.29 against 2.40 ms on the
// 1M-triangle frame): the gathers need the extra wave.
constexpr int kBatchK = 12;   // candidates a lane hands to its wave per iteration

template <int kK>
struct WaveTests {
    float ox[kWave], oy[kWave], oz[kWave];
    float dx[kWave], dy[kWave], dz[kWave];
    float t[kWave];
    unsigned long long key[kWave];
    uint32_t cand[kWave * kK];
    uint8_t own[kWave * kK];
};

// The wave's test batch over candidate references [it, end) of prims: each
// lane for which `want` holds and that has candidates left hands up to kK of
// them to the wave, the wave tests them all (each against its owner's ray in
// W) and every such lane takes its ray's result into t / best and moves it
// past the candidates.  Returns true on a lane whose any-hit ray was hit.
// Every lane of the wave calls it together.  kOwn: the lanes' rays belong to
// scenes with references and records of their own; an item reads its owner's
// from own_rec / own_prims (the wave's 64 entries in LDS), and rec / prims are
// the calling lane's.
template <bool kAny, int kK, bool kOwn = false>
__device__ __forceinline__ bool wave_test_range(const float4* __restrict__ rec, const uint32_t* __restrict__ prims,
                                                WaveTests<kK>& W, uint32_t& it, uint32_t end, float& t,
                                                uint32_t& best, uint32_t lane, bool want, Work& wk,
                                                const float4* const* own_rec = nullptr,
                                                const uint32_t* const* own_prims = nullptr)
{
    bool hit = false;
    uint32_t cnt = 0;
    if (want && it < end) cnt = end - it < (uint32_t)kK ? end - it : (uint32_t)kK;
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
    if (total) {
        const uint32_t ex = incl - cnt;
#pragma unroll
        for (int q = 0; q < kK; ++q) {
            if ((uint32_t)q < cnt) {
                W.cand[ex + q] = it + (uint32_t)q;
                W.own[ex + q] = (uint8_t)lane;
            }
        }
        if (cnt) {
            W.t[lane] = t;
            W.key[lane] = ~0ull;
        }
        wave_sync();
        for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
            const uint32_t k = b0 + lane;
            if (k < total) {
                const uint32_t o = W.own[k];
                const uint32_t ci = W.cand[k];
                f3 v0, v1, v2;
                if (kOwn) load_tri(own_rec[o], own_prims[o][ci], v0, v1, v2);
                else load_tri(rec, prims[ci], v0, v1, v2);
                const float dist = mt_dist(mk(W.ox[o], W.oy[o], W.oz[o]), mk(W.dx[o], W.dy[o], W.dz[o]),
                                           v0, v1, v2);
                if (dist < W.t[o]) {
                    const unsigned long long db = __float_as_uint(dist);
                    atomicMin(&W.key[o], kAny ? ((unsigned long long)ci << 32 | db) : (db << 32 | ci));
                }
            }
        }
        wave_sync();
        if (cnt) {
            const unsigned long long key = W.key[lane];
            uint32_t tested = cnt;
            if (key != ~0ull) {
                const uint32_t ci = kAny ? (uint32_t)(key >> 32) : (uint32_t)key;
                t = __uint_as_float(kAny ? (uint32_t)key : (uint32_t)(key >> 32));
                best = prims[ci];
                if (kAny) {
                    hit = true;
                    tested = ci - it + 1;   // the sequential traversal stops here
                }
            }
            wk.tests += tested;
            it += cnt;
        }
    }
    return hit;
}

template <bool kAny, int kK>
__device__ __forceinline__ bool wave_test_batch(const TraceScene& S, WaveTests<kK>& W, RayState& r, uint32_t lane,
                                                bool want, Work& wk)
{
    return wave_test_range<kAny, kK>(S.rec, S.prims, W, r.it, r.end, r.t, r.best, lane, want, wk);
}

// trace<> of one ray per active lane (RayTracingKernels.h:118-140: rcp = 1 / d,
// traverse, d' = 1 / rcp) with the wave's candidates tested in batches.  Every
// lane of the wave calls it together; an inactive lane takes part in the
// batches only.  kAny: t is both the start and the termination distance (the
// ray stops at its first hit below it).  Returns the hit primitive (S.nt on a
// miss) and d' in dir_out, as trace_ray does; steps_max grid steps per lane
// and batch.
template <bool kTwo, bool kAny, int kK>
__device__ __forceinline__ uint32_t trace_wave(const GridRegs& G, const TraceScene& S, WaveTests<kK>& W, bool active,
                                               f3 org, f3 d, float& t, f3& dir_out, uint32_t steps_max, Work& wk)
{
    const uint32_t lane = __lane_id();
    RayState r;
    bool live = false;
    if (active) {
        set_ray(r, org, d);
        r.t = t;
        r.tterm = t;
        live = ray_begin<kTwo, false>(G, S, r, wk);
        if (live) {
            W.ox[lane] = org.x;
            W.oy[lane] = org.y;
            W.oz[lane] = org.z;
            W.dx[lane] = r.dir.x;
            W.dy[lane] = r.dir.y;
            W.dz[lane] = r.dir.z;
        }
    }
    while (__ballot(live) != 0ull) {
        bool done = false;
        if (live && r.it >= r.end) {
            for (uint32_t m = 0; m < steps_max; ++m) {
                if (kTwo) {
                    bool top = true;
                    if (r.in_leaf) {
                        const int k = leaf_step<false>(S, r, wk);
                        if (k != 2) {
                            top = false;
                            done = k != 1;
                        }
                    }
                    if (top) done = !top_step(G, S, r, org, wk);
                } else {
                    done = !advance_uniform<false>(G, S, r, wk);
                }
                if (done || r.it < r.end) break;
            }
        }
        if (wave_test_batch<kAny>(S, W, r, lane, live && !done, wk)) done = true;
        if (live && done) live = false;
    }
    if (!active) return S.nt;
    t = r.t;
    dir_out = r.dir;
    return r.best;
}

}  // namespace trav
}  // namespace hr
