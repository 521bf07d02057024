// hr_hier.hip -- traversal and AO integrator of the instanced two-level grid
// hierarchy for gfx950 (its build is in hr_build.hip; hr_hier.h lists the
// reference behaviour followed).
//
// Traversal: one ray per lane.  The top grid is stepped as the reference's
// TLGridHierarchyTraverser does (entry EPS past the box, no clamp of the
// entry cell); for every instance of a cell the world ray is clipped to the
// instance box, moved into object coordinates and run through the object's
// uniform grid to its closest hit, which counts only inside the box interval.
// An any-hit ray (AO) stops after the first top cell that leaves it with a
// distance below 0.9999, as the reference's shadow rays do, and is occluded
// when its final distance is below tmax.
#include "hr_flat.h"
#include "hr_hier.h"
#include "hr_traverse.h"

namespace hr {

namespace {

constexpr int kBlock = 256;
using namespace trav;

inline uint32_t cdiv_u(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t device_cus()
{
    static int cached = 0;
    if (!cached) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cached = cus > 0 ? cus : 256;
    }
    return (uint32_t)cached;
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------

// GeometryInstanceMatrix::transformRay (TwoLevelGridHierarchy.h:650-665)
__device__ __forceinline__ void instance_ray(const InstanceDev& I, f3 org, f3 rcp, f3& org_t, f3& rcp_t)
{
    const f3 r0 = mk(I.irot0[0], I.irot0[1], I.irot0[2]);
    const f3 r1 = mk(I.irot1[0], I.irot1[1], I.irot1[2]);
    const f3 r2 = mk(I.irot2[0], I.irot2[1], I.irot2[2]);
    const f3 tr = mk(I.it[0], I.it[1], I.it[2]);
    org_t = r0 * org.x + r1 * org.y + r2 * org.z + tr;
    const f3 dt = r0 / rcp.x + r1 / rcp.y + r2 / rcp.z;
    rcp_t = mk(1.f / dt.x, 1.f / dt.y, 1.f / dt.z);
}

// transformNormalToGlobal (TwoLevelGridHierarchy.h:677-694): the transpose of
// the world-to-object rotation
__device__ __forceinline__ f3 instance_normal(const InstanceDev& I, f3 n)
{
    return mk(I.irot0[0], I.irot1[0], I.irot2[0]) * n.x + mk(I.irot0[1], I.irot1[1], I.irot2[1]) * n.y +
           mk(I.irot0[2], I.irot1[2], I.irot2[2]) * n.z;
}

// GeometryInstanceTraverser (TLGridHierarchyTraverser.h:45-231) for the
// instances of one top cell: clip the world ray to the instance box, move it
// into object coordinates, step the object grid (entry EPS past its box, the
// entry cell clamped on the side the ray leaves by), keep the closest hit
// that lies inside the instance box interval
//
// An instance that spans several top cells is met again in each of them.  A
// second visit cannot change the result: the first one already kept the
// closest hit of the whole object below the ray's distance at the time, the
// distance has only shrunk since, and the box tests that skip an instance
// fail again for a shorter ray.  So each ray remembers the instances it has
// visited in a small direct-mapped table (id & 7) and skips them.
struct Mailbox {
    uint32_t e[8];
};

__device__ __forceinline__ GridRegs inst_grid_regs(const InstGrid& q)
{
    GridRegs g;
    g.bmin = mk(q.bmin[0], q.bmin[1], q.bmin[2]);
    g.bmax = mk(q.bmax[0], q.bmax[1], q.bmax[2]);
    g.cs = mk(q.cs[0], q.cs[1], q.cs[2]);
    g.csr = mk(q.csr[0], q.csr[1], q.csr[2]);
    g.rx = q.res[0];
    g.ry = q.res[1];
    g.rz = q.res[2];
    return g;
}

__device__ __forceinline__ bool mailbox_seen(Mailbox& m, uint32_t id)
{
    const uint32_t k = id & 7u;
    uint32_t v = 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) v = j == k ? m.e[j] : v;
    if (v == id) return true;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) m.e[j] = j == k ? id : m.e[j];
    return false;
}

__device__ void hier_cell(const HierScene& H, uint32_t b, uint32_t e, f3 org, f3 rcp, float& t, uint32_t& best,
                          uint32_t& inst, Mailbox& mb, Work& wk)
{
    TraceScene S{};
    S.prims = H.prims;
    S.rec = H.rec;
    S.nt = H.nt;
    for (uint32_t it = b; it < e; ++it) {
        const uint32_t ii = H.inst_refs[it];
        if (mailbox_seen(mb, ii)) continue;
        const InstanceDev I = H.inst[ii];
        const InstGrid Q = H.inst_grid[ii];
        const f3 t1 = (mk(I.bmin[0], I.bmin[1], I.bmin[2]) - org) * rcp;
        f3 tm = (mk(I.bmax[0], I.bmax[1], I.bmax[2]) - org) * rcp;
        const f3 tn = vmin(t1, tm);
        tm = vmax(t1, tm);
        const float te1 = fmaxf(fmaxf(tn.x, tn.y), tn.z);
        const float tx1 = fminf(fminf(tm.x, tm.y), tm.z);
        if (!(tx1 > te1 && tx1 >= 0.f && te1 < t)) continue;
        f3 ot, rt;
        instance_ray(I, org, rcp, ot, rt);
        const f3 dt = mk(1.f / rt.x, 1.f / rt.y, 1.f / rt.z);
        const GridRegs G = inst_grid_regs(Q);
        const uint32_t base = Q.base;
        float rayT = t;
        uint32_t hit = H.nt;
        const f3 g1 = (G.bmin - ot) * rt;
        f3 gf = (G.bmax - ot) * rt;
        const f3 gn = vmin(g1, gf);
        gf = vmax(g1, gf);
        const float tentry = fmaxf(fmaxf(gn.x, gn.y), gn.z);
        const float texit = fminf(fminf(gf.x, gf.y), gf.z);
        bool flag = texit > tentry && texit >= 0.f && tentry < rayT;
        const f3 entry = (tentry >= 0.f) ? ot + rep(tentry + kEps) / rt : ot;
        f3 cf = (entry - G.bmin) * G.csr;
        cf = mk(floorf(cf.x), floorf(cf.y), floorf(cf.z));
        cf.x = rt.x > 0.f ? fminf((float)G.rx - 1.f, cf.x) : fmaxf(cf.x, 0.f);
        cf.y = rt.y > 0.f ? fminf((float)G.ry - 1.f, cf.y) : fmaxf(cf.y, 0.f);
        cf.z = rt.z > 0.f ? fminf((float)G.rz - 1.f, cf.z) : fmaxf(cf.z, 0.f);
        float tx = ((cf.x + (rt.x > 0.f ? 1.f : 0.f)) * G.cs.x + G.bmin.x - ot.x) * rt.x;
        float ty = ((cf.y + (rt.y > 0.f ? 1.f : 0.f)) * G.cs.y + G.bmin.y - ot.y) * rt.y;
        float tz = ((cf.z + (rt.z > 0.f ? 1.f : 0.f)) * G.cs.z + G.bmin.z - ot.z) * rt.z;
        int cx = to_int_sat(cf.x), cy = to_int_sat(cf.y), cz = to_int_sat(cf.z);
        const int sx = rt.x > 0.f ? 1 : -1, sy = rt.y > 0.f ? 1 : -1, sz = rt.z > 0.f ? 1 : -1;
        flag = flag && (sx > 0 ? cx < G.rx : cx > -1) && (sy > 0 ? cy < G.ry : cy > -1) &&
               (sz > 0 ? cz < G.rz : cz > -1);
        const float ddx = G.cs.x * fabsf(rt.x), ddy = G.cs.y * fabsf(rt.y), ddz = G.cs.z * fabsf(rt.z);
        const uint32_t sxy = (uint32_t)G.rx * (uint32_t)G.ry;
        while (flag) {
            ++wk.steps;
            if (cx < G.rx && cx > -1 && cy < G.ry && cy > -1 && cz < G.rz && cz > -1) {
                const uint2 cell =
                    H.obj_cells[base + (uint32_t)cx + (uint32_t)cy * (uint32_t)G.rx + (uint32_t)cz * sxy];
                isect_range<false, false>(S, cell.x, cell.y, ot, dt, rayT, hit, wk);
            }
            const int d = min_dim(tx, ty, tz);
            if (d == 0) {
                flag = rayT > tx;
                cx += sx; tx += ddx;
                flag = flag && (sx > 0 ? cx < G.rx : cx > -1);
            } else if (d == 1) {
                flag = rayT > ty;
                cy += sy; ty += ddy;
                flag = flag && (sy > 0 ? cy < G.ry : cy > -1);
            } else {
                flag = rayT > tz;
                cz += sz; tz += ddz;
                flag = flag && (sz > 0 ? cz < G.rz : cz > -1);
            }
        }
        if (hit != H.nt && rayT < t && rayT >= te1 && rayT < tx1) {
            t = rayT;
            best = hit;
            inst = ii;
        }
    }
}

// a shadow ray of the hierarchy stops after the top cell in which its
// distance fell below this (TLGridHierarchyTraverser.h:262-267, :366-372)
constexpr float kShadowStop = 0.9999f;

// TLGridHierarchyTraverser::operator() (TLGridHierarchyTraverser.h:235-400):
// the top grid is entered EPS past its box without clamping the entry cell;
// cells outside the grid read as empty.  A shadow ray stops after the top
// cell that leaves its distance below tterm: 0.9999 in the reference; an
// occlusion test against tmax >= 0.9999 may stop once the distance is below
// tmax, since the distance only shrinks and the ray is occluded either way.
template <bool kAny>
__device__ uint32_t trace_hier(const HierScene& H, f3 org, f3 d, float& t, float tterm, f3& dir_out, uint32_t& inst,
                               Work& wk)
{
    Mailbox mb;
#pragma unroll
    for (int j = 0; j < 8; ++j) mb.e[j] = 0xFFFFFFFFu;
    const f3 rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    dir_out = mk(1.f / rcp.x, 1.f / rcp.y, 1.f / rcp.z);
    uint32_t best = H.nt;
    inst = 0xFFFFFFFFu;
    const GridRegs G = load_grid(H.top);
    bool flag = !(kAny && t < tterm);
    const f3 t1 = (G.bmin - org) * rcp;
    f3 tfar = (G.bmax - org) * rcp;
    const f3 tnear = vmin(t1, tfar);
    tfar = vmax(t1, tfar);
    const float tentry = fmaxf(fmaxf(tnear.x, tnear.y), tnear.z);
    const float texit = fminf(fminf(tfar.x, tfar.y), tfar.z);
    flag = flag && texit > tentry && texit >= 0.f;
    const f3 entry = (tentry >= 0.f) ? org + rep(tentry + kEps) / rcp : org;
    f3 cf = (entry - G.bmin) * G.csr;
    cf = mk(floorf(cf.x), floorf(cf.y), floorf(cf.z));
    float tx = ((cf.x + (rcp.x > 0.f ? 1.f : 0.f)) * G.cs.x + G.bmin.x - org.x) * rcp.x;
    float ty = ((cf.y + (rcp.y > 0.f ? 1.f : 0.f)) * G.cs.y + G.bmin.y - org.y) * rcp.y;
    float tz = ((cf.z + (rcp.z > 0.f ? 1.f : 0.f)) * G.cs.z + G.bmin.z - org.z) * rcp.z;
    int cx = to_int_sat(cf.x), cy = to_int_sat(cf.y), cz = to_int_sat(cf.z);
    const int sx = rcp.x > 0.f ? 1 : -1, sy = rcp.y > 0.f ? 1 : -1, sz = rcp.z > 0.f ? 1 : -1;
    flag = flag && (sx > 0 ? cx < G.rx : cx > -1) && (sy > 0 ? cy < G.ry : cy > -1) &&
           (sz > 0 ? cz < G.rz : cz > -1);
    const float ddx = G.cs.x * fabsf(rcp.x), ddy = G.cs.y * fabsf(rcp.y), ddz = G.cs.z * fabsf(rcp.z);
    const uint32_t sxy = (uint32_t)G.rx * (uint32_t)G.ry;
    while (flag) {
        ++wk.steps;
        if (cx < G.rx && cx > -1 && cy < G.ry && cy > -1 && cz < G.rz && cz > -1) {
            const uint2 cell = H.top_cells[(uint32_t)cx + (uint32_t)cy * (uint32_t)G.rx + (uint32_t)cz * sxy];
            hier_cell(H, cell.x, cell.y, org, rcp, t, best, inst, mb, wk);
        }
        if (kAny && t < tterm) flag = false;
        const int dm = min_dim(tx, ty, tz);
        if (dm == 0) {
            flag = flag && t > tx;
            cx += sx; tx += ddx;
            flag = flag && (sx > 0 ? cx < G.rx : cx > -1);
        } else if (dm == 1) {
            flag = flag && t > ty;
            cy += sy; ty += ddy;
            flag = flag && (sy > 0 ? cy < G.ry : cy > -1);
        } else {
            flag = flag && t > tz;
            cz += sz; tz += ddz;
            flag = flag && (sz > 0 ? cz < G.rz : cz > -1);
        }
    }
    return best;
}

template <bool kAny>
__global__ __launch_bounds__(kBlock) void k_trace_hier(HierScene H, const float* __restrict__ org,
                                                       const float* __restrict__ dir, uint32_t n, float tmax,
                                                       float* __restrict__ t_out, uint32_t* __restrict__ prim_out,
                                                       uint32_t* __restrict__ inst_out,
                                                       uint8_t* __restrict__ occ_out, unsigned long long* ray_count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    count_rays(ray_count, i < n);
    if (i >= n) return;
    Work wk;
    const f3 o = mk(org[i], org[(size_t)n + i], org[2 * (size_t)n + i]);
    const f3 d = mk(dir[i], dir[(size_t)n + i], dir[2 * (size_t)n + i]);
    float t = kFltMax;
    f3 d2;
    uint32_t inst;
    const uint32_t best = trace_hier<kAny>(H, o, d, t, fmaxf(kShadowStop, tmax), d2, inst, wk);
    if (kAny) {
        occ_out[i] = t < tmax ? 1u : 0u;
        if (t_out) t_out[i] = t;
    } else {
        t_out[i] = t;
        prim_out[i] = t < kFltMax ? best : H.nt;
        if (inst_out) inst_out[i] = t < kFltMax ? inst : 0xFFFFFFFFu;
    }
}

// ---------------------------------------------------------------------------
// TLGHAOIntegrator
// ---------------------------------------------------------------------------

// the camera ray of slot q of the 8x8 pixel tiles (a wave's rays enter the
// same instances); false for a slot past the image's edge
__device__ __forceinline__ bool hier_camera_ray(const HierScene& H, const CameraDev& cam, const RenderArgs& ra,
                                                uint32_t npix, uint32_t q, float* __restrict__ hits,
                                                uint32_t* __restrict__ hit_prims, uint32_t* __restrict__ hit_inst,
                                                unsigned long long* ray_count)
{
    uint32_t pix;
    const bool active = tiled_pixel(q, (uint32_t)cam.res_x, (uint32_t)cam.res_y, pix);
    count_rays(ray_count, active);
    if (!active) return false;
    Work wk;
    const f3 d0 = primary_direction(cam, ra, pix, npix);
    const f3 org = mk(cam.center[0], cam.center[1], cam.center[2]);
    float t = kFltMax;
    f3 dir;
    uint32_t inst;
    const uint32_t best = trace_hier<false>(H, org, d0, t, 0.f, dir, inst, wk);
    const f3 hp = org + (t - kEps) * dir;
    hits[pix] = hp.x;
    hits[(size_t)npix + pix] = hp.y;
    hits[2 * (size_t)npix + pix] = hp.z;
    hits[3 * (size_t)npix + pix] = dir.x;
    hits[4 * (size_t)npix + pix] = dir.y;
    hits[5 * (size_t)npix + pix] = dir.z;
    hits[6 * (size_t)npix + pix] = t;
    hit_prims[pix] = t < kFltMax ? best : H.nt;
    hit_inst[pix] = t < kFltMax ? inst : 0xFFFFFFFFu;
    return true;
}

// camera rays one per lane.  tiles == 0: workgroup b takes slots 256 b ..;
// otherwise a few workgroups per CU whose waves take 8x8 tiles from the
// counter *next (zero at launch) until all `tiles` are taken.  Six waves per
// SIMD (at most 80 VGPRs): render 32.1 -> 31.6 ms on 10,000 instances
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_primary_hier(
    HierScene H, CameraDev cam, RenderArgs ra, uint32_t npix, float* __restrict__ hits,
    uint32_t* __restrict__ hit_prims, uint32_t* __restrict__ hit_inst, unsigned long long* ray_count,
    uint32_t tiles, uint32_t* __restrict__ next)
{
    if (tiles == 0) {
        (void)hier_camera_ray(H, cam, ra, npix, blockIdx.x * blockDim.x + threadIdx.x, hits, hit_prims, hit_inst,
                              ray_count);
        return;
    }
    while (true) {
        uint32_t tile = 0;
        if (__lane_id() == 0) tile = atomicAdd(next, 1u);
        tile = __shfl(tile, 0, kWave);
        if (tile >= tiles) break;
        (void)hier_camera_ray(H, cam, ra, npix, tile * kWave + __lane_id(), hits, hit_prims, hit_inst, ray_count);
    }
}

// the hit's face normal in world coordinates (TLGridHierarchyAOIntegrator.h:316-357)
__device__ __forceinline__ f3 hier_normal(const HierScene& H, uint32_t prim, uint32_t inst)
{
    f3 v0, v1, v2;
    load_tri(H.rec, prim, v0, v1, v2);
    f3 n = normalize(cross(v1 - v0, v2 - v0));
    if (inst != 0xFFFFFFFFu) n = instance_normal(H.inst[inst], n);
    return n;
}

// AmbientOcclusionRayGenerator of the hierarchy (TLGridHierarchyAOIntegrator.h:296-380)
__global__ __launch_bounds__(kBlock) void k_ao_rays_hier(HierScene H, RenderArgs ra, uint32_t npix,
                                                         const float* __restrict__ hits,
                                                         const uint32_t* __restrict__ hit_prims,
                                                         const uint32_t* __restrict__ hit_inst,
                                                         uint8_t* __restrict__ occ, unsigned long long* ray_count)
{
    const uint32_t S_ = (uint32_t)ra.ao_samples;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = j < npix * S_;
    const uint32_t pix = in ? j / S_ : 0u;
    const uint32_t s = in ? j % S_ : 0u;
    const uint32_t prim = in ? hit_prims[pix] : H.nt;
    const float th = in ? hits[6 * (size_t)npix + pix] : kFltMax;
    const bool active = in && th < kFltMax && prim < H.nt;
    count_rays(ray_count, active);
    if (!in) return;
    if (!active) {
        occ[j] = 1u;
        return;
    }
    Work wk;
    const f3 pdir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
    f3 org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
    f3 n = hier_normal(H, prim, hit_inst[pix]);
    if (dot(n, pdir) > 0.f) n = -n;
    float u[4];
    rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamAO + s, u);
    const f3 h = cosine_dir(u[0], u[1]);
    f3 tg, bn;
    local_frame(n, tg, bn);
    const f3 dir = normalize(n * h.z + tg * h.x + bn * h.y);
    org = org + (0.0001f * ra.ao_dist) * n;
    float t = kFltMax;
    f3 d2;
    uint32_t inst;
    trace_hier<true>(H, org, dir, t, fmaxf(kShadowStop, ra.ao_dist), d2, inst, wk);
    occ[j] = t < ra.ao_dist ? 1u : 0u;
}

// computeAOIlluminationTLGH (TLGridHierarchyAOIntegrator.h:64-204)
__global__ __launch_bounds__(kBlock) void k_ao_shade_hier(HierScene H, RenderArgs ra, uint32_t npix,
                                                          const float* __restrict__ hits,
                                                          const uint32_t* __restrict__ hit_prims,
                                                          const uint32_t* __restrict__ hit_inst,
                                                          const uint8_t* __restrict__ occ, Shading sh,
                                                          float* __restrict__ fb)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const uint32_t S_ = (uint32_t)ra.ao_samples;
    const float th = hits[6 * (size_t)npix + pix];
    const uint32_t prim = hit_prims[pix];
    const bool hit = th < kFltMax;
    f3 n = rep(0.f), kd = rep(0.f), pdir = rep(0.f);
    if (hit && prim < H.nt) {
        n = hier_normal(H, prim, hit_inst[pix]);
        const float* m = sh.mats + 12 * (size_t)(sh.mat_idx ? sh.mat_idx[prim] : 0u);
        kd = mk(m[0], m[1], m[2]) * 3.14159f;
        pdir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
    }
    f3 rad = rep(0.f);
    for (uint32_t s = 0; s < S_; ++s) {
        f3 v;
        if (hit) {
            v = occ[(size_t)pix * S_ + s] ? rep(0.f) : rep(1.f);
            if (v.x + v.y + v.z > 0.f && prim < H.nt) {
                v.x *= kd.x;
                v.y *= kd.y;
                v.z *= kd.z;
                const float sin_zn = fmaxf(0.f, sqrtf(1.f - n.z));
                const float cos_en = fabsf(dot(pdir, n));
                v.x *= (0.7f * cos_en + 0.3f * sin_zn);
                v.y *= (0.7f * cos_en + 0.2f * sin_zn);
                v.z *= (0.9f * cos_en);
            }
        } else {
            v = rep(1.f);
        }
        rad = rad + v * 1.f / (float)S_;
    }
    const float nw = 1.f / (float)(image_of(ra) + 1);
    const float ow = 1.f - nw;
    float* o = fb + 3 * (size_t)pix;
    o[0] = o[0] * ow + rad.x * nw;
    o[1] = o[1] * ow + rad.y * nw;
    o[2] = o[2] * ow + rad.z * nw;
}

// ---------------------------------------------------------------------------
// persistent AO rays: the traversal of trace_hier<true> as a flat state
// machine -- one unit per loop iteration (one triangle test, one object-grid
// step, one instance set-up, one top-grid step) -- with waves refilling idle
// lanes from a global counter, as the uniform / two-level persistent kernels
// do.  The units are those of trace_hier in its order, so every ray's
// occlusion bit is the one-ray-per-lane kernel's.
// ---------------------------------------------------------------------------

struct HierRay {
    f3 org, rcp;
    float t, tterm;
    float tstop;                 // an accepted hit closer than this finishes the ray
    uint32_t best, inst;
    Mailbox mb;
    int cx, cy, cz;              // top grid
    float tx, ty, tz;
    uint32_t it, end;            // instance references of the current top cell
    bool in_inst;                // an instance's object grid is being stepped
    uint32_t ii, base, hit;
    float te1, tx1, rayT;
    f3 ot, rt, dt;               // object-space origin, reciprocal direction, direction
    f3 ocs;                      // object cell size
    int rx, ry, rz;
    int ocx, ocy, ocz;
    float otx, oty, otz;
    uint32_t pit, pend;          // triangle references of the current object cell
};

__device__ __forceinline__ void hp_top_cell(const HierScene& H, const GridRegs& G, HierRay& r)
{
    r.it = r.end = 0u;
    if (r.cx < G.rx && r.cx > -1 && r.cy < G.ry && r.cy > -1 && r.cz < G.rz && r.cz > -1) {
        const uint2 cell =
            H.top_cells[(uint32_t)r.cx + (uint32_t)r.cy * (uint32_t)G.rx + (uint32_t)r.cz * ((uint32_t)G.rx * (uint32_t)G.ry)];
        r.it = cell.x;
        r.end = cell.y;
    }
}

__device__ __forceinline__ void hp_obj_cell(const HierScene& H, HierRay& r)
{
    r.pit = r.pend = 0u;
    if (r.ocx < r.rx && r.ocx > -1 && r.ocy < r.ry && r.ocy > -1 && r.ocz < r.rz && r.ocz > -1) {
        const uint2 cell = H.obj_cells[r.base + (uint32_t)r.ocx + (uint32_t)r.ocy * (uint32_t)r.rx +
                                       (uint32_t)r.ocz * ((uint32_t)r.rx * (uint32_t)r.ry)];
        r.pit = cell.x;
        r.pend = cell.y;
    }
}

// set up the top-grid DDA of a generated ray; false when it does not start
__device__ __forceinline__ bool hp_begin(const HierScene& H, const GridRegs& G, HierRay& r)
{
    r.best = H.nt;
    r.inst = 0xFFFFFFFFu;
    r.in_inst = false;
    r.pit = r.pend = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.mb.e[j] = 0xFFFFFFFFu;
    bool flag = !(r.t < r.tterm);
    const f3 t1 = (G.bmin - r.org) * r.rcp;
    f3 tfar = (G.bmax - r.org) * r.rcp;
    const f3 tnear = vmin(t1, tfar);
    tfar = vmax(t1, tfar);
    const float tentry = fmaxf(fmaxf(tnear.x, tnear.y), tnear.z);
    const float texit = fminf(fminf(tfar.x, tfar.y), tfar.z);
    flag = flag && texit > tentry && texit >= 0.f;
    const f3 entry = (tentry >= 0.f) ? r.org + rep(tentry + kEps) / r.rcp : r.org;
    f3 cf = (entry - G.bmin) * G.csr;
    cf = mk(floorf(cf.x), floorf(cf.y), floorf(cf.z));
    r.tx = ((cf.x + (r.rcp.x > 0.f ? 1.f : 0.f)) * G.cs.x + G.bmin.x - r.org.x) * r.rcp.x;
    r.ty = ((cf.y + (r.rcp.y > 0.f ? 1.f : 0.f)) * G.cs.y + G.bmin.y - r.org.y) * r.rcp.y;
    r.tz = ((cf.z + (r.rcp.z > 0.f ? 1.f : 0.f)) * G.cs.z + G.bmin.z - r.org.z) * r.rcp.z;
    r.cx = to_int_sat(cf.x);
    r.cy = to_int_sat(cf.y);
    r.cz = to_int_sat(cf.z);
    flag = flag && (r.rcp.x > 0.f ? r.cx < G.rx : r.cx > -1) && (r.rcp.y > 0.f ? r.cy < G.ry : r.cy > -1) &&
           (r.rcp.z > 0.f ? r.cz < G.rz : r.cz > -1);
    if (!flag) return false;
    hp_top_cell(H, G, r);
    return true;
}

// one unit of work other than a triangle test (an object-grid step, an
// instance set-up or a top-grid step); true when the ray is finished.
// entered: an instance's object grid was entered (r.ot / r.dt set)
__device__ __forceinline__ bool hp_unit(const HierScene& H, const GridRegs& G, HierRay& r, bool& entered)
{
    if (r.in_inst) {
        // the object DDA's step after a cell
        const int d = min_dim(r.otx, r.oty, r.otz);
        bool flag;
        if (d == 0) {
            flag = r.rayT > r.otx;
            r.ocx += r.rt.x > 0.f ? 1 : -1;
            r.otx += r.ocs.x * fabsf(r.rt.x);
            flag = flag && (r.rt.x > 0.f ? r.ocx < r.rx : r.ocx > -1);
        } else if (d == 1) {
            flag = r.rayT > r.oty;
            r.ocy += r.rt.y > 0.f ? 1 : -1;
            r.oty += r.ocs.y * fabsf(r.rt.y);
            flag = flag && (r.rt.y > 0.f ? r.ocy < r.ry : r.ocy > -1);
        } else {
            flag = r.rayT > r.otz;
            r.ocz += r.rt.z > 0.f ? 1 : -1;
            r.otz += r.ocs.z * fabsf(r.rt.z);
            flag = flag && (r.rt.z > 0.f ? r.ocz < r.rz : r.ocz > -1);
        }
        if (flag) {
            hp_obj_cell(H, r);
            return false;
        }
        r.in_inst = false;
        if (r.hit != H.nt && r.rayT < r.t && r.rayT >= r.te1 && r.rayT < r.tx1) {
            r.t = r.rayT;
            r.best = r.hit;
            r.inst = r.ii;
            // the distance only shrinks from here: an occlusion test against
            // tstop is decided
            if (r.t < r.tstop) return true;
        }
        return false;
    }
    // the next instance of the top cell that the ray meets
    while (r.it < r.end) {
        const uint32_t ii = H.inst_refs[r.it];
        ++r.it;
        if (mailbox_seen(r.mb, ii)) continue;
        const InstanceDev I = H.inst[ii];
        const InstGrid Q = H.inst_grid[ii];
        const f3 b1 = (mk(I.bmin[0], I.bmin[1], I.bmin[2]) - r.org) * r.rcp;
        f3 bm = (mk(I.bmax[0], I.bmax[1], I.bmax[2]) - r.org) * r.rcp;
        const f3 bn = vmin(b1, bm);
        bm = vmax(b1, bm);
        const float te1 = fmaxf(fmaxf(bn.x, bn.y), bn.z);
        const float tx1 = fminf(fminf(bm.x, bm.y), bm.z);
        if (!(tx1 > te1 && tx1 >= 0.f && te1 < r.t)) continue;
        f3 ot, rt;
        instance_ray(I, r.org, r.rcp, ot, rt);
        const GridRegs O = inst_grid_regs(Q);
        const f3 g1 = (O.bmin - ot) * rt;
        f3 gf = (O.bmax - ot) * rt;
        const f3 gn = vmin(g1, gf);
        gf = vmax(g1, gf);
        const float tentry = fmaxf(fmaxf(gn.x, gn.y), gn.z);
        const float texit = fminf(fminf(gf.x, gf.y), gf.z);
        bool flag = texit > tentry && texit >= 0.f && tentry < r.t;
        const f3 entry = (tentry >= 0.f) ? ot + rep(tentry + kEps) / rt : ot;
        f3 cf = (entry - O.bmin) * O.csr;
        cf = mk(floorf(cf.x), floorf(cf.y), floorf(cf.z));
        cf.x = rt.x > 0.f ? fminf((float)O.rx - 1.f, cf.x) : fmaxf(cf.x, 0.f);
        cf.y = rt.y > 0.f ? fminf((float)O.ry - 1.f, cf.y) : fmaxf(cf.y, 0.f);
        cf.z = rt.z > 0.f ? fminf((float)O.rz - 1.f, cf.z) : fmaxf(cf.z, 0.f);
        const int ocx = to_int_sat(cf.x), ocy = to_int_sat(cf.y), ocz = to_int_sat(cf.z);
        flag = flag && (rt.x > 0.f ? ocx < O.rx : ocx > -1) && (rt.y > 0.f ? ocy < O.ry : ocy > -1) &&
               (rt.z > 0.f ? ocz < O.rz : ocz > -1);
        if (!flag) continue;          // no object cell is stepped: no hit from this instance
        r.ii = ii;
        r.te1 = te1;
        r.tx1 = tx1;
        r.ot = ot;
        r.rt = rt;
        r.dt = mk(1.f / rt.x, 1.f / rt.y, 1.f / rt.z);
        r.ocs = O.cs;
        r.rx = O.rx;
        r.ry = O.ry;
        r.rz = O.rz;
        r.base = Q.base;
        r.otx = ((cf.x + (rt.x > 0.f ? 1.f : 0.f)) * O.cs.x + O.bmin.x - ot.x) * rt.x;
        r.oty = ((cf.y + (rt.y > 0.f ? 1.f : 0.f)) * O.cs.y + O.bmin.y - ot.y) * rt.y;
        r.otz = ((cf.z + (rt.z > 0.f ? 1.f : 0.f)) * O.cs.z + O.bmin.z - ot.z) * rt.z;
        r.ocx = ocx;
        r.ocy = ocy;
        r.ocz = ocz;
        r.rayT = r.t;
        r.hit = H.nt;
        r.in_inst = true;
        entered = true;
        hp_obj_cell(H, r);
        return false;
    }
    // the top cell is done: the shadow stop, then the top DDA's step
    if (r.t < r.tterm) return true;
    const int dm = min_dim(r.tx, r.ty, r.tz);
    bool flag;
    if (dm == 0) {
        flag = r.t > r.tx;
        r.cx += r.rcp.x > 0.f ? 1 : -1;
        r.tx += G.cs.x * fabsf(r.rcp.x);
        flag = flag && (r.rcp.x > 0.f ? r.cx < G.rx : r.cx > -1);
    } else if (dm == 1) {
        flag = r.t > r.ty;
        r.cy += r.rcp.y > 0.f ? 1 : -1;
        r.ty += G.cs.y * fabsf(r.rcp.y);
        flag = flag && (r.rcp.y > 0.f ? r.cy < G.ry : r.cy > -1);
    } else {
        flag = r.t > r.tz;
        r.cz += r.rcp.z > 0.f ? 1 : -1;
        r.tz += G.cs.z * fabsf(r.rcp.z);
        flag = flag && (r.rcp.z > 0.f ? r.cz < G.rz : r.cz > -1);
    }
    if (!flag) return true;
    hp_top_cell(H, G, r);
    return false;
}

// one unit of work; true when the ray is finished
__device__ __forceinline__ bool hp_step(const HierScene& H, const GridRegs& G, HierRay& r)
{
    if (r.pit < r.pend) {
        // one triangle of the object cell (MollerTrumboreIntersectionTest)
        const uint32_t p = H.prims[r.pit];
        ++r.pit;
        f3 v0, v1, v2;
        load_tri(H.rec, p, v0, v1, v2);
        const float dist = mt_dist(r.ot, r.dt, v0, v1, v2);
        if (dist < r.rayT) {
            r.rayT = dist;
            r.hit = p;
        }
        return false;
    }
    bool entered = false;
    return hp_unit(H, G, r, entered);
}

// AmbientOcclusionRayGenerator of the hierarchy (TLGridHierarchyAOIntegrator.h:296-380)
// for the persistent kernel: ray j = sample j % S of pixel j / S
struct GenAOHier {
    RenderArgs ra;
    uint32_t npix;
    const float* hits;
    const uint32_t* hit_prims;
    const uint32_t* hit_inst;
    uint8_t* occ;

    __device__ __forceinline__ uint32_t count() const { return npix * (uint32_t)ra.ao_samples; }

    // false: ray j has no traversal (its result is written); *traced: it counts as a ray
    __device__ __forceinline__ bool gen(const HierScene& H, const GridRegs& G, uint32_t j, HierRay& r,
                                        bool& traced) const
    {
        const uint32_t S_ = (uint32_t)ra.ao_samples;
        const uint32_t pix = j / S_, s = j % S_;
        const uint32_t prim = hit_prims[pix];
        const float th = hits[6 * (size_t)npix + pix];
        traced = th < kFltMax && prim < H.nt;
        if (!traced) {
            occ[j] = 1u;
            return false;
        }
        const f3 pdir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
        f3 o = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        f3 n = hier_normal(H, prim, hit_inst[pix]);
        if (dot(n, pdir) > 0.f) n = -n;
        float u[4];
        rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamAO + s, u);
        const f3 h = cosine_dir(u[0], u[1]);
        f3 tg, bn;
        local_frame(n, tg, bn);
        const f3 dir = normalize(n * h.z + tg * h.x + bn * h.y);
        r.org = o + (0.0001f * ra.ao_dist) * n;
        r.rcp = mk(1.f / dir.x, 1.f / dir.y, 1.f / dir.z);
        r.t = kFltMax;
        r.tterm = fmaxf(kShadowStop, ra.ao_dist);
        r.tstop = ra.ao_dist;
        if (hp_begin(H, G, r)) return true;
        finish(H, j, r);
        return false;
    }

    __device__ __forceinline__ void finish(const HierScene&, uint32_t j, const HierRay& r) const
    {
        occ[j] = r.t < ra.ao_dist ? 1u : 0u;
    }
};

// camera rays of the hierarchy (k_primary_hier's outputs)
struct GenPrimaryHier {
    CameraDev cam;
    RenderArgs ra;
    uint32_t npix;
    float* hits;
    uint32_t* hit_prims;
    uint32_t* hit_inst;

    __device__ __forceinline__ uint32_t count() const { return npix; }

    __device__ __forceinline__ bool gen(const HierScene& H, const GridRegs& G, uint32_t j, HierRay& r,
                                        bool& traced) const
    {
        traced = true;
        const f3 d0 = primary_direction(cam, ra, j, npix);
        r.org = mk(cam.center[0], cam.center[1], cam.center[2]);
        r.rcp = mk(1.f / d0.x, 1.f / d0.y, 1.f / d0.z);
        r.t = kFltMax;
        r.tterm = 0.f;
        r.tstop = -1.f;
        if (hp_begin(H, G, r)) return true;
        finish(H, j, r);
        return false;
    }

    __device__ __forceinline__ void finish(const HierScene& H, uint32_t pix, const HierRay& r) const
    {
        const f3 dir = mk(1.f / r.rcp.x, 1.f / r.rcp.y, 1.f / r.rcp.z);
        const f3 hp = r.org + (r.t - kEps) * dir;
        hits[pix] = hp.x;
        hits[(size_t)npix + pix] = hp.y;
        hits[2 * (size_t)npix + pix] = hp.z;
        hits[3 * (size_t)npix + pix] = dir.x;
        hits[4 * (size_t)npix + pix] = dir.y;
        hits[5 * (size_t)npix + pix] = dir.z;
        hits[6 * (size_t)npix + pix] = r.t;
        hit_prims[pix] = r.t < kFltMax ? r.best : H.nt;
        hit_inst[pix] = r.t < kFltMax ? r.inst : 0xFFFFFFFFu;
    }
};

// Waves fetch rays from *pool when `refill` of their lanes are idle (or all
// are); every wave leaves the loop once the pool is exhausted and its lanes
// are idle, so the grid drains.
template <class Gen>
__global__ __launch_bounds__(kBlock) void k_hier_persist(HierScene H, Gen g, uint32_t* __restrict__ pool,
                                                         uint32_t refill, unsigned long long* ray_count)
{
    const GridRegs G = load_grid(H.top);
    const uint64_t lt = __lanemask_lt();
    const uint32_t nrays = g.count();
    HierRay r;
    uint32_t j = 0, started = 0;
    bool live = false, exhausted = false;
    while (true) {
        const uint64_t idle = __ballot(!live);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && (nidle >= refill || idle == ~0ull)) {
            uint32_t base = 0;
            if (__lane_id() == 0) base = atomicAdd(pool, nidle);
            base = __shfl(base, 0, kWave);
            if (base + nidle >= nrays) exhausted = true;
            if (!live) {
                j = base + (uint32_t)__popcll(idle & lt);
                if (j < nrays) {
                    bool traced = false;
                    live = g.gen(H, G, j, r, traced);
                    started += traced ? 1u : 0u;
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (live && hp_step(H, G, r)) {
            live = false;
            g.finish(H, j, r);
        }
    }
    if (ray_count) {
        unsigned long long c = started;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (__lane_id() == 0 && c) atomicAdd(ray_count, c);
    }
}

// k_hier_persist with the triangle tests batched across the wave (the
// candidate references of each lane's current object cell, up to kBatchK per
// lane and iteration, tested at full width against the lanes' object-space
// rays kept in LDS; see wave_test_range).  A lane whose cell has no candidate
// left takes one other unit per iteration.  Every ray's units and their order
// are trace_hier's, so the outputs are the same.  Four waves per SIMD (at most
// 128 VGPRs): render 25.6 -> 24.6 ms against the compiler's 144 VGPRs and three
// waves, 31.6 ms with per-lane tests, on 10,000 instances at 1024^2 x 4 AO.
template <class Gen>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_hier_persist_bt(HierScene H, Gen g, uint32_t* __restrict__ pool,
                                                            uint32_t refill, unsigned long long* ray_count)
{
    __shared__ WaveTests<kBatchK> lds[kBlock / kWave];
    WaveTests<kBatchK>& W = lds[__builtin_amdgcn_readfirstlane((int)threadIdx.x) / kWave];
    const GridRegs G = load_grid(H.top);
    const uint32_t lane = __lane_id();
    const uint64_t lt = __lanemask_lt();
    const uint32_t nrays = g.count();
    HierRay r;
    Work wk;
    uint32_t j = 0, started = 0;
    bool live = false, exhausted = false;
    while (true) {
        const uint64_t idle = __ballot(!live);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && (nidle >= refill || idle == ~0ull)) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(pool, nidle);
            base = __shfl(base, 0, kWave);
            if (base + nidle >= nrays) exhausted = true;
            if (!live) {
                j = base + (uint32_t)__popcll(idle & lt);
                if (j < nrays) {
                    bool traced = false;
                    live = g.gen(H, G, j, r, traced);
                    started += traced ? 1u : 0u;
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (exhausted) break;
            continue;
        }
        bool done = false;
        if (live && r.pit >= r.pend) {
            bool entered = false;
            done = hp_unit(H, G, r, entered);
            if (entered) {
                W.ox[lane] = r.ot.x;
                W.oy[lane] = r.ot.y;
                W.oz[lane] = r.ot.z;
                W.dx[lane] = r.dt.x;
                W.dy[lane] = r.dt.y;
                W.dz[lane] = r.dt.z;
            }
        }
        wave_test_range<false, kBatchK>(H.rec, H.prims, W, r.pit, r.pend, r.rayT, r.hit, lane, live && !done, wk);
        if (live && done) {
            live = false;
            g.finish(H, j, r);
        }
    }
    if (ray_count) {
        unsigned long long c = started;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (lane == 0 && c) atomicAdd(ray_count, c);
    }
}

template <class K>
uint32_t hier_blocks(K kernel)
{
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlock, 0);
    return (uint32_t)((cus > 0 ? cus : 256) * (per > 0 ? per : 1));
}

template <class Gen>
void launch_hier_persist(const HierScene& H, const Gen& g, uint32_t nrays, uint32_t* pool, uint32_t refill,
                         unsigned long long* ray_count, hipStream_t s, bool batch)
{
    static uint32_t blocks_max[2] = { 0, 0 };
    if (!blocks_max[batch]) blocks_max[batch] = batch ? hier_blocks(k_hier_persist_bt<Gen>) : hier_blocks(k_hier_persist<Gen>);
    uint32_t blocks = cdiv_u(nrays, kBlock);
    if (blocks > blocks_max[batch]) blocks = blocks_max[batch];
    if (!blocks) return;
    fill_u32(pool, 0u, 1, nullptr, 1, s);
    if (batch)
        hipLaunchKernelGGL(k_hier_persist_bt<Gen>, dim3(blocks), dim3(kBlock), 0, s, H, g, pool, refill, ray_count);
    else
        hipLaunchKernelGGL(k_hier_persist<Gen>, dim3(blocks), dim3(kBlock), 0, s, H, g, pool, refill, ray_count);
}

}  // namespace

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

void launch_trace_hier(const HierScene& H, const float* org, const float* dir, uint32_t n, int anyhit, float tmax,
                       float* t_out, uint32_t* prim_out, uint32_t* inst_out, uint8_t* occ_out,
                       unsigned long long* ray_count, hipStream_t s)
{
    if (n == 0) return;
    const dim3 g(cdiv_u(n, kBlock)), b(kBlock);
    if (anyhit)
        hipLaunchKernelGGL(k_trace_hier<true>, g, b, 0, s, H, org, dir, n, tmax, t_out, prim_out, inst_out, occ_out,
                           ray_count);
    else
        hipLaunchKernelGGL(k_trace_hier<false>, g, b, 0, s, H, org, dir, n, tmax, t_out, prim_out, inst_out, occ_out,
                           ray_count);
}

void launch_render_ao_hier(const HierScene& H, const CameraDev& cam, const RenderArgs& ra, float* hits,
                           uint32_t* hit_prims, uint32_t* hit_inst, uint8_t* occ, const Shading& sh, float* fb,
                           bool pers_primary, bool pers_ao, uint32_t* pool, uint32_t refill,
                           unsigned long long* ray_count, hipStream_t s, bool batch, hipEvent_t after_primary,
                           hipEvent_t after_ao, uint32_t camera_blocks)
{
    const uint32_t npix = (uint32_t)cam.res_x * (uint32_t)cam.res_y;
    if (pers_primary) {
        const GenPrimaryHier g{ cam, ra, npix, hits, hit_prims, hit_inst };
        launch_hier_persist(H, g, npix, pool, refill, ray_count, s, batch);
    } else {
        const uint32_t slots = tiled_slots((uint32_t)cam.res_x, (uint32_t)cam.res_y);
        uint32_t blocks = cdiv_u(slots, kBlock), tiles = 0;
        if (camera_blocks > 0 && (size_t)camera_blocks * device_cus() < blocks) {
            blocks = camera_blocks * device_cus();
            tiles = slots / kWave;
            fill_u32(pool, 0u, 1, nullptr, 1, s);
        }
        hipLaunchKernelGGL(k_primary_hier, dim3(blocks), dim3(kBlock), 0, s, H, cam, ra, npix, hits, hit_prims,
                           hit_inst, ray_count, tiles, pool);
    }
    if (after_primary) (void)hipEventRecord(after_primary, s);
    const size_t nrays = (size_t)npix * (size_t)ra.ao_samples;
    if (pers_ao) {
        const GenAOHier g{ ra, npix, hits, hit_prims, hit_inst, occ };
        launch_hier_persist(H, g, (uint32_t)nrays, pool, refill, ray_count, s, batch);
    } else {
        hipLaunchKernelGGL(k_ao_rays_hier, dim3(cdiv_u(nrays, kBlock)), dim3(kBlock), 0, s, H, ra, npix, hits,
                           hit_prims, hit_inst, occ, ray_count);
    }
    if (after_ao) (void)hipEventRecord(after_ao, s);
    hipLaunchKernelGGL(k_ao_shade_hier, dim3(cdiv_u(npix, kBlock)), dim3(kBlock), 0, s, H, ra, npix, hits, hit_prims,
                       hit_inst, occ, sh, fb);
}

}  // namespace hr
