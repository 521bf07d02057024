// hr_trace.hip -- 3D-DDA traversal of uniform and two-level grids with
// Moller-Trumbore intersection, primary rays, ambient occlusion and the
// path-tracing megakernel, for gfx950.
//
// Reference behaviour followed:
//   MollerTrumboreIntersectionTest   src/RT/Algorithm/RayTriangleIntersector.h:36-92
//   UGridTraverser                   src/RT/Algorithm/UGridTraverser.h:44-177
//   TLGridTraverser                  src/RT/Algorithm/TLGridTraverser.h:45-354
//   trace<>                          src/RT/Algorithm/RayTracingKernels.h:47-152
//   SimpleRayBuffer::store           src/RT/Structure/RayBuffers.h:62-87
//   Camera / pixel samplers          src/RT/Primitive/Camera.h:42-132
//   RandomPrimaryRayGenerator        src/RT/Algorithm/RayGenerators.h:75-117
//   AmbientOcclusionRayGenerator     src/RT/Algorithm/RayGenerators.h:273-344
//   computeAOIllumination            src/RT/Integrator/AOIntegrator.h:64-191
//   pathTrace                        src/RT/Integrator/PathTracer.h:41-372
//
// Departures, shared with the CPU oracle: the counter-based RNG keyed by
// (pixel, image, stream); any-hit rays start at t = tmax and stop at the
// first accepted hit (occluded iff a hit lies in (1e-4, tmax)); paths stop
// after pt_depth vertices; hit candidates store the primitive id directly
// (the reference stores the position in primitives[] and maps it after the
// traversal: the same result).
//
// Kernel shape: one ray per lane, 256-lane workgroups; the grid parameters
// are loaded once per lane from the device-resident GridParams block that
// the builder wrote, so a trace needs no host knowledge of the grid.
#include "hr_trace.h"

namespace hr {

namespace {

constexpr int kBlock = 256;

struct GridRegs {
    f3 bmin, bmax, cs, csr;
    int rx, ry, rz;
};

__device__ __forceinline__ GridRegs load_grid(const GridParams* gp)
{
    GridRegs g;
    g.bmin = mk(gp->bmin[0], gp->bmin[1], gp->bmin[2]);
    g.bmax = mk(gp->bmax[0], gp->bmax[1], gp->bmax[2]);
    g.cs = mk(gp->cs[0], gp->cs[1], gp->cs[2]);
    g.csr = mk(gp->csr[0], gp->csr[1], gp->csr[2]);
    g.rx = gp->res[0];
    g.ry = gp->res[1];
    g.rz = gp->res[2];
    return g;
}

// MIN_DIMENSION (UGridTraverser.h:40-41)
__device__ __forceinline__ int min_dim(float x, float y, float z)
{
    return (x < y) ? ((x < z) ? 0 : 2) : ((y < z) ? 1 : 2);
}

struct Work {
    uint32_t steps = 0, tests = 0;
};

template <bool kAny>
__device__ __forceinline__ bool isect_range(const TraceScene& S, uint32_t b, uint32_t e, f3 org, f3 dir,
                                            float& t, uint32_t& best, Work& wk)
{
    for (uint32_t it = b; it < e; ++it) {
        ++wk.tests;
        const uint32_t p = S.prims[it];
        f3 v0, v1, v2;
        load_tri(S.rec, p, v0, v1, v2);
        const f3 e1 = v1 - v0;
        const f3 e2 = v2 - v0;
        const f3 pv = cross(dir, e2);
        const float det_rcp = 1.f / dot(e1, pv);
        f3 tv = org - v0;
        const float alpha = det_rcp * dot(tv, pv);
        tv = cross(tv, e1);
        const float beta = det_rcp * dot(tv, dir);
        const float dist = det_rcp * dot(e2, tv);
        if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > kIsectEps && dist < t) {
            t = dist;
            best = p;
            if (kAny) return true;
        }
    }
    return false;
}

// entry cell and first crossing distances (UGridTraverser.h:71-126 /
// TLGridTraverser.h:72-126); the uniform grid steps EPS past the entry
template <bool kUniform>
__device__ __forceinline__ bool dda_setup(const GridRegs& G, f3 org, f3 rcp, float& tentry, int& cx,
                                          int& cy, int& cz, float& tx, float& ty, float& tz)
{
    const f3 t1 = (G.bmin - org) * rcp;
    f3 tfar = (G.bmax - org) * rcp;
    const f3 tnear = vmin(t1, tfar);
    tfar = vmax(t1, tfar);
    tentry = fmaxf(fmaxf(tnear.x, tnear.y), tnear.z);
    const float texit = fminf(fminf(tfar.x, tfar.y), tfar.z);
    bool ok = texit > tentry && texit >= 0.f;
    f3 entry;
    if (kUniform) {
        entry = (tentry >= 0.f) ? org + rep(tentry + kEps) / rcp : org;
    } else {
        tentry = fmaxf(0.f, tentry);
        entry = org + rep(tentry) / rcp;
    }
    f3 cf = (entry - G.bmin) * G.csr;
    cf = mk(floorf(cf.x), floorf(cf.y), floorf(cf.z));
    cf.x = fminf((float)G.rx - 1.f, fmaxf(cf.x, 0.f));
    cf.y = fminf((float)G.ry - 1.f, fmaxf(cf.y, 0.f));
    cf.z = fminf((float)G.rz - 1.f, fmaxf(cf.z, 0.f));
    tx = ((cf.x + (rcp.x > 0.f ? 1.f : 0.f)) * G.cs.x + G.bmin.x - org.x) * rcp.x;
    ty = ((cf.y + (rcp.y > 0.f ? 1.f : 0.f)) * G.cs.y + G.bmin.y - org.y) * rcp.y;
    tz = ((cf.z + (rcp.z > 0.f ? 1.f : 0.f)) * G.cs.z + G.bmin.z - org.z) * rcp.z;
    cx = (int)cf.x;
    cy = (int)cf.y;
    cz = (int)cf.z;
    return ok && cx < G.rx && cx > -1 && cy < G.ry && cy > -1 && cz < G.rz && cz > -1;
}

template <bool kAny>
__device__ void trav_uniform(const GridRegs& G, const TraceScene& S, f3 org, f3 rcp, f3 dir, float& t,
                             uint32_t& best, float tterm, Work& wk)
{
    float tentry, tx, ty, tz;
    int cx, cy, cz;
    bool flag = !(kAny && t < tterm);
    flag = dda_setup<true>(G, org, rcp, tentry, cx, cy, cz, tx, ty, tz) && flag;
    const int sx = rcp.x > 0.f ? 1 : -1, sy = rcp.y > 0.f ? 1 : -1, sz = rcp.z > 0.f ? 1 : -1;
    const float dx = G.cs.x * fabsf(rcp.x), dy = G.cs.y * fabsf(rcp.y), dz = G.cs.z * fabsf(rcp.z);
    const uint32_t sxy = (uint32_t)G.rx * (uint32_t)G.ry;
    while (flag) {
        ++wk.steps;
        const uint2 cell = S.cells[(uint32_t)cx + (uint32_t)cy * (uint32_t)G.rx + (uint32_t)cz * sxy];
        isect_range<kAny>(S, cell.x, cell.y, org, dir, t, best, wk);
        const int d = min_dim(tx, ty, tz);
        if (kAny && t < tterm) flag = false;
        if (d == 0) {
            flag = flag && t > tx;
            cx += sx; tx += dx;
            flag = flag && cx < G.rx && cx > -1;
        } else if (d == 1) {
            flag = flag && t > ty;
            cy += sy; ty += dy;
            flag = flag && cy < G.ry && cy > -1;
        } else {
            flag = flag && t > tz;
            cz += sz; tz += dz;
            flag = flag && cz < G.rz && cz > -1;
        }
    }
}

template <bool kAny>
__device__ void trav_two_level(const GridRegs& G, const TraceScene& S, f3 org, f3 rcp, f3 dir, float& t,
                               uint32_t& best, float tterm, Work& wk)
{
    float tentry, tx, ty, tz;
    int cx, cy, cz;
    bool flag = !(kAny && t < tterm);
    flag = dda_setup<false>(G, org, rcp, tentry, cx, cy, cz, tx, ty, tz) && flag;
    const int sx = rcp.x > 0.f ? 1 : -1, sy = rcp.y > 0.f ? 1 : -1, sz = rcp.z > 0.f ? 1 : -1;
    const float dx = G.cs.x * fabsf(rcp.x), dy = G.cs.y * fabsf(rcp.y), dz = G.cs.z * fabsf(rcp.z);
    const uint32_t sxy = (uint32_t)G.rx * (uint32_t)G.ry;
    while (flag) {
        ++wk.steps;
        const uint2 cell = S.cells[(uint32_t)cx + (uint32_t)cy * (uint32_t)G.rx + (uint32_t)cz * sxy];
        bool second = (cell.y & 0x40000000u) != 0u;
        if (second) {
            const int r0 = (int)(cell.y & 0xFFu), r1 = (int)((cell.y >> 8) & 0xFFu),
                      r2 = (int)((cell.y >> 16) & 0xFFu);
            const f3 ep = org + rep(tentry) / rcp;
            const f3 cres = mk((float)r0, (float)r1, (float)r2);
            const f3 srcp = G.csr * cres;
            f3 mb;
            mb.x = (float)cx * G.cs.x + G.bmin.x;
            mb.y = (float)cy * G.cs.y + G.bmin.y;
            mb.z = (float)cz * G.cs.z + G.bmin.z;
            f3 cf;
            cf.x = floorf((ep - mb).x * srcp.x);
            cf.y = floorf((ep - mb).y * srcp.y);
            cf.z = floorf((ep - mb).z * srcp.z);
            const float scx = 1.f / srcp.x, scy = 1.f / srcp.y, scz = 1.f / srcp.z;
            cf = vmin(cres - rep(1.f), vmax(cf, rep(0.f)));
            float ux = ((cf.x + (rcp.x > 0.f ? 1.f : 0.f)) * scx + mb.x - org.x) * rcp.x;
            float uy = ((cf.y + (rcp.y > 0.f ? 1.f : 0.f)) * scy + mb.y - org.y) * rcp.y;
            float uz = ((cf.z + (rcp.z > 0.f ? 1.f : 0.f)) * scz + mb.z - org.z) * rcp.z;
            int lx = (int)cf.x, ly = (int)cf.y, lz = (int)cf.z;
            const float ddx = scx * fabsf(rcp.x), ddy = scy * fabsf(rcp.y), ddz = scz * fabsf(rcp.z);
            const uint32_t r01 = (uint32_t)r0 * (uint32_t)r1;
            while (second) {
                ++wk.steps;
                const uint2 leaf = S.leaves[cell.x + (uint32_t)lx + (uint32_t)r0 * (uint32_t)ly +
                                            r01 * (uint32_t)lz];
                isect_range<kAny>(S, leaf.x, leaf.y, org, dir, t, best, wk);
                if (kAny && t < tterm) { second = false; flag = false; }
                const int d = min_dim(ux, uy, uz);
                if (d == 0) {
                    if (t < ux) { second = false; flag = false; }
                    lx += sx; ux += ddx;
                    second = second && lx != r0 && lx != -1;
                } else if (d == 1) {
                    if (t < uy) { second = false; flag = false; }
                    ly += sy; uy += ddy;
                    second = second && ly != r1 && ly != -1;
                } else {
                    if (t < uz) { second = false; flag = false; }
                    lz += sz; uz += ddz;
                    second = second && lz != r2 && lz != -1;
                }
            }
        }
        if (flag) {
            const int d = min_dim(tx, ty, tz);
            if (d == 0) {
                flag = t > tx;
                tentry = tx; cx += sx; tx += dx;
                flag = flag && cx != G.rx && cx != -1;
            } else if (d == 1) {
                flag = t > ty;
                tentry = ty; cy += sy; ty += dy;
                flag = flag && cy != G.ry && cy != -1;
            } else {
                flag = t > tz;
                tentry = tz; cz += sz; tz += dz;
                flag = flag && cz != G.rz && cz != -1;
            }
        }
    }
}

// trace<> body: rcp = 1 / d, traverse, d' = 1 / rcp (RayTracingKernels.h:118-140)
template <bool kTwo, bool kAny>
__device__ __forceinline__ uint32_t trace_ray(const GridRegs& G, const TraceScene& S, f3 org, f3 d,
                                              float& t, float tterm, f3& dir_out, Work& wk)
{
    const f3 rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    const f3 dir = mk(1.f / rcp.x, 1.f / rcp.y, 1.f / rcp.z);
    uint32_t best = S.nt;
    if (kTwo) trav_two_level<kAny>(G, S, org, rcp, dir, t, best, tterm, wk);
    else trav_uniform<kAny>(G, S, org, rcp, dir, t, best, tterm, wk);
    dir_out = dir;
    return best;
}

__device__ __forceinline__ void count_rays(unsigned long long* counter, bool traced)
{
    const uint64_t b = __ballot(traced);
    if (counter && __lane_id() == 0 && b) atomicAdd(counter, (unsigned long long)__popcll(b));
}

// {rays, steps, tests} work counters, one atomic per wave and counter;
// every lane of the wave must call it
__device__ __forceinline__ void add_work(unsigned long long* stats, uint32_t rays, const Work& wk)
{
    if (!stats) return;
    unsigned long long r = rays, s = wk.steps, t = wk.tests;
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) {
        r += __shfl_xor(r, d, kWave);
        s += __shfl_xor(s, d, kWave);
        t += __shfl_xor(t, d, kWave);
    }
    if (__lane_id() == 0) {
        atomicAdd(stats + 0, r);
        atomicAdd(stats + 1, s);
        atomicAdd(stats + 2, t);
    }
}

// ---------------------------------------------------------------------------
// primary rays
// ---------------------------------------------------------------------------

__device__ __forceinline__ f3 primary_direction(const CameraDev& cam, const RenderArgs& ra,
                                                uint32_t pix, uint32_t npix)
{
    const float pid = (float)(pix < npix - 1u ? pix : npix - 1u);
    const float resx = (float)cam.res_x;
    float sx, sy;
    if (!ra.jitter) {
        // RegularPixelSampler<1,1>, sample 0 (Camera.h:82-97)
        sy = 0.5f;
        sx = 0.5f;
        const float tmp = pid / resx;
        sy += truncf(tmp);
        sx += (tmp - truncf(tmp)) * resx;
    } else {
        // GaussianPixelSampler (Camera.h:121-131)
        float u[4];
        rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)ra.image, kStreamPrimary, u);
        const float tmp = pid / resx;
        const float rad = 0.5f * sqrtf(-spec_logf(1 - u[0]));
        const float angle = 2.f * kPi * u[1];
        float s, c;
        spec_sincosf(angle, s, c);
        sy = truncf(tmp) + rad * s;
        sx = (tmp - truncf(tmp)) * resx + rad * c;
    }
    const f3 tl = mk(cam.top_left[0], cam.top_left[1], cam.top_left[2]);
    const f3 stx = mk(cam.step_x[0], cam.step_x[1], cam.step_x[2]);
    const f3 sty = mk(cam.step_y[0], cam.step_y[1], cam.step_y[2]);
    return tl + sx * stx + sy * sty;
}

template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_primary(TraceScene S, CameraDev cam, RenderArgs ra,
                                                    uint32_t npix, float* __restrict__ hits,
                                                    uint32_t* __restrict__ hit_prims,
                                                    unsigned long long* ray_count)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = pix < npix;
    count_rays(ray_count, active);
    Work wk;
    if (active) {
        const GridRegs G = load_grid(S.gp);
        const f3 d0 = primary_direction(cam, ra, pix, npix);
        const f3 org = mk(cam.center[0], cam.center[1], cam.center[2]);
        float t = kFltMax;
        f3 dir;
        const uint32_t best = trace_ray<kTwo, false>(G, S, org, d0, t, 0.f, dir, wk);
        const uint32_t prim = t < kFltMax ? best : S.nt;
        const f3 hp = org + (t - kEps) * dir;
        hits[pix] = hp.x;
        hits[(size_t)npix + pix] = hp.y;
        hits[2 * (size_t)npix + pix] = hp.z;
        hits[3 * (size_t)npix + pix] = dir.x;
        hits[4 * (size_t)npix + pix] = dir.y;
        hits[5 * (size_t)npix + pix] = dir.z;
        hits[6 * (size_t)npix + pix] = t;
        hit_prims[pix] = prim;
    }
    add_work(ra.stats, active ? 1u : 0u, wk);
}

// ---------------------------------------------------------------------------
// ambient occlusion
// ---------------------------------------------------------------------------

// CosineHemisphereSampler (HemisphereSamplers.hpp:58-71)
__device__ __forceinline__ f3 cosine_dir(float u1, float u2)
{
    float s, c;
    spec_sincosf(2.f * kPi * u1, s, c);
    f3 r;
    r.x = c * sqrtf(1 - u2);
    r.y = s * sqrtf(1 - u2);
    r.z = sqrtf(u2);
    return r;
}

template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_ao_rays(TraceScene S, RenderArgs ra, uint32_t npix,
                                                    const float* __restrict__ hits,
                                                    const uint32_t* __restrict__ hit_prims,
                                                    uint8_t* __restrict__ occ,
                                                    unsigned long long* ray_count)
{
    const uint32_t S_ = (uint32_t)ra.ao_samples;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = j < npix * S_;
    const uint32_t pix = in ? j / S_ : 0u;
    const uint32_t s = in ? j % S_ : 0u;
    const uint32_t prim = in ? hit_prims[pix] : S.nt;
    const float th = in ? hits[6 * (size_t)npix + pix] : kFltMax;
    const bool active = in && th < kFltMax && prim < S.nt;
    count_rays(ray_count, active);
    Work wk;
    if (in && !active) occ[j] = 1u;
    if (active) {
    const GridRegs G = load_grid(S.gp);
    // AmbientOcclusionRayGenerator::operator() (RayGenerators.h:281-343)
    f3 org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
    const f3 pdir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix],
                       hits[5 * (size_t)npix + pix]);
    f3 v0, v1, v2;
    load_tri(S.rec, prim, v0, v1, v2);
    f3 n = normalize(cross(v1 - v0, v2 - v0));
    if (dot(n, pdir) > 0.f) n = -n;
    float u[4];
    rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)ra.image, kStreamAO + s, u);
    const f3 h = cosine_dir(u[0], u[1]);
    f3 tg, bn;
    local_frame(n, tg, bn);
    const f3 dir = normalize(n * h.z + tg * h.x + bn * h.y);
    org = org + (0.01f * ra.ao_dist) * dir;
    float t = ra.ao_dist;
    f3 d2;
    trace_ray<kTwo, true>(G, S, org, dir, t, ra.ao_dist, d2, wk);
    occ[j] = t < ra.ao_dist ? 1u : 0u;
    }
    add_work(ra.stats, active ? 1u : 0u, wk);
}

// computeAOIllumination (AOIntegrator.h:82-188), one lane per pixel
__global__ __launch_bounds__(kBlock) void k_ao_shade(TraceScene S, RenderArgs ra, uint32_t npix,
                                                     const float* __restrict__ hits,
                                                     const uint32_t* __restrict__ hit_prims,
                                                     const uint8_t* __restrict__ occ, Shading sh,
                                                     float* __restrict__ fb)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const uint32_t S_ = (uint32_t)ra.ao_samples;
    const float th = hits[6 * (size_t)npix + pix];
    const uint32_t prim = hit_prims[pix];
    f3 n = rep(0.f), kd = rep(0.f), pdir = rep(0.f);
    const bool hit = th < kFltMax;
    if (hit && prim < S.nt) {
        f3 v0, v1, v2;
        load_tri(S.rec, prim, v0, v1, v2);
        n = normalize(cross(v1 - v0, v2 - v0));
        const float* m = sh.mats + 12 * (size_t)(sh.mat_idx ? sh.mat_idx[prim] : 0u);
        kd = mk(m[0], m[1], m[2]);
        pdir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix],
                  hits[5 * (size_t)npix + pix]);
    }
    f3 rad = rep(0.f);
    for (uint32_t s = 0; s < S_; ++s) {
        f3 v;
        if (hit) {
            v = occ[(size_t)pix * S_ + s] ? rep(0.f) : rep(1.f);
            if (v.x + v.y + v.z > 0.f && prim < S.nt) {
                v.x *= kd.x * kPi;
                v.y *= kd.y * kPi;
                v.z *= kd.z * kPi;
                if (kd.x * kPi < 0.9f || kd.y * kPi < 0.9f || kd.z * kPi < 0.9f) {
                    const float sin_zn = fmaxf(0.f, sqrtf(1.f - n.z));
                    const float cos_en = fabsf(dot(pdir, n));
                    v.x *= (0.7f * cos_en + 0.3f * sin_zn);
                    v.y *= (0.7f * cos_en + 0.2f * sin_zn);
                    v.z *= (0.9f * cos_en);
                } else {
                    v = rep(1.f);
                }
            }
        } else {
            v = rep(1.f);
        }
        rad = rad + (v * 1.f) / (float)S_;
    }
    const float nw = 1.f / (float)(ra.image + 1);
    const float ow = 1.f - nw;
    float* o = fb + 3 * (size_t)pix;
    o[0] = o[0] * ow + rad.x * nw;
    o[1] = o[1] * ow + rad.y * nw;
    o[2] = o[2] * ow + rad.z * nw;
}

// ---------------------------------------------------------------------------
// path tracing megakernel (PathTracer.h:41-372), one lane per pixel
// ---------------------------------------------------------------------------

template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_pt(TraceScene S, RenderArgs ra, uint32_t npix,
                                               const float* __restrict__ hits,
                                               const uint32_t* __restrict__ hit_prims, Shading sh,
                                               float* __restrict__ fb, unsigned long long* ray_count)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = pix < npix;
    const GridRegs G = load_grid(S.gp);
    f3 radiance = rep(0.f);
    float ray_t = in ? hits[6 * (size_t)npix + pix] : kFltMax;
    uint32_t prim = in ? hit_prims[pix] : S.nt;
    uint32_t traced = 0;
    Work wk;
    if (ray_t < kFltMax && prim < S.nt) {
        f3 org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        f3 dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix],
                    hits[5 * (size_t)npix + pix]);
        f3 att = rep(1.f);
        bool last_spec = true;
        for (int depth = 0; depth < ra.pt_depth; ++depth) {
            float u[4], w[4];
            rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)ra.image, kStreamPT + 2u * (uint32_t)depth, u);
            rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)ra.image, kStreamPT + 2u * (uint32_t)depth + 1u,
                  w);
            f3 v0, v1, v2;
            load_tri(S.rec, prim, v0, v1, v2);
            f3 n = normalize(cross(v1 - v0, v2 - v0));
            float cos_ne = -dot(n, dir);
            const float* m = sh.mats + 12 * (size_t)(sh.mat_idx ? sh.mat_idx[prim] : 0u);
            const f3 kd = mk(m[0], m[1], m[2]);
            const f3 ks = mk(m[4], m[5], m[6]);
            const f3 ke = mk(m[8], m[9], m[10]);
            const float se = m[7];
            const float ior = m[3] - 10.f;
            bool end = false;
            if (ior > 0.f) {
                // transparent (PathTracer.h:133-183)
                last_spec = true;
                float n1 = 1.000293f, n2 = ior;
                if (cos_ne < 0.f) {
                    n1 = ior;
                    n2 = 1.000293f;
                    n = -n;
                    cos_ne = -cos_ne;
                    att.x *= spec_expf(-ray_t * kd.x * kPi);
                    att.y *= spec_expf(-ray_t * kd.y * kPi);
                    att.z *= spec_expf(-ray_t * kd.z * kPi);
                }
                const float f0 = (n1 - n2) * (n1 - n2) / ((n1 + n2) * (n1 + n2));
                const float fres = f0 + (1.f - f0) * spec_powf(1.f - fabsf(cos_ne), 5.f);
                const float nr = n1 / n2;
                const float sin2 = nr * nr * (1.f - cos_ne * cos_ne);
                if (sin2 >= 1.f || w[2] < fres) {
                    org = org + (5.f * kEps) * n;
                    dir = dir + (2 * cos_ne) * n;
                    if (sin2 >= 1.f) {
                        if (w[3] < 0.5f) att = att * 2.f;
                        else end = true;
                    }
                } else {
                    org = org - (5.f * kEps) * n;
                    const float ns = nr * cos_ne - sqrtf(1.f - sin2);
                    dir = nr * dir + ns * n;
                }
            } else {
                // opaque (PathTracer.h:184-313)
                if (cos_ne < 0.f) { n = -n; cos_ne = -cos_ne; }
                org = org + (5.f * kEps) * n;
                if (last_spec) radiance = radiance + att * ke;
                last_spec = false;
                if (sh.nlights > 0) {
                    // AreaLightSourceCollection::getLight (LightSource.hpp:116-128)
                    int li = sh.nlights - 1;
                    for (int i = 0; i < sh.nlights - 1; ++i)
                        if (u[0] < sh.light_cdf[i] / sh.light_cdf[sh.nlights - 1]) { li = i; break; }
                    const float* L = sh.lights + 15 * (size_t)li;
                    const f3 lpos = mk(L[0], L[1], L[2]), lnrm = mk(L[3], L[4], L[5]);
                    const f3 lint = mk(L[6], L[7], L[8]), le1 = mk(L[9], L[10], L[11]),
                             le2 = mk(L[12], L[13], L[14]);
                    const f3 lp = lpos + u[1] * le1 + u[2] * le2;
                    const f3 sdir = lp - org;
                    const float d2r = 1.f / dot(sdir, sdir);
                    const f3 sdn = sdir * sqrtf(d2r);
                    const float cos_nl = dot(n, sdn);
                    const float cos_hl = fmaxf(0.f, dot(normalize(sdn - dir), n));
                    const float cos_ln = dot(-lnrm, sdn);
                    if (cos_nl > 0.f && cos_ln > 0.f) {
                        float st = 0.9999f;
                        f3 dd;
                        trace_ray<kTwo, true>(G, S, org, sdir, st, 0.9999f, dd, wk);
                        ++traced;
                        if (!(st < 0.9999f)) {
                            const float area = len(cross(le1, le2));
                            const f3 ls = ((lint * area) * d2r) * cos_ln;
                            radiance = radiance + ((att * kd) * ls) * cos_nl;
                            if (!last_spec)
                                radiance = radiance +
                                           (((att * ks) * fmaxf(0.f, spec_powf(cos_hl, se))) * ls) * cos_nl;
                        }
                    }
                }
                if (depth + 1 >= ra.pt_depth) break;
                const float albedo_d = hmax(kd) * kPi;
                const float albedo_s = hmax(ks) * 2.f * kPi / (se + 2);
                const float rnum = u[3];
                if (rnum < albedo_d) {
                    att = ((att * kd) * kPi) / albedo_d;
                    const f3 rd = cosine_dir(w[0], w[1]);
                    f3 tg, bn;
                    local_frame(n, tg, bn);
                    dir = tg * rd.x + bn * rd.y + n * rd.z;
                } else if (rnum < albedo_d + albedo_s) {
                    // PowerCosineHemisphereSampler (HemisphereSamplers.hpp:78-94)
                    f3 rd;
                    rd.z = spec_powf(w[1], 1.f / (se + 1.f));
                    float s_, c_;
                    spec_sincosf(2.f * kPi * w[0], s_, c_);
                    rd.x = c_ * sqrtf(1 - rd.z * rd.z);
                    rd.y = s_ * sqrtf(1 - rd.z * rd.z);
                    const f3 upv = dir + (2.f * cos_ne) * n;
                    f3 rt, fw;
                    local_frame(upv, rt, fw);
                    dir = rt * rd.x + fw * rd.y + upv * rd.z;
                    last_spec = true;
                    const float cnl = dot(dir, n);
                    if (cnl < 0.f) end = true;
                    else att = (((att * cnl) * ks) * kPi) / ((se + 1.f) * 0.5f * albedo_s);
                } else {
                    end = true;
                }
            }
            if (end || depth + 1 >= ra.pt_depth) break;
            // extension ray (PathTracer.h:318-332)
            float t = kFltMax;
            f3 d2;
            const uint32_t hit = trace_ray<kTwo, false>(G, S, org, dir, t, 0.f, d2, wk);
            ++traced;
            dir = d2;
            if (!(t < kFltMax)) break;
            org = org + (t - kEps) * dir;
            prim = hit;
            ray_t = t;
        }
    }
    if (ray_count && traced) atomicAdd(ray_count, (unsigned long long)traced);
    add_work(ra.stats, traced, wk);
    if (!in) return;
    const float nw = 1.f / (float)(ra.image + 1);
    const float ow = 1.f - nw;
    float* o = fb + 3 * (size_t)pix;
    o[0] = o[0] * ow + radiance.x * nw;
    o[1] = o[1] * ow + radiance.y * nw;
    o[2] = o[2] * ow + radiance.z * nw;
}

// ---------------------------------------------------------------------------
// ray batches (trace<> with RayLoader, RayGenerators.h:347-364)
// ---------------------------------------------------------------------------

template <bool kTwo, bool kAny>
__global__ __launch_bounds__(kBlock) void k_trace_batch(TraceScene S, const float* __restrict__ org,
                                                        const float* __restrict__ dir, uint32_t n,
                                                        float tmax, float* __restrict__ t_out,
                                                        uint32_t* __restrict__ prim_out,
                                                        uint8_t* __restrict__ occ_out,
                                                        unsigned long long* ray_count,
                                                        unsigned long long* stats)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    count_rays(ray_count, i < n);
    Work wk;
    if (i < n) {
        const GridRegs G = load_grid(S.gp);
        const f3 o = mk(org[i], org[(size_t)n + i], org[2 * (size_t)n + i]);
        const f3 d = mk(dir[i], dir[(size_t)n + i], dir[2 * (size_t)n + i]);
        float t = kAny ? tmax : kFltMax;
        f3 d2;
        const uint32_t best = trace_ray<kTwo, kAny>(G, S, o, d, t, tmax, d2, wk);
        if (kAny) {
            occ_out[i] = t < tmax ? 1u : 0u;
            if (t_out) t_out[i] = t;
        } else {
            t_out[i] = t;
            prim_out[i] = t < kFltMax ? best : S.nt;
        }
    }
    add_work(stats, i < n ? 1u : 0u, wk);
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_primary(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, float* hits,
                    uint32_t* hit_prims, unsigned long long* ray_count, hipStream_t s)
{
    const uint32_t npix = (uint32_t)cam.res_x * (uint32_t)cam.res_y;
    if (S.two_level)
        hipLaunchKernelGGL(k_primary<true>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, cam, ra, npix,
                           hits, hit_prims, ray_count);
    else
        hipLaunchKernelGGL(k_primary<false>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, cam, ra, npix,
                           hits, hit_prims, ray_count);
}

void launch_ao_rays(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                    const uint32_t* hit_prims, uint8_t* occ, unsigned long long* ray_count, hipStream_t s)
{
    const size_t nrays = (size_t)npix * (size_t)ra.ao_samples;
    if (S.two_level)
        hipLaunchKernelGGL(k_ao_rays<true>, dim3(blocks_for(nrays)), dim3(kBlock), 0, s, S, ra, npix, hits,
                           hit_prims, occ, ray_count);
    else
        hipLaunchKernelGGL(k_ao_rays<false>, dim3(blocks_for(nrays)), dim3(kBlock), 0, s, S, ra, npix, hits,
                           hit_prims, occ, ray_count);
}

void launch_ao_shade(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                     const uint32_t* hit_prims, const uint8_t* occ, const Shading& sh, float* fb,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_ao_shade, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits, hit_prims,
                       occ, sh, fb);
}

void launch_pt(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
               const uint32_t* hit_prims, const Shading& sh, float* fb,
               unsigned long long* ray_count, hipStream_t s)
{
    if (S.two_level)
        hipLaunchKernelGGL(k_pt<true>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits,
                           hit_prims, sh, fb, ray_count);
    else
        hipLaunchKernelGGL(k_pt<false>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits,
                           hit_prims, sh, fb, ray_count);
}

void launch_trace_batch(const TraceScene& S, const float* org, const float* dir, uint32_t n,
                        int anyhit, float tmax, float* t_out, uint32_t* prim_out, uint8_t* occ_out,
                        unsigned long long* ray_count, unsigned long long* stats, hipStream_t s)
{
    const dim3 g(blocks_for(n > 0 ? n : 1)), b(kBlock);
    if (S.two_level) {
        if (anyhit)
            hipLaunchKernelGGL((k_trace_batch<true, true>), g, b, 0, s, S, org, dir, n, tmax, t_out, prim_out,
                               occ_out, ray_count, stats);
        else
            hipLaunchKernelGGL((k_trace_batch<true, false>), g, b, 0, s, S, org, dir, n, tmax, t_out, prim_out,
                               occ_out, ray_count, stats);
    } else {
        if (anyhit)
            hipLaunchKernelGGL((k_trace_batch<false, true>), g, b, 0, s, S, org, dir, n, tmax, t_out, prim_out,
                               occ_out, ray_count, stats);
        else
            hipLaunchKernelGGL((k_trace_batch<false, false>), g, b, 0, s, S, org, dir, n, tmax, t_out, prim_out,
                               occ_out, ray_count, stats);
    }
}

}  // namespace hr
