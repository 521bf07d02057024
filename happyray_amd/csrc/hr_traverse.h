// hr_traverse.h -- device-side grid traversal shared by the one-ray-per-lane
// kernels (hr_trace.hip) and the persistent kernels (hr_persist.hip).
//
// Reference behaviour followed:
//   MollerTrumboreIntersectionTest   src/RT/Algorithm/RayTriangleIntersector.h:36-92
//   UGridTraverser                   src/RT/Algorithm/UGridTraverser.h:44-177
//   TLGridTraverser                  src/RT/Algorithm/TLGridTraverser.h:45-354
//   trace<> (rcp, output)            src/RT/Algorithm/RayTracingKernels.h:118-149
#pragma once

#include "hr_trace.h"

namespace hr {
namespace trav {

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


// RegularPixelSampler<1,1> / GaussianPixelSampler and Camera::getDirection
// (src/RT/Primitive/Camera.h:68-131)
__device__ __forceinline__ f3 primary_direction(const CameraDev& cam, const RenderArgs& ra,
                                                uint32_t pix, uint32_t npix)
{
    const float pid = (float)(pix < npix - 1u ? pix : npix - 1u);
    const float resx = (float)cam.res_x;
    float sx, sy;
    const int image = image_of(ra);
    const bool regular = ra.simple_sampler ? image < 4 : !ra.jitter;
    if (regular) {
        // RegularPixelSampler<k,k> (Camera.h:82-97): k = 2 with sample id =
        // image for the simple integrator, else k = 1 (the pixel centre)
        const int k = ra.simple_sampler ? 2 : 1;
        float tmp = (float)((int)(float)image % (k * k)) / (float)k;
        if (!ra.simple_sampler) tmp = 0.f;
        sy = (truncf(tmp) + 0.5f) / (float)k;
        sx = tmp - truncf(tmp) + 0.5f / (float)k;
        tmp = pid / resx;
        sy += truncf(tmp);
        sx += (tmp - truncf(tmp)) * resx;
    } else {
        // GaussianPixelSampler (Camera.h:121-131)
        float u[4];
        rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image, kStreamPrimary, u);
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

// Shading normal of the path tracers (PathTracer.h:109-135, :526-553): barycentric
// weights from the sub-triangle areas seen from the hit point, then the
// normalized blend of the corner normals; the geometric normal when the scene
// has no normals.
__device__ __forceinline__ f3 shading_normal(const Shading& sh, uint32_t prim, f3 org, f3 v0, f3 v1, f3 v2)
{
    const f3 real_normal = cross(v1 - v0, v2 - v0);
    if (!sh.nidx) return normalize(real_normal);
    const f3 a = v0 - org, b = v1 - org, c = v2 - org;
    const f3 n0 = cross(b, c), n1 = cross(c, a);
    const float rcp = 1.f / sqrtf(dot(real_normal, real_normal));
    const float u = len(n0) * rcp, v = len(n1) * rcp;
    const float4 c0 = sh.normals[sh.nidx[3 * (size_t)prim]];
    const float4 c1 = sh.normals[sh.nidx[3 * (size_t)prim + 1]];
    const float4 c2 = sh.normals[sh.nidx[3 * (size_t)prim + 2]];
    return normalize(u * mk(c0.x, c0.y, c0.z) + v * mk(c1.x, c1.y, c1.z) + (1.f - u - v) * mk(c2.x, c2.y, c2.z));
}

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

}  // namespace trav
}  // namespace hr
