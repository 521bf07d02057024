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

// a pointer into LDS (the workgroup's shared memory)
typedef __attribute__((address_space(3))) const float* lds_cfloat_ptr;

struct GridRegs {
    f3 bmin, bmax, cs, csr;
    int rx, ry, rz;
    lds_cfloat_ptr sub;   // nullptr, or the workgroup's table of leaf cell sizes (sub_table)
};

// Leaf cell size of a top cell per axis: 1 / (cell_size_rcp * leaf
// resolution), the three divisions of a top-cell visit (TLGridTraverser.h:
// 160-172).  The resolution is an integer in [1, 255], so a workgroup keeps
// the values of one grid in a table of 3 x 256 entries in LDS, computed by the
// same expression, and a visit reads them instead of dividing.
constexpr int kSubTable = 3 * 256;

// every thread of the workgroup must call this (it synchronizes)
__device__ __forceinline__ void sub_table(GridRegs& g, float* tab)
{
    for (int i = (int)threadIdx.x; i < kSubTable; i += (int)blockDim.x) {
        const int a = i >> 8;
        const float c = a == 0 ? g.csr.x : (a == 1 ? g.csr.y : g.csr.z);
        tab[i] = 1.f / (c * (float)(i & 255));
    }
    __syncthreads();
    g.sub = (lds_cfloat_ptr)tab;
}

// kTab: from the table (G.sub must be set), else by division
template <bool kTab = true>
__device__ __forceinline__ f3 leaf_cell_size(const GridRegs& G, f3 srcp, int r0, int r1, int r2)
{
    if (kTab) return mk(G.sub[r0], G.sub[256 + r1], G.sub[512 + r2]);
    return mk(1.f / srcp.x, 1.f / srcp.y, 1.f / srcp.z);
}

__device__ __forceinline__ GridRegs load_grid(const GridParams* gp)
{
    GridRegs g;
    g.sub = nullptr;
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

// candidates [b, e) of one cell or leaf.  kPipe: software pipeline -- the
// next candidate's triangle and the id after it are in flight while the
// current triangle is tested; otherwise only the next id is fetched ahead
// (fewer live registers, for callers deep in a loop nest)
template <bool kAny, bool kPipe = true>
__device__ __forceinline__ bool isect_range(const TraceScene& S, uint32_t b, uint32_t e, f3 org, f3 dir,
                                            float& t, uint32_t& best, Work& wk)
{
    if (b >= e) return false;
    if (!kPipe) {
        uint32_t pn = S.prims[b];
        for (uint32_t it = b; it < e; ++it) {
            ++wk.tests;
            const uint32_t p = pn;
            if (it + 1 < e) pn = S.prims[it + 1];
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
    uint32_t p = S.prims[b];
    f3 n0, n1, n2;
    load_tri(S.rec, p, n0, n1, n2);
    uint32_t pn = b + 1 < e ? S.prims[b + 1] : 0u;
    for (uint32_t it = b; it < e; ++it) {
        ++wk.tests;
        const uint32_t pc = p;
        const f3 v0 = n0, v1 = n1, v2 = n2;
        if (it + 1 < e) {
            p = pn;
            load_tri(S.rec, p, n0, n1, n2);
            if (it + 2 < e) pn = S.prims[it + 2];
        }
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
            best = pc;
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
            const f3 sc = leaf_cell_size(G, srcp, r0, r1, r2);
            const float scx = sc.x, scy = sc.y, scz = sc.z;
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


// Tiled pixel order for the one-ray-per-lane camera kernel: 8x8-pixel tiles
// in row-major order, the pixels of a tile in Z order, so that a wave's 64
// camera rays cover one tile and visit nearby cells (camera rays of the 1M-
// triangle soup 0.679 -> 0.640 ms; the wave-batched camera kernel and the AO
// rays measured slower in this order and keep row order).  Slots past the
// image in its last row and column of tiles map to no pixel (false).  Only
// the assignment of rays to lanes changes.
__host__ __device__ __forceinline__ uint32_t tiled_slots(uint32_t res_x, uint32_t res_y)
{
    return ((res_x + 7u) >> 3) * ((res_y + 7u) >> 3) * 64u;
}

__device__ __forceinline__ bool tiled_pixel(uint32_t q, uint32_t res_x, uint32_t res_y, uint32_t& pix)
{
    const uint32_t tiles_x = (res_x + 7u) >> 3;
    const uint32_t t = q >> 6, w = q & 63u;
    const uint32_t lx = (w & 1u) | ((w >> 1) & 2u) | ((w >> 2) & 4u);
    const uint32_t ly = ((w >> 1) & 1u) | ((w >> 2) & 2u) | ((w >> 3) & 4u);
    const uint32_t x = (t % tiles_x) * 8u + lx, y = (t / tiles_x) * 8u + ly;
    pix = y * res_x + x;
    return x < res_x && y < res_y;
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

// ---------------------------------------------------------------------------
// path tracer vertex (PathTracer.h:41-372), shared by the one-lane-per-pixel
// megakernel and the persistent path kernel so both apply the same operations
// ---------------------------------------------------------------------------

struct PathState {
    f3 org, dir, att, radiance;
    float ray_t;
    uint32_t prim;
    int depth;
    bool last_spec;
};

// what a vertex leaves for its shadow ray and its bounce
struct PathVertex {
    bool opaque, end, shadow;
    f3 n;
    float cos_ne;
    f3 kd, ks;
    float se, u3, w0, w1;
    f3 sdir;           // shadow ray direction (to the light point), from P.org
    f3 add_d, add_s;   // next-event terms added when the shadow ray is unoccluded
};

// one path vertex up to its shadow ray: random numbers, material, the
// transparent branch in full (PathTracer.h:133-183), the opaque branch's
// offset, emission and light sample (PathTracer.h:184-260)
__device__ __forceinline__ void pt_vertex(const TraceScene& S, const Shading& sh, const RenderArgs& ra, uint32_t pix,
                                          PathState& P, PathVertex& V)
{
    float u[4], w[4];
    rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamPT + 2u * (uint32_t)P.depth, u);
    rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamPT + 2u * (uint32_t)P.depth + 1u, w);
    f3 v0, v1, v2;
    load_tri(S.rec, P.prim, v0, v1, v2);
    f3 n = shading_normal(sh, P.prim, P.org, v0, v1, v2);
    float cos_ne = -dot(n, P.dir);
    const float* m = sh.mats + 12 * (size_t)(sh.mat_idx ? sh.mat_idx[P.prim] : 0u);
    const f3 kd = mk(m[0], m[1], m[2]);
    const f3 ks = mk(m[4], m[5], m[6]);
    const f3 ke = mk(m[8], m[9], m[10]);
    const float se = m[7];
    const float ior = m[3] - 10.f;
    V.end = false;
    V.shadow = false;
    V.opaque = !(ior > 0.f);
    if (!V.opaque) {
        P.last_spec = true;
        float n1 = 1.000293f, n2 = ior;
        if (cos_ne < 0.f) {
            n1 = ior;
            n2 = 1.000293f;
            n = -n;
            cos_ne = -cos_ne;
            P.att.x *= spec_expf(-P.ray_t * kd.x * kPi);
            P.att.y *= spec_expf(-P.ray_t * kd.y * kPi);
            P.att.z *= spec_expf(-P.ray_t * kd.z * kPi);
        }
        const float f0 = (n1 - n2) * (n1 - n2) / ((n1 + n2) * (n1 + n2));
        const float fres = f0 + (1.f - f0) * spec_powf(1.f - fabsf(cos_ne), 5.f);
        const float nr = n1 / n2;
        const float sin2 = nr * nr * (1.f - cos_ne * cos_ne);
        if (sin2 >= 1.f || w[2] < fres) {
            P.org = P.org + (5.f * kEps) * n;
            P.dir = P.dir + (2 * cos_ne) * n;
            if (sin2 >= 1.f) {
                if (w[3] < 0.5f) P.att = P.att * 2.f;
                else V.end = true;
            }
        } else {
            P.org = P.org - (5.f * kEps) * n;
            const float ns = nr * cos_ne - sqrtf(1.f - sin2);
            P.dir = nr * P.dir + ns * n;
        }
        return;
    }
    if (cos_ne < 0.f) { n = -n; cos_ne = -cos_ne; }
    P.org = P.org + (5.f * kEps) * n;
    if (P.last_spec) P.radiance = P.radiance + P.att * ke;
    P.last_spec = false;
    V.n = n;
    V.cos_ne = cos_ne;
    V.kd = kd;
    V.ks = ks;
    V.se = se;
    V.u3 = u[3];
    V.w0 = w[0];
    V.w1 = w[1];
    if (sh.nlights > 0) {
        // AreaLightSourceCollection::getLight (LightSource.hpp:116-128)
        int li = sh.nlights - 1;
        for (int i = 0; i < sh.nlights - 1; ++i)
            if (u[0] < sh.light_cdf[i] / sh.light_cdf[sh.nlights - 1]) { li = i; break; }
        const float* L = sh.lights + 15 * (size_t)li;
        const f3 lpos = mk(L[0], L[1], L[2]), lnrm = mk(L[3], L[4], L[5]);
        const f3 lint = mk(L[6], L[7], L[8]), le1 = mk(L[9], L[10], L[11]), le2 = mk(L[12], L[13], L[14]);
        const f3 lp = lpos + u[1] * le1 + u[2] * le2;
        const f3 sdir = lp - P.org;
        const float d2r = 1.f / dot(sdir, sdir);
        const f3 sdn = sdir * sqrtf(d2r);
        const float cos_nl = dot(n, sdn);
        const float cos_hl = fmaxf(0.f, dot(normalize(sdn - P.dir), n));
        const float cos_ln = dot(-lnrm, sdn);
        if (cos_nl > 0.f && cos_ln > 0.f) {
            V.shadow = true;
            V.sdir = sdir;
            const float area = len(cross(le1, le2));
            const f3 ls = ((lint * area) * d2r) * cos_ln;
            V.add_d = ((P.att * kd) * ls) * cos_nl;
            V.add_s = (((P.att * ks) * fmaxf(0.f, spec_powf(cos_hl, se))) * ls) * cos_nl;
        }
    }
}

// the shadow ray's outcome: the light sample counts when nothing is closer
// than 0.9999 (PathTracer.h:244-258)
__device__ __forceinline__ void pt_light(PathState& P, const PathVertex& V, float shadow_t)
{
    if (!(shadow_t < 0.9999f)) {
        P.radiance = P.radiance + V.add_d;
        if (!P.last_spec) P.radiance = P.radiance + V.add_s;
    }
}

// Russian roulette and the bounce of an opaque vertex (PathTracer.h:262-313)
__device__ __forceinline__ void pt_bounce(PathState& P, PathVertex& V)
{
    const f3 kd = V.kd, ks = V.ks, n = V.n;
    const float se = V.se;
    const float albedo_d = hmax(kd) * kPi;
    const float albedo_s = hmax(ks) * 2.f * kPi / (se + 2);
    const float rnum = V.u3;
    if (rnum < albedo_d) {
        P.att = ((P.att * kd) * kPi) / albedo_d;
        const f3 rd = cosine_dir(V.w0, V.w1);
        f3 tg, bn;
        local_frame(n, tg, bn);
        P.dir = tg * rd.x + bn * rd.y + n * rd.z;
    } else if (rnum < albedo_d + albedo_s) {
        // PowerCosineHemisphereSampler (HemisphereSamplers.hpp:78-94)
        f3 rd;
        rd.z = spec_powf(V.w1, 1.f / (se + 1.f));
        float s_, c_;
        spec_sincosf(2.f * kPi * V.w0, s_, c_);
        rd.x = c_ * sqrtf(1 - rd.z * rd.z);
        rd.y = s_ * sqrtf(1 - rd.z * rd.z);
        const f3 upv = P.dir + (2.f * V.cos_ne) * n;
        f3 rt, fw;
        local_frame(upv, rt, fw);
        P.dir = rt * rd.x + fw * rd.y + upv * rd.z;
        P.last_spec = true;
        const float cnl = dot(P.dir, n);
        if (cnl < 0.f) V.end = true;
        else P.att = (((P.att * cnl) * ks) * kPi) / ((se + 1.f) * 0.5f * albedo_s);
    } else {
        V.end = true;
    }
}

}  // namespace trav
}  // namespace hr
