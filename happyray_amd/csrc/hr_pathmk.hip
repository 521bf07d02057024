// hr_pathmk.hip -- the multi-kernel path tracer (RTEngine::renderFrame mode 1)
// for gfx950: path tracing to the end vertex, one shadow ray per pixel there,
// direct light weighted by the path throughput, and the blur / residue merge
// into the framebuffer.
//
// Reference behaviour followed:
//   tracePath                        src/RT/Integrator/PathTracer.h:374-748
//   AreaLightShadowRayGenerator      src/RT/Algorithm/RayGenerators.h:118-206
//   AreaLightSource::isOnLS          src/RT/Primitive/LightSource.hpp:50-76
//   DirectIlluminationBuffer::store  src/RT/Structure/RayBuffers.h:211-235
//   computeDirectIllumination        src/RT/Integrator/PathTracer.h:751-880
//   addIndirectIllumination          src/RT/Integrator/PathTracer.h:889-990
//   PathTracer::integrate            src/RT/Integrator/PathTracer.h:1038-1215
//
// Departures, shared with the CPU oracle: the counter-based RNG keyed by
// (pixel, image, stream) in place of the thread-seeded XorShift32+; the
// shadow ray is an any-hit ray occluded by a hit in (1e-4, 0.9999) as in the
// megakernel; the merge reads the framebuffer and residue of the previous
// image only (the reference's 16x16 blocks overlap by a one-pixel apron and
// read neighbours that other blocks may already have rewritten, and its
// threads at x = resX or y = resY rewrite the last column / row a second
// time), so here it is two launches: the blended values of all pixels, then
// the blur of each pixel by its own lane.
#include "hr_traverse.h"

namespace hr {

namespace {

constexpr int kBlock = 256;
using namespace trav;

struct Mat {
    f3 kd, ks;
    float se, ior;
};

__device__ __forceinline__ Mat material(const Shading& sh, uint32_t prim)
{
    const float* m = sh.mats + 12 * (size_t)(sh.mat_idx ? sh.mat_idx[prim] : 0u);
    Mat r;
    r.kd = mk(m[0], m[1], m[2]);
    r.ks = mk(m[4], m[5], m[6]);
    r.se = m[7];
    r.ior = m[3] - 10.f;   // PhongMaterial::isTransparent (Material.hpp:76-82)
    return r;
}

__device__ __forceinline__ f3 ld_rgb(const float* p, uint32_t i)
{
    return mk(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]);
}

__device__ __forceinline__ void st_rgb(float* p, uint32_t i, f3 v)
{
    p[3 * (size_t)i] = v.x;
    p[3 * (size_t)i + 1] = v.y;
    p[3 * (size_t)i + 2] = v.z;
}

// Russian roulette at an opaque vertex (PathTracer.h:489-520, :704-735):
// 0 diffuse, 1 specular, 2 absorbed (throughput already divided)
__device__ __forceinline__ int roulette(const Mat& m, float rnum, f3& att, bool at_primary)
{
    const float albedo_d = hmax(m.kd) * kPi;
    const float albedo_s = hmax(m.ks) * 2.f * kPi / (m.se + 2.f);
    if (rnum < albedo_d) {
        att = ((att * m.kd) * kPi) / albedo_d;
        return 0;
    }
    if (rnum < albedo_d + albedo_s) {
        att = ((att * m.ks) * kPi) / ((m.se + 1.f) * 0.5f * albedo_s);
        return 1;
    }
    if (at_primary)
        att = att * rep(1.f / (1.f - albedo_s - albedo_d));
    else
        att = att / (1.f - albedo_s - albedo_d);
    return 2;
}

// tracePath: each path runs from its primary hit until the roulette absorbs
// it, it leaves the scene, total internal reflection kills it or it has made
// kPathMaxBounces bounces; the end vertex (SimpleRayBuffer::store) and the
// throughput are written
template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_trace_path(TraceScene S, RenderArgs ra, uint32_t npix,
                                                       const float* __restrict__ hits,
                                                       const uint32_t* __restrict__ hit_prims, Shading sh,
                                                       PathBuffers pb, unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
    __shared__ float sub[kSubTable];
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = pix < npix;   // every lane stays for add_work
    const uint32_t image = (uint32_t)image_of(ra);
    f3 org = rep(0.f), dir = rep(0.f);
    float t = kFltMax;
    uint32_t prim = S.nt;
    if (in) {
        org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
        t = hits[6 * (size_t)npix + pix];
        prim = hit_prims[pix];
    }
    // the end vertex starts as the primary hit
    f3 e_hp = org, e_dir = dir;
    float e_t = t;
    uint32_t e_prim = prim;
    f3 att = rep(1.f);
    bool alive = t < kFltMax && prim < S.nt;
    bool diffuse = false;
    Mat m;
    uint32_t traced = 0;
    Work wk;
    GridRegs G = load_grid(S.gp);
    if (kTwo) sub_table(G, sub);
    if (alive) {
        m = material(sh, prim);
        if (!(m.ior > 0.f)) {
            float u[4];
            rand4(ra.seed_lo, ra.seed_hi, pix, image, kStreamPathMK, u);
            const int r = roulette(m, u[0], att, true);
            diffuse = r == 0;
            if (r == 2) alive = false;
        }
    }
    int depth = 0;
    while (alive) {
        float a[4], b[4];
        rand4(ra.seed_lo, ra.seed_hi, pix, image, kStreamPathMK + 2u * (uint32_t)depth, a);
        rand4(ra.seed_lo, ra.seed_hi, pix, image, kStreamPathMK + 2u * (uint32_t)depth + 1u, b);
        f3 v0, v1, v2;
        load_tri(S.rec, prim, v0, v1, v2);
        f3 n = shading_normal(sh, prim, org, v0, v1, v2);
        float dnd = -dot(n, dir);
        if (m.ior > 0.f) {
            // transparent (PathTracer.h:556-622)
            float n1 = 1.000293f, n2 = m.ior;
            if (dnd < 0.f) {
                n1 = m.ior;
                n2 = 1.000293f;
                n = -n;
                dnd = -dnd;
                att.x *= spec_expf(-t * m.kd.x * kPi);
                att.y *= spec_expf(-t * m.kd.y * kPi);
                att.z *= spec_expf(-t * m.kd.z * kPi);
            }
            const float f0 = (n1 - n2) * (n1 - n2) / ((n1 + n2) * (n1 + n2));
            const float fres = f0 + (1.f - f0) * spec_powf(1.f - fabsf(dnd), 5.f);
            const float ni = n1 / n2;
            const float sin2 = ni * ni * (1 - dnd * dnd);
            if (sin2 > 1.f || a[1] < fres) {
                org = org + (5.f * kEps) * n;
                dir = dir + (2 * dnd) * n;
                if (sin2 >= 1.f) {
                    if (a[2] < 0.5f) {
                        att = att * 2.f;
                    } else {
                        att = rep(0.f);
                        e_hp = org + (t - kEps) * dir;
                        e_dir = dir;
                        e_t = t;
                        e_prim = prim;
                        break;
                    }
                }
            } else {
                const float nf = ni * dnd - sqrtf(1.f - sin2);
                org = org - (10.f * kEps) * n;
                dir = ni * dir + nf * n;
            }
        } else if (diffuse) {
            // PathTracer.h:623-641
            if (dnd < 0.f) n = -n;
            const f3 rd = cosine_dir(b[0], b[1]);
            f3 tg, bn;
            local_frame(n, tg, bn);
            dir = tg * rd.x + bn * rd.y + n * rd.z;
            org = org + (5.f * kEps) * n;
        } else {
            // specular (PathTracer.h:642-672), PowerCosineHemisphereSampler
            // (HemisphereSamplers.hpp:78-94)
            if (dnd < 0.f) {
                n = -n;
                dnd = -dnd;
            }
            f3 rd;
            rd.z = spec_powf(b[1], 1.f / (m.se + 1.f));
            float s_, c_;
            spec_sincosf(2.f * kPi * b[0], s_, c_);
            rd.x = c_ * sqrtf(1 - rd.z * rd.z);
            rd.y = s_ * sqrtf(1 - rd.z * rd.z);
            const f3 up = dir + (2.f * dnd) * n;
            f3 rt, fw;
            local_frame(up, rt, fw);
            dir = rt * rd.x + fw * rd.y + up * rd.z;
            org = org + (5.f * kEps) * n;
            const float c = dot(dir, n);
            att = c < kEps ? att * 0.f : att * c;
        }
        // extension ray (PathTracer.h:676-700)
        t = kFltMax;
        f3 d2;
        const uint32_t hit = trace_ray<kTwo, false>(G, S, org, dir, t, 0.f, d2, wk);
        ++traced;
        dir = d2;
        if (!(t < kFltMax) || depth >= kPathMaxBounces) {
            e_hp = org + (t - kEps) * dir;
            e_dir = dir;
            e_t = t;
            e_prim = t < kFltMax ? hit : prim;
            break;
        }
        prim = hit;
        m = material(sh, prim);
        if (!(m.ior > 0.f)) {
            float u[4];
            rand4(ra.seed_lo, ra.seed_hi, pix, image, kStreamPathMK + 2u * (uint32_t)(depth + 1), u);
            const int r = roulette(m, u[0], att, false);
            diffuse = r == 0;
            if (r == 2) {
                e_hp = org + (t - kEps) * dir;
                e_dir = dir;
                e_t = t;
                e_prim = prim;
                break;
            }
        }
        ++depth;
        org = org + (t - kEps) * dir;
    }
    if (ray_count && traced) atomicAdd(ray_count, (unsigned long long)traced);
    add_work(ra.stats, traced, wk);
    if (!in) return;
    pb.end[pix] = e_hp.x;
    pb.end[(size_t)npix + pix] = e_hp.y;
    pb.end[2 * (size_t)npix + pix] = e_hp.z;
    pb.end[3 * (size_t)npix + pix] = e_dir.x;
    pb.end[4 * (size_t)npix + pix] = e_dir.y;
    pb.end[5 * (size_t)npix + pix] = e_dir.z;
    pb.end[6 * (size_t)npix + pix] = e_t;
    pb.end_prims[pix] = e_prim;
    st_rgb(pb.importance, pix, att);
}

// AreaLightSource::isOnLS (LightSource.hpp:50-76)
__device__ __forceinline__ bool on_light(f3 pos, f3 nrm, f3 e1, f3 e2, f3 p)
{
    const float eps = 0.001f * fmaxf(hmax(vabs(e1)), hmax(vabs(e2)));
    if (!(fabsf(dot(p - pos, nrm)) < eps)) return false;
    const f3 a1 = vabs(e1), a2 = vabs(e2);
    const int d1 = (a1.x > a1.y) ? ((a1.x > a1.z) ? 0 : 2) : ((a1.y > a1.z) ? 1 : 2);
    const int d2 = (a2.x > a2.y) ? ((a2.x > a2.z) ? 0 : 2) : ((a2.y > a2.z) ? 1 : 2);
    const float e1x = get(e1, d1), e1y = get(e1, d2), e2x = get(e2, d1), e2y = get(e2, d2);
    const float px = get(pos, d1), py = get(pos, d2), xx = get(p, d1), xy = get(p, d2);
    float beta = (xy - py) - (e1y / e1x) * (xx - px);
    beta /= e2y - e2x * (e1y / e1x);
    const float alpha = ((xx - px) - beta * e2x) / e1x;
    return alpha > -eps && alpha < 1.f + eps && beta > -eps && beta < 1.f + eps;
}

// AreaLightShadowRayGenerator<1,1> + trace<..., shadow> + DirectIlluminationBuffer::store
template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_path_shadow(TraceScene S, RenderArgs ra, uint32_t npix, Shading sh,
                                                        PathBuffers pb, unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
    __shared__ float sub[kSubTable];
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = pix < npix;   // every lane stays for add_work
    const float t = in ? pb.end[6 * (size_t)npix + pix] : kFltMax;
    f3 ls = rep(0.f), dl = rep(0.f);
    bool traced = false;
    Work wk;
    GridRegs G = load_grid(S.gp);
    if (kTwo) sub_table(G, sub);
    if (t < kFltMax) {
        float u[4];
        rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamPathLight, u);
        int li = sh.nlights - 1;
        for (int i = 0; i < sh.nlights - 1; ++i)
            if (u[0] < sh.light_cdf[i] / sh.light_cdf[sh.nlights - 1]) { li = i; break; }
        const float* L = sh.lights + 15 * (size_t)li;
        const f3 lpos = mk(L[0], L[1], L[2]), lnrm = mk(L[3], L[4], L[5]);
        const f3 lint = mk(L[6], L[7], L[8]), le1 = mk(L[9], L[10], L[11]), le2 = mk(L[12], L[13], L[14]);
        const f3 org = mk(pb.end[pix], pb.end[(size_t)npix + pix], pb.end[2 * (size_t)npix + pix]);
        if (on_light(lpos, lnrm, le1, le2, org)) {
            const f3 dir = mk(pb.end[3 * (size_t)npix + pix], pb.end[4 * (size_t)npix + pix],
                              pb.end[5 * (size_t)npix + pix]);
            const float cos_ln = -dot(dir, lnrm);
            ls = lint * (cos_ln > 0.f ? .5f : 0.f);
        } else {
            // StratifiedSampleGenerator<1,1> (RandomNumberGenerators.hpp:188-205)
            const float r1 = 1.f * u[1] + 0.f * 1.f;
            const float r2 = 1.f * u[2] + 0.f * 1.f;
            const f3 dir = (lpos + r1 * le1 + r2 * le2) - org;
            const float a = 1.f / dot(dir, dir);
            const f3 dn = dir * sqrtf(a);
            const float cos_ln = dot(lnrm, -dn);
            ls = lint * len(cross(le1, le2)) * a * fmaxf(0.f, cos_ln) * 0.5f;
            float st = 0.9999f;
            f3 d2;
            trace_ray<kTwo, true>(G, S, org, dir, st, 0.9999f, d2, wk);
            traced = true;
            if (st < 0.9999f) ls = rep(0.f);
            else dl = normalize(d2);
        }
    }
    count_rays(ray_count, traced);
    add_work(ra.stats, traced ? 1u : 0u, wk);
    if (!in) return;
    st_rgb(pb.ls, pix, ls);
    st_rgb(pb.dir_ls, pix, dl);
}

// computeDirectIllumination, one sample per pixel: the end vertex's direct
// light times the throughput, into the running mean new_radiance
__global__ __launch_bounds__(kBlock) void k_path_direct(TraceScene S, RenderArgs ra, uint32_t npix, Shading sh,
                                                        PathBuffers pb)
{
    if (frame_skipped(ra)) return;
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const float t = pb.end[6 * (size_t)npix + pix];
    const uint32_t prim = pb.end_prims[pix];
    f3 v = ld_rgb(pb.ls, pix);
    if (t < kFltMax) {
        if (v.x + v.y + v.z > 0.f && prim < S.nt) {
            const f3 org = mk(pb.end[pix], pb.end[(size_t)npix + pix], pb.end[2 * (size_t)npix + pix]);
            const f3 dir = mk(pb.end[3 * (size_t)npix + pix], pb.end[4 * (size_t)npix + pix],
                              pb.end[5 * (size_t)npix + pix]);
            f3 v0, v1, v2;
            load_tri(S.rec, prim, v0, v1, v2);
            const f3 n = shading_normal(sh, prim, org, v0, v1, v2);
            const Mat m = material(sh, prim);
            const f3 dl = ld_rgb(pb.dir_ls, pix);
            const float cos_nl = dot(n, dl);
            const float cos_ne = -dot(n, dir);
            const float cos_hl = fmaxf(0.f, dot(normalize(dl - dir), n));
            if (fabsf(dl.x) + fabsf(dl.y) + fabsf(dl.z) > 0.f && cos_ne > 0.f && cos_nl > 0.f) {
                const f3 lsr = v;
                v = (m.kd * lsr) * cos_nl;
                v = v + ((m.ks * fmaxf(0.f, spec_powf(cos_hl, m.se))) * lsr) * cos_nl;
            }
        }
    } else {
        v = rep(0.f);
    }
    const f3 rad = rep(0.f) + v * 1.f / (float)1;
    const float nw = 1.f / (float)(image_of(ra) + 1);
    const float ow = 1.f - nw;
    const f3 att = ld_rgb(pb.importance, pix);
    st_rgb(pb.new_radiance, pix, ld_rgb(pb.new_radiance, pix) * ow + rad * att * nw);
}

// addIndirectIllumination, first half: each pixel's contribution and its
// blended value
__global__ __launch_bounds__(kBlock) void k_path_blend(RenderArgs ra, uint32_t npix, PathBuffers pb,
                                                       const float* __restrict__ fb)
{
    if (frame_skipped(ra)) return;
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const float nw = 1.f / (float)(image_of(ra) + 1);
    const float ow = 1.f - nw;
    f3 nc = ld_rgb(pb.importance, pix) * ld_rgb(pb.new_radiance, pix);
    nc = nc + ld_rgb(pb.residue, pix);
    st_rgb(pb.contrib, pix, nc);
    st_rgb(pb.blended, pix, ld_rgb(fb, pix) * ow + nc * nw);
}

// second half: the 3x3 blur (with the reference's weights: its centre tap
// reads the pixel below, PathTracer.h:954-968) where the contribution is
// bright, the residue keeping what the blur removed
__global__ __launch_bounds__(kBlock) void k_path_merge(RenderArgs ra, int resx, int resy, PathBuffers pb,
                                                       float* __restrict__ fb)
{
    if (frame_skipped(ra)) return;
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (uint32_t)resx * (uint32_t)resy) return;
    const int x = (int)(pix % (uint32_t)resx), y = (int)(pix / (uint32_t)resx);
    const int xl = x > 0 ? x - 1 : 0, xr = x + 1 < resx ? x + 1 : resx - 1;
    const int yb = y > 0 ? y - 1 : 0, yt = y + 1 < resy ? y + 1 : resy - 1;
    auto at = [&](int px, int py) { return ld_rgb(pb.blended, (uint32_t)py * (uint32_t)resx + (uint32_t)px); };
    const int image = image_of(ra);
    const float nw = 1.f / (float)(image + 1);
    const float ow = 1.f - nw;
    const f3 pv = ld_rgb(fb, pix);
    const f3 nc = ld_rgb(pb.contrib, pix);
    const f3 top = at(x, yt);
    const f3 blurred = (1.f * at(xl, yt) + 2.f * top + 1.f * at(xr, yt) + 2.f * at(xl, y) + 4.f * top +
                        2.f * at(xr, y) + 1.f * at(xl, yb) + 2.f * at(x, yb) + 1.f * at(xr, yb)) /
                       16.f;
    float variance = dot(nc, rep(1.f));
    variance *= variance;
    if (variance < fmaxf(0.25f / (float)image, 0.05f)) {
        st_rgb(fb, pix, pv * ow + nc * nw);
        st_rgb(pb.residue, pix, rep(0.f));
    } else {
        const f3 val = vmin(blurred + 0.5f * nc, nc);
        st_rgb(fb, pix, pv * ow + val * nw);
        st_rgb(pb.residue, pix, nc - val);
    }
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_path_multi(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, const float* hits,
                       const uint32_t* hit_prims, const Shading& sh, const PathBuffers& pb, float* fb,
                       unsigned long long* ray_count, hipStream_t s, hipEvent_t after_trace)
{
    const uint32_t npix = (uint32_t)cam.res_x * (uint32_t)cam.res_y;
    const dim3 g(blocks_for(npix)), b(kBlock);
    if (S.two_level) {
        hipLaunchKernelGGL(k_trace_path<true>, g, b, 0, s, S, ra, npix, hits, hit_prims, sh, pb, ray_count);
        hipLaunchKernelGGL(k_path_shadow<true>, g, b, 0, s, S, ra, npix, sh, pb, ray_count);
    } else {
        hipLaunchKernelGGL(k_trace_path<false>, g, b, 0, s, S, ra, npix, hits, hit_prims, sh, pb, ray_count);
        hipLaunchKernelGGL(k_path_shadow<false>, g, b, 0, s, S, ra, npix, sh, pb, ray_count);
    }
    if (after_trace) (void)hipEventRecord(after_trace, s);
    hipLaunchKernelGGL(k_path_direct, g, b, 0, s, S, ra, npix, sh, pb);
    hipLaunchKernelGGL(k_path_blend, g, b, 0, s, ra, npix, pb, fb);
    hipLaunchKernelGGL(k_path_merge, g, b, 0, s, ra, cam.res_x, cam.res_y, pb, fb);
}

}  // namespace hr
