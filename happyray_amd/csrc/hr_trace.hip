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
//   simpleShade                      src/RT/Integrator/SimpleIntegrator.h:42-125
//   dumpRays                         src/RT/Integrator/AORayExporter.h:50-68
//   uploadObjFrameNormalData         src/Application/WFObjectUploader.h:121-165
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
#include "hr_flat.h"
#include "hr_traverse.h"

namespace hr {

namespace {

constexpr int kBlock = 256;
using namespace trav;

// ---------------------------------------------------------------------------
// primary rays
// ---------------------------------------------------------------------------

// one camera ray per lane: slot q of the 8x8 pixel tiles; false for a slot
// past the image's edge
template <bool kTwo>
__device__ __forceinline__ bool camera_ray(const TraceScene& S, const GridRegs& G, const CameraDev& cam,
                                           const RenderArgs& ra, uint32_t npix, uint32_t q, float* __restrict__ hits,
                                           uint32_t* __restrict__ hit_prims, unsigned long long* ray_count, Work& wk)
{
    uint32_t pix;
    const bool active = tiled_pixel(q, (uint32_t)cam.res_x, (uint32_t)cam.res_y, pix);
    count_rays(ray_count, active);
    if (!active) return false;
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
    return true;
}

// Camera rays, one per lane, in 8x8 pixel tiles (a wave's 64 rays cover one
// tile).  tiles == 0: workgroup b takes the four tiles of slots 256 b ..;
// otherwise the grid is a few workgroups per CU whose waves take tiles from
// the counter *next (zero at launch) until all `tiles` are taken, so the
// kernel leaves wave slots to work on other streams (a neighbouring frame's
// rebuild or secondary rays).  Six waves per SIMD (at most 80 VGPRs).
template <bool kTwo>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_primary(
    TraceScene S, CameraDev cam, RenderArgs ra, uint32_t npix, float* __restrict__ hits,
    uint32_t* __restrict__ hit_prims, unsigned long long* ray_count, uint32_t tiles, uint32_t* __restrict__ next)
{
    if (frame_skipped(ra)) return;
    __shared__ float sub[kSubTable];
    Work wk;
    GridRegs G = load_grid(S.gp);
    if (kTwo) sub_table(G, sub);
    uint32_t traced = 0;
    if (tiles == 0) {
        traced = camera_ray<kTwo>(S, G, cam, ra, npix, blockIdx.x * blockDim.x + threadIdx.x, hits, hit_prims,
                                  ray_count, wk) ? 1u : 0u;
    } else {
        while (true) {
            uint32_t tile = 0;
            if (__lane_id() == 0) tile = atomicAdd(next, 1u);
            tile = __shfl(tile, 0, kWave);
            if (tile >= tiles) break;
            traced += camera_ray<kTwo>(S, G, cam, ra, npix, tile * kWave + __lane_id(), hits, hit_prims, ray_count,
                                       wk) ? 1u : 0u;
        }
    }
    add_work(ra.stats, traced, wk);
}

// camera rays through trace_wave: one ray per lane as in k_primary, with the
// wave's candidate tests batched (a lane whose ray has ended lends its slot to
// the others' tests); the same hits as k_primary
template <bool kTwo>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_primary_wave(
    TraceScene S, CameraDev cam, RenderArgs ra, uint32_t npix, float* __restrict__ hits,
    uint32_t* __restrict__ hit_prims, uint32_t steps_max, unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
    __shared__ WaveTests<kBatchK> lds[kBlock / kWave];
    WaveTests<kBatchK>& W = lds[__builtin_amdgcn_readfirstlane((int)threadIdx.x) / kWave];
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;   // row order: tiles measured slower here
    const bool active = pix < npix;
    count_rays(ray_count, active);
    Work wk;
    const GridRegs G = load_grid(S.gp);
    f3 d0 = rep(0.f);
    if (active) d0 = primary_direction(cam, ra, pix, npix);
    const f3 org = mk(cam.center[0], cam.center[1], cam.center[2]);
    float t = kFltMax;
    f3 dir;
    const uint32_t best = trace_wave<kTwo, false>(G, S, W, active, org, d0, t, dir, steps_max, wk);
    if (active) {
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

template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_ao_rays(TraceScene S, RenderArgs ra, uint32_t npix,
                                                    const float* __restrict__ hits,
                                                    const uint32_t* __restrict__ hit_prims,
                                                    uint8_t* __restrict__ occ,
                                                    unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
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
    __shared__ float sub[kSubTable];
    GridRegs G = load_grid(S.gp);
    if (kTwo) sub_table(G, sub);
    if (active) {
    // AmbientOcclusionRayGenerator::operator() (RayGenerators.h:281-343)
    f3 org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
    const f3 pdir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix],
                       hits[5 * (size_t)npix + pix]);
    f3 v0, v1, v2;
    load_tri(S.rec, prim, v0, v1, v2);
    f3 n = normalize(cross(v1 - v0, v2 - v0));
    if (dot(n, pdir) > 0.f) n = -n;
    float u[4];
    rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamAO + s, u);
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
    if (frame_skipped(ra)) return;
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
    fb_put(ra, fb + 3 * (size_t)pix, rad.x, rad.y, rad.z);
}

// ---------------------------------------------------------------------------
// path tracing megakernel (PathTracer.h:41-372), one lane per pixel
// ---------------------------------------------------------------------------

// k_pt with the wave-cooperative trace: the same path loop, with the depth
// iterations run in step by the whole wave (a finished path sits out the rest)
// so that each shadow and extension ray goes through trace_wave, whose
// candidate tests run in full-width batches.  The bounce of an opaque vertex
// is computed before its shadow ray is traced (it reads nothing the shadow ray
// decides, and the light term keeps the vertex's own last_spec), so only the
// light term waits across the trace.  Per path the operations and their order
// are k_pt's: the same image bit for bit.
template <bool kTwo>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_pt_wave(TraceScene S, RenderArgs ra, uint32_t npix,
                                                    const float* __restrict__ hits,
                                                    const uint32_t* __restrict__ hit_prims, Shading sh,
                                                    float* __restrict__ fb, uint32_t steps_max,
                                                    unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
    __shared__ WaveTests<kBatchK> lds[kBlock / kWave];
    __shared__ float light_terms[6][kBlock];   // a vertex's light terms wait here across its shadow ray
    WaveTests<kBatchK>& W = lds[__builtin_amdgcn_readfirstlane((int)threadIdx.x) / kWave];
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = pix < npix;
    const GridRegs G = load_grid(S.gp);
    PathState P;
    P.radiance = rep(0.f);
    P.ray_t = in ? hits[6 * (size_t)npix + pix] : kFltMax;
    P.prim = in ? hit_prims[pix] : S.nt;
    uint32_t traced = 0;
    Work wk;
    bool alive = P.ray_t < kFltMax && P.prim < S.nt;
    if (alive) {
        P.org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        P.dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
        P.att = rep(1.f);
        P.last_spec = true;
    }
    for (int depth = 0; depth < ra.pt_depth; ++depth) {
        if (__ballot(alive) == 0ull) break;
        PathVertex V;
        bool shadow = false;
        bool light_spec = true;
        if (alive) {
            P.depth = depth;
            pt_vertex(S, sh, ra, pix, P, V);
            if (V.opaque) {
                shadow = V.shadow;
                if (shadow) {
                    light_terms[0][threadIdx.x] = V.add_d.x;
                    light_terms[1][threadIdx.x] = V.add_d.y;
                    light_terms[2][threadIdx.x] = V.add_d.z;
                    light_terms[3][threadIdx.x] = V.add_s.x;
                    light_terms[4][threadIdx.x] = V.add_s.y;
                    light_terms[5][threadIdx.x] = V.add_s.z;
                }
                light_spec = P.last_spec;
                if (depth + 1 >= ra.pt_depth) alive = false;
                else pt_bounce(P, V);   // leaves P.org, the shadow ray's origin, as it is
            }
            if (alive && (V.end || depth + 1 >= ra.pt_depth)) alive = false;
        }
        // the shadow ray (PathTracer.h:236-258) and its light term
        float st = 0.9999f;
        f3 dd;
        trace_wave<kTwo, true>(G, S, W, shadow, P.org, V.sdir, st, dd, steps_max, wk);
        if (shadow) {
            ++traced;
            if (!(st < 0.9999f)) {
                P.radiance = P.radiance + mk(light_terms[0][threadIdx.x], light_terms[1][threadIdx.x],
                                             light_terms[2][threadIdx.x]);
                if (!light_spec)
                    P.radiance = P.radiance + mk(light_terms[3][threadIdx.x], light_terms[4][threadIdx.x],
                                                 light_terms[5][threadIdx.x]);
            }
        }
        // the extension ray (PathTracer.h:318-332)
        float t = kFltMax;
        f3 d2;
        const uint32_t hit = trace_wave<kTwo, false>(G, S, W, alive, P.org, P.dir, t, d2, steps_max, wk);
        if (alive) {
            ++traced;
            P.dir = d2;
            if (!(t < kFltMax)) {
                alive = false;
            } else {
                P.org = P.org + (t - kEps) * P.dir;
                P.prim = hit;
                P.ray_t = t;
            }
        }
    }
    if (ray_count && traced) atomicAdd(ray_count, (unsigned long long)traced);
    add_work(ra.stats, traced, wk);
    if (!in) return;
    fb_put(ra, fb + 3 * (size_t)pix, P.radiance.x, P.radiance.y, P.radiance.z);
}

template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_pt(TraceScene S, RenderArgs ra, uint32_t npix,
                                               const float* __restrict__ hits,
                                               const uint32_t* __restrict__ hit_prims, Shading sh,
                                               float* __restrict__ fb, unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
    __shared__ float sub[kSubTable];
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = pix < npix;
    GridRegs G = load_grid(S.gp);
    if (kTwo) sub_table(G, sub);
    PathState P;
    P.radiance = rep(0.f);
    P.ray_t = in ? hits[6 * (size_t)npix + pix] : kFltMax;
    P.prim = in ? hit_prims[pix] : S.nt;
    uint32_t traced = 0;
    Work wk;
    if (P.ray_t < kFltMax && P.prim < S.nt) {
        P.org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        P.dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
        P.att = rep(1.f);
        P.last_spec = true;
        for (P.depth = 0; P.depth < ra.pt_depth; ++P.depth) {
            PathVertex V;
            pt_vertex(S, sh, ra, pix, P, V);
            if (V.opaque) {
                if (V.shadow) {
                    float st = 0.9999f;
                    f3 dd;
                    trace_ray<kTwo, true>(G, S, P.org, V.sdir, st, 0.9999f, dd, wk);
                    ++traced;
                    pt_light(P, V, st);
                }
                if (P.depth + 1 >= ra.pt_depth) break;
                pt_bounce(P, V);
            }
            if (V.end || P.depth + 1 >= ra.pt_depth) break;
            // extension ray (PathTracer.h:318-332)
            float t = kFltMax;
            f3 d2;
            const uint32_t hit = trace_ray<kTwo, false>(G, S, P.org, P.dir, t, 0.f, d2, wk);
            ++traced;
            P.dir = d2;
            if (!(t < kFltMax)) break;
            P.org = P.org + (t - kEps) * P.dir;
            P.prim = hit;
            P.ray_t = t;
        }
    }
    if (ray_count && traced) atomicAdd(ray_count, (unsigned long long)traced);
    add_work(ra.stats, traced, wk);
    if (!in) return;
    fb_put(ra, fb + 3 * (size_t)pix, P.radiance.x, P.radiance.y, P.radiance.z);
}

// ---------------------------------------------------------------------------
// simple integrator (SimpleIntegrator.h:42-125), one lane per pixel
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_simple_shade(TraceScene S, RenderArgs ra, uint32_t npix,
                                                         const float* __restrict__ hits,
                                                         const uint32_t* __restrict__ hit_prims,
                                                         float* __restrict__ fb)
{
    if (frame_skipped(ra)) return;
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const float t = hits[6 * (size_t)npix + pix];
    const uint32_t prim = hit_prims[pix];
    f3 rad;
    if (t < kFltMax && prim < S.nt) {
        f3 v0, v1, v2;
        load_tri(S.rec, prim, v0, v1, v2);
        const f3 org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        const f3 dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix],
                          hits[5 * (size_t)npix + pix]);
        const f3 real_normal = cross(v1 - v0, v2 - v0);
        const f3 a = v0 - org, b = v1 - org, c = v2 - org;
        const f3 n0 = cross(b, c), n1 = cross(c, a);
        const float rcp = 1.f / sqrtf(dot(real_normal, real_normal));
        const float u = len(n0) * rcp, v = len(n1) * rcp;
        const f3 n = normalize(real_normal);
        const float w = fmaxf(0.f, fabsf(dot(-n, normalize(dir))));
        rad = mk(u, v, 1.f - u - v) * w;
    } else {
        rad = mk(0.15f, 0.2f, 0.3f);
    }
    fb_put(ra, fb + 3 * (size_t)pix, rad.x, rad.y, rad.z);
}

// ---------------------------------------------------------------------------
// ray records
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_wire_to_soa(const float* __restrict__ rec, uint32_t n,
                                                        float* __restrict__ org, float* __restrict__ dir,
                                                        uint32_t stride)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rec + 6 * (size_t)i;
    org[i] = r[0];
    org[(size_t)stride + i] = r[1];
    org[2 * (size_t)stride + i] = r[2];
    dir[i] = r[3];
    dir[(size_t)stride + i] = r[4];
    dir[2 * (size_t)stride + i] = r[5];
}

// dumpRays<PrimaryGenerator> then dumpRays<AmbientOcclusionRayGenerator>
// (AORayExporter.h:50-68): record j < npix is pixel j's camera ray; record
// npix + j is AO ray j (pixel j / S, sample j % S)
__global__ __launch_bounds__(kBlock) void k_ao_dump(TraceScene S, CameraDev cam, RenderArgs ra, uint32_t npix,
                                                    const float* __restrict__ hits,
                                                    const uint32_t* __restrict__ hit_prims,
                                                    float* __restrict__ out)
{
    const uint32_t S_ = (uint32_t)ra.ao_samples;
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (uint64_t)npix * (S_ + 1u)) return;
    f3 org, dir;
    if (j < npix) {
        org = mk(cam.center[0], cam.center[1], cam.center[2]);
        dir = primary_direction(cam, ra, (uint32_t)j, npix);
    } else {
        const uint32_t a = (uint32_t)(j - npix);
        const uint32_t pix = a / S_, s = a % S_;
        // AmbientOcclusionRayGenerator::operator() (RayGenerators.h:272-344)
        org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
        dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix], hits[5 * (size_t)npix + pix]);
        const float th = hits[6 * (size_t)npix + pix];
        const uint32_t prim = hit_prims[pix];
        if (th < kFltMax && prim < S.nt) {
            f3 v0, v1, v2;
            load_tri(S.rec, prim, v0, v1, v2);
            f3 n = normalize(cross(v1 - v0, v2 - v0));
            if (dot(n, dir) > 0.f) n = -n;
            float u[4];
            rand4(ra.seed_lo, ra.seed_hi, pix, (uint32_t)image_of(ra), kStreamAO + s, u);
            const f3 h = cosine_dir(u[0], u[1]);
            f3 tg, bn;
            local_frame(n, tg, bn);
            dir = normalize(n * h.z + tg * h.x + bn * h.y);
            org = org + (0.01f * ra.ao_dist) * dir;
        }
    }
    float* o = out + 6 * j;
    o[0] = org.x;
    o[1] = org.y;
    o[2] = org.z;
    o[3] = dir.x;
    o[4] = dir.y;
    o[5] = dir.z;
}

// ---------------------------------------------------------------------------
// per-corner normals
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_frame_normals(const float* __restrict__ a,
                                                          const float* __restrict__ b, uint32_t nn,
                                                          const FrameSel* __restrict__ fs,
                                                          const float* __restrict__ keys,
                                                          float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    if (fs) {
        a = keys + (size_t)fs->k1 * 3 * nn;
        b = keys + (size_t)fs->k2 * 3 * nn;
    }
    const float c = 0.f;
    const float x = a[i] * (1.f - c) + b[i] * c;
    const float y = a[(size_t)nn + i] * (1.f - c) + b[(size_t)nn + i] * c;
    const float z = a[2 * (size_t)nn + i] * (1.f - c) + b[2 * (size_t)nn + i] * c;
    out[i] = make_float4(x, y, z, 0.f);
}

// clears the framebuffer when the captured frame is the first image of a
// running mean (the uncaptured path clears it from the host)
__global__ __launch_bounds__(kBlock) void k_clear_first(float* __restrict__ fb, size_t n,
                                                        const FrameSel* __restrict__ fs)
{
    if (fs->image != 0 || fs->skip) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        fb[i] = 0.f;
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
    __shared__ float sub[kSubTable];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    count_rays(ray_count, i < n);
    Work wk;
    GridRegs G = load_grid(S.gp);
    if (kTwo) sub_table(G, sub);
    if (i < n) {
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

}  // namespace

void launch_primary(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, float* hits,
                    uint32_t* hit_prims, unsigned long long* ray_count, hipStream_t s, uint32_t wave_steps,
                    uint32_t blocks_per_cu, uint32_t* next)
{
    const uint32_t npix = (uint32_t)cam.res_x * (uint32_t)cam.res_y;
    if (wave_steps) {
        if (S.two_level)
            hipLaunchKernelGGL(k_primary_wave<true>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, cam, ra, npix,
                               hits, hit_prims, wave_steps, ray_count);
        else
            hipLaunchKernelGGL(k_primary_wave<false>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, cam, ra, npix,
                               hits, hit_prims, wave_steps, ray_count);
        return;
    }
    const uint32_t slots = tiled_slots((uint32_t)cam.res_x, (uint32_t)cam.res_y);
    uint32_t blocks = blocks_for(slots), tiles = 0;
    if (blocks_per_cu > 0 && next && (size_t)blocks_per_cu * device_cus() < blocks) {
        blocks = blocks_per_cu * device_cus();
        tiles = slots / kWave;
        fill_u32(next, 0u, 1, nullptr, 1, s);
    }
    if (S.two_level)
        hipLaunchKernelGGL(k_primary<true>, dim3(blocks), dim3(kBlock), 0, s, S, cam, ra, npix,
                           hits, hit_prims, ray_count, tiles, next);
    else
        hipLaunchKernelGGL(k_primary<false>, dim3(blocks), dim3(kBlock), 0, s, S, cam, ra, npix,
                           hits, hit_prims, ray_count, tiles, next);
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
               unsigned long long* ray_count, hipStream_t s, uint32_t wave_steps)
{
    if (wave_steps) {
        if (S.two_level)
            hipLaunchKernelGGL(k_pt_wave<true>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits,
                               hit_prims, sh, fb, wave_steps, ray_count);
        else
            hipLaunchKernelGGL(k_pt_wave<false>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits,
                               hit_prims, sh, fb, wave_steps, ray_count);
        return;
    }
    if (S.two_level)
        hipLaunchKernelGGL(k_pt<true>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits,
                           hit_prims, sh, fb, ray_count);
    else
        hipLaunchKernelGGL(k_pt<false>, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits,
                           hit_prims, sh, fb, ray_count);
}

void launch_simple_shade(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                         const uint32_t* hit_prims, float* fb, hipStream_t s)
{
    hipLaunchKernelGGL(k_simple_shade, dim3(blocks_for(npix)), dim3(kBlock), 0, s, S, ra, npix, hits, hit_prims,
                       fb);
}

void launch_wire_to_soa(const float* rec, uint32_t n, float* org, float* dir, uint32_t stride, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_wire_to_soa, dim3(blocks_for(n)), dim3(kBlock), 0, s, rec, n, org, dir, stride);
}

void launch_ao_dump(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, const float* hits,
                    const uint32_t* hit_prims, float* out, hipStream_t s)
{
    const uint32_t npix = (uint32_t)cam.res_x * (uint32_t)cam.res_y;
    const size_t n = (size_t)npix * ((size_t)ra.ao_samples + 1);
    hipLaunchKernelGGL(k_ao_dump, dim3(blocks_for(n)), dim3(kBlock), 0, s, S, cam, ra, npix, hits, hit_prims, out);
}

__global__ void k_frame_gate(const BuildStatus* __restrict__ st, uint32_t* __restrict__ halt,
                             FrameSel* __restrict__ fs, unsigned long long* __restrict__ frame_rays)
{
    if (threadIdx.x != 0) return;
    const uint32_t seq = fs->seq;
    const uint32_t h = st->overflow != 0u ? atomicMin(halt, seq) : atomicAdd(halt, 0u);
    fs->skip = (h <= seq || st->overflow != 0u) ? 1u : 0u;
    if (frame_rays) *frame_rays = 0ull;
}

void launch_frame_gate(const BuildStatus* st, uint32_t* halt, FrameSel* fs, unsigned long long* frame_rays,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_frame_gate, dim3(1), dim3(64), 0, s, st, halt, fs, frame_rays);
}

__global__ __launch_bounds__(kBlock) void k_frame_blend(const FrameSel* __restrict__ fs,
                                                        const uint32_t* __restrict__ halt,
                                                        const float* __restrict__ rad, float* __restrict__ fb,
                                                        uint32_t npix, const unsigned long long* frame_rays,
                                                        unsigned long long* ray_count)
{
    const uint32_t seq = fs->seq;
    if (__hip_atomic_load(halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= seq) return;
    const int image = fs->image;
    const float nw = 1.f / (float)(image + 1);
    const float ow = 1.f - nw;
    const size_t n = 3 * (size_t)npix;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float f = image == 0 ? 0.f : fb[i];
        fb[i] = f * ow + rad[i] * nw;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(ray_count, *frame_rays);
}

void launch_frame_blend(const FrameSel* fs, const uint32_t* halt, const float* rad, float* fb, uint32_t npix,
                        const unsigned long long* frame_rays, unsigned long long* ray_count, hipStream_t s)
{
    const size_t n = 3 * (size_t)npix;
    const size_t want = (n + kBlock - 1) / kBlock;
    const uint32_t blocks = (uint32_t)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(k_frame_blend, dim3(blocks), dim3(kBlock), 0, s, fs, halt, rad, fb, npix, frame_rays,
                       ray_count);
}

void launch_clear_first(float* fb, size_t n, const FrameSel* fs, hipStream_t s)
{
    uint32_t blocks = blocks_for(n > 0 ? n : 1);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_clear_first, dim3(blocks), dim3(kBlock), 0, s, fb, n, fs);
}

void launch_frame_normals(const float* a, const float* b, uint32_t nn, float4* out, hipStream_t s,
                          const FrameSel* fs, const float* keys)
{
    if (nn == 0) return;
    hipLaunchKernelGGL(k_frame_normals, dim3(blocks_for(nn)), dim3(kBlock), 0, s, a, b, nn, fs, keys, out);
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
