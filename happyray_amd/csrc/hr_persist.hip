// hr_persist.hip -- persistent traversal kernels with dynamic ray fetch.
//
// The reference's trace<> kernel uses a persistent warp pool only on old
// architectures and otherwise a grid-stride loop with one ray per thread
// (src/RT/Algorithm/RayTracingKernels.h:65-106); its traversers nest a cell
// loop, a leaf loop and a candidate loop (src/RT/Algorithm/TLGridTraverser.h:
// 132-350, RayTriangleIntersector.h:52-90), so a wave waits for its slowest
// lane at every level.  Here each lane runs a flat state machine in which one
// loop iteration does one unit of work -- one ray/triangle test, or one DDA
// advance that loads the next cell or leaf range -- and a wave refills idle
// lanes with new rays from a global counter whenever at least `refill` of its
// 64 lanes are idle.  The operations applied to each ray, and their order,
// are those of the reference traversal (same cells, same candidates, same
// termination tests), so results are bit-identical to the one-ray-per-lane
// kernels and to the CPU oracle.
#include "hr_flat.h"
#include "hr_persist.h"

namespace hr {

namespace {

using namespace trav;
constexpr int kBlock = 256;
// ---------------------------------------------------------------------------
// ray generators / sinks
// ---------------------------------------------------------------------------

struct GenPrimary {
    CameraDev cam;
    RenderArgs ra{};
    uint32_t npix;
    float* hits;
    uint32_t* hit_prims;

    __device__ bool skipped() const { return frame_skipped(ra); }

    __device__ bool start(uint32_t j, RayState& r, const TraceScene&) const
    {
        const f3 d0 = primary_direction(cam, ra, j, npix);
        set_ray(r, mk(cam.center[0], cam.center[1], cam.center[2]), d0);
        r.t = kFltMax;
        r.tterm = 0.f;
        return true;
    }
    __device__ void finish(uint32_t j, const RayState& r, bool started, const TraceScene& S) const
    {
        (void)started;
        const uint32_t prim = r.t < kFltMax ? r.best : S.nt;
        const f3 hp = r.org + (r.t - kEps) * r.dir;
        hits[j] = hp.x;
        hits[(size_t)npix + j] = hp.y;
        hits[2 * (size_t)npix + j] = hp.z;
        hits[3 * (size_t)npix + j] = r.dir.x;
        hits[4 * (size_t)npix + j] = r.dir.y;
        hits[5 * (size_t)npix + j] = r.dir.z;
        hits[6 * (size_t)npix + j] = r.t;
        hit_prims[j] = prim;
    }
};

struct GenAO {
    RenderArgs ra{};
    uint32_t npix;
    const float* hits;
    const uint32_t* hit_prims;
    uint8_t* occ;

    __device__ bool skipped() const { return frame_skipped(ra); }

    // AmbientOcclusionRayGenerator::operator() (RayGenerators.h:281-343);
    // rays in pixel row order (the tiled order measured slower here)
    __device__ bool start(uint32_t j, RayState& r, const TraceScene& S) const
    {
        const uint32_t ns = (uint32_t)ra.ao_samples;
        const uint32_t pix = j / ns, s = j % ns;
        const uint32_t prim = hit_prims[pix];
        const float th = hits[6 * (size_t)npix + pix];
        if (!(th < kFltMax) || prim >= S.nt) return false;
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
        set_ray(r, org, dir);
        r.t = ra.ao_dist;
        r.tterm = ra.ao_dist;
        return true;
    }
    __device__ void finish(uint32_t j, const RayState& r, bool started, const TraceScene&) const
    {
        occ[j] = (!started || r.t < ra.ao_dist) ? 1u : 0u;
    }
};

// the primary-hit record of a finished closest-hit ray (SimpleRayBuffer::store,
// RayBuffers.h:62-87): hit point org + (t - EPS) dir, direction, t
__device__ __forceinline__ void store_hit(float* hits, uint32_t n, uint32_t j, const RayState& r)
{
    const f3 hp = r.org + (r.t - kEps) * r.dir;
    hits[j] = hp.x;
    hits[(size_t)n + j] = hp.y;
    hits[2 * (size_t)n + j] = hp.z;
    hits[3 * (size_t)n + j] = r.dir.x;
    hits[4 * (size_t)n + j] = r.dir.y;
    hits[5 * (size_t)n + j] = r.dir.z;
    hits[6 * (size_t)n + j] = r.t;
}

struct GenBatch {
    const float* org;
    const float* dir;
    uint32_t n;
    float tmax;
    int any;
    float* t_out;
    uint32_t* prim_out;
    uint8_t* occ_out;
    float* hits;

    __device__ bool skipped() const { return false; }
    __device__ bool start(uint32_t j, RayState& r, const TraceScene&) const
    {
        set_ray(r, mk(org[j], org[(size_t)n + j], org[2 * (size_t)n + j]),
                mk(dir[j], dir[(size_t)n + j], dir[2 * (size_t)n + j]));
        r.t = any ? tmax : kFltMax;
        r.tterm = tmax;
        return true;
    }
    __device__ void finish(uint32_t j, const RayState& r, bool, const TraceScene& S) const
    {
        if (any) {
            occ_out[j] = r.t < tmax ? 1u : 0u;
            if (t_out) t_out[j] = r.t;
        } else {
            t_out[j] = r.t;
            prim_out[j] = r.t < kFltMax ? r.best : S.nt;
            if (hits) store_hit(hits, n, j, r);
        }
    }
};

// ---------------------------------------------------------------------------
// the persistent kernel
// ---------------------------------------------------------------------------

// at most 64 VGPRs, so eight waves per SIMD hide the gathers' latency
// (AO 3.19 -> 3.12 ms against the compiler's 70 VGPRs and seven waves).  The
// leaf cell sizes of a top-cell visit are divided here, not read from the
// LDS table of the one-ray-per-lane kernels: under the 64-register cap the
// table's address spills (8 -> 40..120 bytes of scratch per lane, AO
// 3.02 -> 3.06..4.08 ms).
// kFuse: a lane whose cell or leaf has no candidates left advances and, when
// the new one has candidates, tests the first of them in the same iteration
// (one or two units per iteration instead of one: AO 3.08 -> 3.01 ms on the
// 1M-triangle frame).  Every ray runs the same units in the same order
// either way.
template <bool kTwo, bool kAny, class Gen, bool kFuse = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_persist(
    TraceScene S, Gen gen, uint32_t nrays, uint32_t* __restrict__ pool, uint32_t refill,
    unsigned long long* ray_count, unsigned long long* stats)
{
    if (gen.skipped()) return;
    const GridRegs G = load_grid(S.gp);
    const uint64_t lt = __lanemask_lt();
    RayState r;
    Work wk;
    uint32_t j = 0, started = 0;
    bool live = false;
    bool exhausted = false;
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
                    const bool ok = gen.start(j, r, S);
                    started += ok ? 1u : 0u;
                    live = ok && ray_begin<kTwo, kAny>(G, S, r, wk);
                    if (!live) gen.finish(j, r, ok, S);
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (live) {
            bool done = false;
            if (kFuse) {
                if (r.it >= r.end)
                    done = kTwo ? !advance_two_level<kAny>(G, S, r, wk) : !advance_uniform<kAny>(G, S, r, wk);
                if (!done && r.it < r.end) done = test_one<kAny>(S, r, wk);
            } else {
                if (r.it < r.end) done = test_one<kAny>(S, r, wk);
                else done = kTwo ? !advance_two_level<kAny>(G, S, r, wk) : !advance_uniform<kAny>(G, S, r, wk);
            }
            if (done) {
                live = false;
                gen.finish(j, r, true, S);
            }
        }
    }
    if (ray_count) {
        unsigned long long c = started;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (ray_count && __lane_id() == 0 && c) atomicAdd(ray_count, c);
    }
    add_work(stats, started, wk);
}

template <bool kTwo, bool kAny, class Gen, int kK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_persist_bt(
    TraceScene S, Gen gen, uint32_t nrays, uint32_t* __restrict__ pool, uint32_t refill, uint32_t steps_max,
    unsigned long long* ray_count, unsigned long long* stats)
{
    if (gen.skipped()) return;
    __shared__ WaveTests<kK> lds[kBlock / kWave];
    WaveTests<kK>& W = lds[__builtin_amdgcn_readfirstlane((int)threadIdx.x) / kWave];
    const GridRegs G = load_grid(S.gp);
    const uint32_t lane = __lane_id();
    const uint64_t lt = __lanemask_lt();
    RayState r;
    Work wk;
    uint32_t j = 0, started = 0;
    bool live = false;
    bool exhausted = false;
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
                    const bool ok = gen.start(j, r, S);
                    started += ok ? 1u : 0u;
                    live = ok && ray_begin<kTwo, kAny>(G, S, r, wk);
                    if (live) {
                        W.ox[lane] = r.org.x;
                        W.oy[lane] = r.org.y;
                        W.oz[lane] = r.org.z;
                        W.dx[lane] = r.dir.x;
                        W.dy[lane] = r.dir.y;
                        W.dz[lane] = r.dir.z;
                    } else {
                        gen.finish(j, r, ok, S);
                    }
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (exhausted) break;
            continue;
        }
        bool done = false;
        if (live && r.it >= r.end) {
            for (uint32_t m = 0; m < steps_max; ++m) {
                if (kTwo) {
                    bool top = true;
                    if (r.in_leaf) {
                        const int k = leaf_step<kAny>(S, r, wk);
                        if (k != 2) {
                            top = false;
                            done = k != 1;
                        }
                    }
                    if (top) done = !top_step(G, S, r, mk(W.ox[lane], W.oy[lane], W.oz[lane]), wk);
                } else {
                    done = !advance_uniform<kAny>(G, S, r, wk);
                }
                if (done || r.it < r.end) break;
            }
        }
        if (wave_test_batch<kAny>(S, W, r, lane, live && !done, wk)) done = true;
        if (live && done) {
            live = false;
            r.org = mk(W.ox[lane], W.oy[lane], W.oz[lane]);
            r.dir = mk(W.dx[lane], W.dy[lane], W.dz[lane]);
            gen.finish(j, r, true, S);
        }
    }
    if (ray_count) {
        unsigned long long c = started;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (lane == 0 && c) atomicAdd(ray_count, c);
    }
    add_work(stats, started, wk);
}

// The multi-scene kernel: the loop of k_persist with a per-lane scene.  A
// lane that starts a ray finds its batch by binary search over the ray
// prefix, then loads that scene's descriptor and grid parameters.
template <bool kAny>
__global__ __launch_bounds__(kBlock) void k_persist_multi(const SceneDesc* __restrict__ scenes,
                                                          const RayBatchDev* __restrict__ batches,
                                                          const uint32_t* __restrict__ prefix, uint32_t nb,
                                                          uint32_t nrays, uint32_t* __restrict__ pool,
                                                          uint32_t refill, unsigned long long* ray_count,
                                                          unsigned long long* stats)
{
    const uint64_t lt = __lanemask_lt();
    RayState r;
    Work wk;
    TraceScene S = scenes[0].S;
    GridRegs G = load_grid(S.gp);
    uint32_t j = 0, bi = 0, started = 0;
    bool live = false;
    bool exhausted = false;
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
                    uint32_t lo = 0, hi = nb;            // prefix[lo] <= j < prefix[hi]
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (prefix[mid] <= j) lo = mid;
                        else hi = mid;
                    }
                    bi = lo;
                    const RayBatchDev& B = batches[bi];
                    const uint32_t jl = j - prefix[bi];
                    S = scenes[B.scene].S;
                    G = load_grid(S.gp);
                    const size_t st = B.stride ? B.stride : B.n;
                    set_ray(r, mk(B.org[jl], B.org[st + jl], B.org[2 * st + jl]),
                            mk(B.dir[jl], B.dir[st + jl], B.dir[2 * st + jl]));
                    r.t = kAny ? B.tmax : kFltMax;
                    r.tterm = B.tmax;
                    ++started;
                    live = ray_begin<true, kAny>(G, S, r, wk);
                    j = jl;
                    if (!live) {
                        if (kAny) {
                            B.occ_out[jl] = r.t < B.tmax ? 1u : 0u;
                            if (B.t_out) B.t_out[jl] = r.t;
                        } else {
                            B.t_out[jl] = r.t;
                            B.prim_out[jl] = scenes[B.scene].nt;
                            if (B.hits) store_hit(B.hits, B.n, jl, r);
                        }
                    }
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (live) {
            bool done;
            if (r.it < r.end) done = test_one<kAny>(S, r, wk);
            else done = !advance_two_level<kAny>(G, S, r, wk);
            if (done) {
                live = false;
                const RayBatchDev& B = batches[bi];
                if (kAny) {
                    B.occ_out[j] = r.t < B.tmax ? 1u : 0u;
                    if (B.t_out) B.t_out[j] = r.t;
                } else {
                    const SceneDesc& D = scenes[B.scene];
                    B.t_out[j] = r.t;
                    B.prim_out[j] = r.t < kFltMax ? r.best - D.tri_base : D.nt;
                    if (B.hits) store_hit(B.hits, B.n, j, r);
                }
            }
        }
    }
    if (ray_count) {
        unsigned long long c = started;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (__lane_id() == 0 && c) atomicAdd(ray_count, c);
    }
    add_work(stats, started, wk);
}

// k_persist_multi with the candidate tests batched across the wave (see
// k_persist_bt): a lane's ray may belong to any scene, so each lane also keeps
// its scene's triangle records and references in LDS for the items it hands
// out.  Up to steps_max grid steps per lane and iteration.
template <bool kAny>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_persist_multi_bt(
    const SceneDesc* __restrict__ scenes, const RayBatchDev* __restrict__ batches, const uint32_t* __restrict__ prefix,
    uint32_t nb, uint32_t nrays, uint32_t* __restrict__ pool, uint32_t refill, uint32_t steps_max,
    unsigned long long* ray_count, unsigned long long* stats)
{
    __shared__ WaveTests<kBatchK> lds[kBlock / kWave];
    __shared__ const float4* recs[kBlock];
    __shared__ const uint32_t* prims[kBlock];
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)threadIdx.x) / kWave;
    WaveTests<kBatchK>& W = lds[wave];
    const float4* const* own_rec = recs + wave * kWave;
    const uint32_t* const* own_prims = prims + wave * kWave;
    const uint32_t lane = __lane_id();
    const uint64_t lt = __lanemask_lt();
    RayState r;
    Work wk;
    TraceScene S = scenes[0].S;
    GridRegs G = load_grid(S.gp);
    uint32_t j = 0, bi = 0, started = 0;
    bool live = false;
    bool exhausted = false;
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
                    uint32_t lo = 0, hi = nb;            // prefix[lo] <= j < prefix[hi]
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (prefix[mid] <= j) lo = mid;
                        else hi = mid;
                    }
                    bi = lo;
                    const RayBatchDev& B = batches[bi];
                    const uint32_t jl = j - prefix[bi];
                    S = scenes[B.scene].S;
                    G = load_grid(S.gp);
                    const size_t st = B.stride ? B.stride : B.n;
                    set_ray(r, mk(B.org[jl], B.org[st + jl], B.org[2 * st + jl]),
                            mk(B.dir[jl], B.dir[st + jl], B.dir[2 * st + jl]));
                    r.t = kAny ? B.tmax : kFltMax;
                    r.tterm = B.tmax;
                    ++started;
                    live = ray_begin<true, kAny>(G, S, r, wk);
                    j = jl;
                    if (live) {
                        W.ox[lane] = r.org.x;
                        W.oy[lane] = r.org.y;
                        W.oz[lane] = r.org.z;
                        W.dx[lane] = r.dir.x;
                        W.dy[lane] = r.dir.y;
                        W.dz[lane] = r.dir.z;
                        recs[threadIdx.x] = S.rec;
                        prims[threadIdx.x] = S.prims;
                    } else if (kAny) {
                        B.occ_out[jl] = r.t < B.tmax ? 1u : 0u;
                        if (B.t_out) B.t_out[jl] = r.t;
                    } else {
                        B.t_out[jl] = r.t;
                        B.prim_out[jl] = scenes[B.scene].nt;
                        if (B.hits) store_hit(B.hits, B.n, jl, r);
                    }
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (exhausted) break;
            continue;
        }
        bool done = false;
        if (live && r.it >= r.end) {
            for (uint32_t m = 0; m < steps_max; ++m) {
                bool top = true;
                if (r.in_leaf) {
                    const int k = leaf_step<kAny>(S, r, wk);
                    if (k != 2) {
                        top = false;
                        done = k != 1;
                    }
                }
                if (top) done = !top_step(G, S, r, mk(W.ox[lane], W.oy[lane], W.oz[lane]), wk);
                if (done || r.it < r.end) break;
            }
        }
        if (wave_test_range<kAny, kBatchK, true>(S.rec, S.prims, W, r.it, r.end, r.t, r.best, lane, live && !done,
                                                 wk, own_rec, own_prims))
            done = true;
        if (live && done) {
            live = false;
            const RayBatchDev& B = batches[bi];
            if (kAny) {
                B.occ_out[j] = r.t < B.tmax ? 1u : 0u;
                if (B.t_out) B.t_out[j] = r.t;
            } else {
                const SceneDesc& D = scenes[B.scene];
                B.t_out[j] = r.t;
                B.prim_out[j] = r.t < kFltMax ? r.best - D.tri_base : D.nt;
                if (B.hits) {
                    r.org = mk(W.ox[lane], W.oy[lane], W.oz[lane]);
                    r.dir = mk(W.dx[lane], W.dy[lane], W.dz[lane]);
                    store_hit(B.hits, B.n, j, r);
                }
            }
        }
    }
    if (ray_count) {
        unsigned long long c = started;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (lane == 0 && c) atomicAdd(ray_count, c);
    }
    add_work(stats, started, wk);
}

// ---------------------------------------------------------------------------
// persistent path tracer
// ---------------------------------------------------------------------------
//
// The megakernel's paths (k_pt, PathTracer.h:41-372) with the traversal of
// k_persist: each lane holds one path and runs, per loop iteration, one unit
// of its current ray (a triangle test or a DDA advance).  A lane whose ray has
// finished waits in the shading state; the wave runs the shading code (path
// vertex, light sample, bounce, the next ray's set-up) once at least
// `shade_min` of its lanes wait or none is tracing, and fetches new pixels
// once `refill` of its lanes are idle.  Per path the operations and their
// order are k_pt's, except that the bounce of an opaque vertex is computed
// before its shadow ray is traced: the bounce reads nothing the shadow ray
// changes, and the light sample is added as k_pt adds it (diffuse, then
// specular part: after the vertex the last bounce is never specular), so
// the radiance is bit-identical.

constexpr uint32_t kPathIdle = 0, kPathShade = 1, kPathTrace = 2;

// the path's state between shading steps; the path's position and direction
// live in the lane's ray (org, dir) while it shades
struct PathLane {
    f3 att, rad;
    f3 pdir;             // extension direction waiting for the shadow ray
    f3 add_d, add_s;     // light sample added when the shadow ray is unoccluded
    uint32_t pix;
    int depth;
    bool last_spec;
    bool after_shadow;   // the finished ray was the vertex's shadow ray
    bool end_after;      // the path ends after that shadow ray
};

__device__ __forceinline__ void pt_store(const RenderArgs& ra, float* fb, uint32_t pix, f3 rad)
{
    fb_put(ra, fb + 3 * (size_t)pix, rad.x, rad.y, rad.z);
}

// a ray of the path: shadow rays stop at the first hit closer than 0.9999,
// extension rays find the closest hit (tterm 0)
template <bool kTwo>
__device__ __forceinline__ bool pt_begin_ray(const GridRegs& G, const TraceScene& S, RayState& r, f3 org, f3 d,
                                             bool shadow, Work& wk)
{
    set_ray(r, org, d);
    r.t = shadow ? 0.9999f : kFltMax;
    r.tterm = shadow ? 0.9999f : 0.f;
    return ray_begin<kTwo, false, true>(G, S, r, wk);
}

// a finished extension ray: the next vertex, or the end of the path
__device__ __forceinline__ bool pt_extension_done(RayState& r, PathLane& L, const TraceScene& S)
{
    if (!(r.t < kFltMax)) return false;
    r.org = r.org + (r.t - kEps) * r.dir;
    ++L.depth;
    return true;
}

template <bool kTwo>
__global__ __launch_bounds__(kBlock) void k_persist_pt(TraceScene S, RenderArgs ra, uint32_t npix,
                                                       const float* __restrict__ hits,
                                                       const uint32_t* __restrict__ hit_prims, Shading sh,
                                                       float* __restrict__ fb, uint32_t* __restrict__ pool,
                                                       uint32_t refill, uint32_t shade_min,
                                                       unsigned long long* ray_count)
{
    if (frame_skipped(ra)) return;
    const GridRegs G = load_grid(S.gp);
    const uint64_t lt = __lanemask_lt();
    RayState r;
    PathLane L;
    Work wk;
    uint32_t state = kPathIdle, traced = 0;
    bool exhausted = false;
    while (true) {
        // new pixels for idle lanes
        const uint64_t idle = __ballot(state == kPathIdle);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && (nidle >= refill || idle == ~0ull)) {
            uint32_t base = 0;
            if (__lane_id() == 0) base = atomicAdd(pool, nidle);
            base = __shfl(base, 0, kWave);
            if (base + nidle >= npix) exhausted = true;
            if (state == kPathIdle) {
                const uint32_t pix = base + (uint32_t)__popcll(idle & lt);
                if (pix < npix) {
                    const float th = hits[6 * (size_t)npix + pix];
                    const uint32_t prim = hit_prims[pix];
                    L.pix = pix;
                    if (th < kFltMax && prim < S.nt) {
                        r.org = mk(hits[pix], hits[(size_t)npix + pix], hits[2 * (size_t)npix + pix]);
                        r.dir = mk(hits[3 * (size_t)npix + pix], hits[4 * (size_t)npix + pix],
                                   hits[5 * (size_t)npix + pix]);
                        r.t = th;
                        r.best = prim;
                        L.att = rep(1.f);
                        L.rad = rep(0.f);
                        L.depth = 0;
                        L.last_spec = true;
                        L.after_shadow = false;
                        state = kPathShade;
                    } else {
                        pt_store(ra, fb, pix, rep(0.f));
                    }
                }
            }
        }
        const uint64_t shading = __ballot(state == kPathShade);
        const uint64_t tracing = __ballot(state == kPathTrace);
        if (!shading && !tracing) {
            if (exhausted) break;
            continue;
        }
        if (shading && ((uint32_t)__popcll(shading) >= shade_min || !tracing) && state == kPathShade) {
            bool done = false;      // the path has ended
            bool shadow = false;    // the next ray is a shadow ray
            f3 ndir;
            if (L.after_shadow) {
                // PathTracer.h:244-258 with the shadow ray's outcome
                L.after_shadow = false;
                if (!(r.t < 0.9999f)) {
                    L.rad = L.rad + L.add_d;
                    L.rad = L.rad + L.add_s;
                }
                done = L.end_after;
                ndir = L.pdir;
            } else {
                PathState P;
                P.org = r.org;
                P.dir = r.dir;
                P.att = L.att;
                P.radiance = L.rad;
                P.ray_t = r.t;
                P.prim = r.best;
                P.depth = L.depth;
                P.last_spec = L.last_spec;
                PathVertex V;
                pt_vertex(S, sh, ra, L.pix, P, V);
                const bool last = P.depth + 1 >= ra.pt_depth;
                if (V.opaque && !last) pt_bounce(P, V);
                const bool end = V.end || last;
                L.att = P.att;
                L.rad = P.radiance;
                L.last_spec = P.last_spec;
                r.org = P.org;
                ndir = P.dir;
                if (V.opaque && V.shadow) {
                    L.add_d = V.add_d;
                    L.add_s = V.add_s;
                    L.pdir = P.dir;
                    L.end_after = end;
                    L.after_shadow = true;
                    shadow = true;
                    ndir = V.sdir;
                } else {
                    done = end;
                }
            }
            if (done) {
                pt_store(ra, fb, L.pix, L.rad);
                state = kPathIdle;
            } else {
                ++traced;
                if (pt_begin_ray<kTwo>(G, S, r, r.org, ndir, shadow, wk)) {
                    state = kPathTrace;
                } else if (!shadow) {
                    // the extension ray misses the grid: the path ends
                    pt_store(ra, fb, L.pix, L.rad);
                    state = kPathIdle;
                }
                // a shadow ray outside the grid is unoccluded: shade again
            }
        }
        if (state == kPathTrace) {
            bool fin;
            if (r.it < r.end) fin = test_one<false, true>(S, r, wk);
            else fin = kTwo ? !advance_two_level<false, true>(G, S, r, wk)
                            : !advance_uniform<false, true>(G, S, r, wk);
            if (fin) {
                if (L.after_shadow) {
                    state = kPathShade;
                } else if (pt_extension_done(r, L, S)) {
                    state = kPathShade;
                } else {
                    pt_store(ra, fb, L.pix, L.rad);
                    state = kPathIdle;
                }
            }
        }
    }
    if (ray_count) {
        unsigned long long c = traced;
#pragma unroll
        for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
        if (__lane_id() == 0 && c) atomicAdd(ray_count, c);
    }
    add_work(ra.stats, traced, wk);
}

template <class K>
uint32_t persistent_blocks(K kernel)
{
    static uint32_t cached = 0;
    if (cached) return cached;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlock, 0);
    if (cus <= 0) cus = 256;
    if (per <= 0) per = 1;
    cached = (uint32_t)(cus * per);
    return cached;
}

template <bool kTwo, bool kAny, class Gen>
void launch(const TraceScene& S, const Gen& gen, uint32_t nrays, uint32_t* pool, uint32_t refill,
            unsigned long long* ray_count, unsigned long long* stats, hipStream_t s, const PersistOpts& o)
{
    if (nrays == 0) return;
    fill_u32(pool, 0u, 1, nullptr, 1, s);
    const uint32_t need = (nrays + kBlock - 1) / kBlock;
    if (o.batch) {
        auto kern = k_persist_bt<kTwo, kAny, Gen, kBatchK>;
        uint32_t blocks = persistent_blocks(kern);
        if (blocks > need) blocks = need;
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, S, gen, nrays, pool, refill, o.steps, ray_count,
                           stats);
        return;
    }
    auto kern = o.fuse ? k_persist<kTwo, kAny, Gen, true> : k_persist<kTwo, kAny, Gen, false>;
    uint32_t blocks = o.fuse ? persistent_blocks(k_persist<kTwo, kAny, Gen, true>)
                             : persistent_blocks(k_persist<kTwo, kAny, Gen, false>);
    if (blocks > need) blocks = need;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, S, gen, nrays, pool, refill, ray_count, stats);
}

}  // namespace

void persist_primary(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, float* hits,
                     uint32_t* hit_prims, uint32_t* pool, uint32_t refill, unsigned long long* ray_count,
                     hipStream_t s, const PersistOpts& o)
{
    const uint32_t npix = (uint32_t)cam.res_x * (uint32_t)cam.res_y;
    GenPrimary g{ cam, ra, npix, hits, hit_prims };
    if (S.two_level) launch<true, false>(S, g, npix, pool, refill, ray_count, ra.stats, s, o);
    else launch<false, false>(S, g, npix, pool, refill, ray_count, ra.stats, s, o);
}

void persist_ao(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                const uint32_t* hit_prims, uint8_t* occ, uint32_t* pool, uint32_t refill,
                unsigned long long* ray_count, hipStream_t s, const PersistOpts& o)
{
    GenAO g{ ra, npix, hits, hit_prims, occ };
    const uint32_t n = npix * (uint32_t)ra.ao_samples;
    if (S.two_level) launch<true, true>(S, g, n, pool, refill, ray_count, ra.stats, s, o);
    else launch<false, true>(S, g, n, pool, refill, ray_count, ra.stats, s, o);
}

void persist_batch(const TraceScene& S, const float* org, const float* dir, uint32_t n, int anyhit, float tmax,
                   float* t_out, uint32_t* prim_out, uint8_t* occ_out, uint32_t* pool, uint32_t refill,
                   unsigned long long* ray_count, unsigned long long* stats, hipStream_t s, float* hits,
                   const PersistOpts& o)
{
    GenBatch g{ org, dir, n, tmax, anyhit, t_out, prim_out, occ_out, anyhit ? nullptr : hits };
    if (S.two_level) {
        if (anyhit) launch<true, true>(S, g, n, pool, refill, ray_count, stats, s, o);
        else launch<true, false>(S, g, n, pool, refill, ray_count, stats, s, o);
    } else {
        if (anyhit) launch<false, true>(S, g, n, pool, refill, ray_count, stats, s, o);
        else launch<false, false>(S, g, n, pool, refill, ray_count, stats, s, o);
    }
}

void persist_pt(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                const uint32_t* hit_prims, const Shading& sh, float* fb, uint32_t* pool, uint32_t refill,
                uint32_t shade_min, unsigned long long* ray_count, hipStream_t s)
{
    if (npix == 0) return;
    fill_u32(pool, 0u, 1, nullptr, 1, s);
    const uint32_t need = (npix + kBlock - 1) / kBlock;
    if (S.two_level) {
        uint32_t blocks = persistent_blocks(k_persist_pt<true>);
        if (blocks > need) blocks = need;
        hipLaunchKernelGGL(k_persist_pt<true>, dim3(blocks), dim3(kBlock), 0, s, S, ra, npix, hits, hit_prims, sh,
                           fb, pool, refill, shade_min, ray_count);
    } else {
        uint32_t blocks = persistent_blocks(k_persist_pt<false>);
        if (blocks > need) blocks = need;
        hipLaunchKernelGGL(k_persist_pt<false>, dim3(blocks), dim3(kBlock), 0, s, S, ra, npix, hits, hit_prims, sh,
                           fb, pool, refill, shade_min, ray_count);
    }
}

void persist_multi(const SceneDesc* scenes, const RayBatchDev* batches, const uint32_t* prefix, uint32_t nb,
                   uint32_t nrays, int anyhit, uint32_t* pool, uint32_t refill, uint32_t blocks_per_cu,
                   unsigned long long* ray_count, unsigned long long* stats, hipStream_t s, const PersistOpts& o)
{
    if (nrays == 0 || nb == 0) return;
    fill_u32(pool, 0u, 1, nullptr, 1, s);
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
    uint32_t blocks = (uint32_t)cus * (blocks_per_cu ? blocks_per_cu : 1u);
    const uint32_t need = (nrays + kBlock - 1) / kBlock;
    if (o.batch) {
        if (blocks_per_cu == 0)
            blocks = anyhit ? persistent_blocks(k_persist_multi_bt<true>) : persistent_blocks(k_persist_multi_bt<false>);
        if (blocks > need) blocks = need;
        if (anyhit)
            hipLaunchKernelGGL(k_persist_multi_bt<true>, dim3(blocks), dim3(kBlock), 0, s, scenes, batches, prefix, nb,
                               nrays, pool, refill, o.steps, ray_count, stats);
        else
            hipLaunchKernelGGL(k_persist_multi_bt<false>, dim3(blocks), dim3(kBlock), 0, s, scenes, batches, prefix,
                               nb, nrays, pool, refill, o.steps, ray_count, stats);
        return;
    }
    if (blocks_per_cu == 0) {
        blocks = anyhit ? persistent_blocks(k_persist_multi<true>) : persistent_blocks(k_persist_multi<false>);
    }
    if (blocks > need) blocks = need;
    if (anyhit)
        hipLaunchKernelGGL(k_persist_multi<true>, dim3(blocks), dim3(kBlock), 0, s, scenes, batches, prefix, nb, nrays,
                           pool, refill, ray_count, stats);
    else
        hipLaunchKernelGGL(k_persist_multi<false>, dim3(blocks), dim3(kBlock), 0, s, scenes, batches, prefix, nb,
                           nrays, pool, refill, ray_count, stats);
}

}  // namespace hr
