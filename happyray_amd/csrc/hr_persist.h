// hr_persist.h -- host interface of the persistent traversal kernels.
#pragma once

#include "hr_trace.h"

namespace hr {

// Kernel choice of a persistent launch; results are the same for every choice.
//   batch 0: each lane tests its own candidates, one per loop iteration, and
//     fuse 1 lets a lane that steps to a new cell or leaf test its first
//     candidate in the same iteration (fuse 0: one unit of work per lane and
//     iteration);
//   batch 1: each lane hands up to 12 candidates per iteration to its wave,
//     which tests them all at full width, and takes up to `steps` grid steps
//     (1..4) per iteration to reach its next candidates.
struct PersistOpts {
    uint32_t fuse = 1;
    uint32_t batch = 1;
    uint32_t steps = 2;
};

// `pool` is one device word (reset by the launch); `refill` is the number of
// idle lanes of a wave that triggers fetching new rays (1..64).
void persist_primary(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, float* hits,
                     uint32_t* hit_prims, uint32_t* pool, uint32_t refill, unsigned long long* ray_count,
                     hipStream_t s, const PersistOpts& o = PersistOpts{});
void persist_ao(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                const uint32_t* hit_prims, uint8_t* occ, uint32_t* pool, uint32_t refill,
                unsigned long long* ray_count, hipStream_t s, const PersistOpts& o = PersistOpts{});
// hits != nullptr (closest hit only): also the 7-plane hit buffer of the rays
void persist_batch(const TraceScene& S, const float* org, const float* dir, uint32_t n, int anyhit, float tmax,
                   float* t_out, uint32_t* prim_out, uint8_t* occ_out, uint32_t* pool, uint32_t refill,
                   unsigned long long* ray_count, unsigned long long* stats, hipStream_t s,
                   float* hits = nullptr, const PersistOpts& o = PersistOpts{});

// The path-tracing megakernel's paths (one per pixel of the hit buffer) in a
// persistent launch: `refill` idle lanes fetch new pixels, `shade_min` lanes
// waiting to shade trigger the wave's shading step (1..64).  Results equal
// launch_pt's.
void persist_pt(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                const uint32_t* hit_prims, const Shading& sh, float* fb, uint32_t* pool, uint32_t refill,
                uint32_t shade_min, unsigned long long* ray_count, hipStream_t s);

// A scene of a batch as the traversal sees it: its grid (cells from the
// scene's first cell; leaves, references and triangle records shared by the
// batch, with global indices) and the offset of its triangle ids.
struct SceneDesc {
    TraceScene S;
    uint32_t tri_base;   // global id of the scene's triangle 0
    uint32_t nt;         // the scene's triangle count (miss id of its rays)
};

// One ray batch of a multi-scene launch (device pointers, SoA rays).
struct RayBatchDev {
    uint32_t scene;
    uint32_t n;
    float tmax;
    uint32_t stride;     // elements between the x, y and z planes of org / dir (0: n)
    const float* org;
    const float* dir;
    float* t_out;        // closest hit: distance (FLT_MAX on a miss); any hit: optional
    uint32_t* prim_out;  // closest hit: scene-local triangle id, the scene's nt on a miss
    uint8_t* occ_out;    // any hit: 1 when occluded
    float* hits;         // closest hit, optional: 7 SoA planes of n floats (hit point
                         // org + (t - EPS) dir, direction, t), the layout of the
                         // primary-hit buffer the path tracer starts from
};

// One persistent launch over the rays of nb batches of many scenes (all
// closest-hit or all any-hit): rays [prefix[b], prefix[b+1]) belong to batch
// b; lanes refill from a global counter and may hold rays of different scenes.
void persist_multi(const SceneDesc* scenes, const RayBatchDev* batches, const uint32_t* prefix, uint32_t nb,
                   uint32_t nrays, int anyhit, uint32_t* pool, uint32_t refill, uint32_t blocks_per_cu,
                   unsigned long long* ray_count, unsigned long long* stats, hipStream_t s,
                   const PersistOpts& o = PersistOpts{});

}  // namespace hr
