// hr_persist.h -- host interface of the persistent traversal kernels.
#pragma once

#include "hr_trace.h"

namespace hr {

// `pool` is one device word (reset by the launch); `refill` is the number of
// idle lanes of a wave that triggers fetching new rays (1..64).
void persist_primary(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, float* hits,
                     uint32_t* hit_prims, uint32_t* pool, uint32_t refill, unsigned long long* ray_count,
                     hipStream_t s);
void persist_ao(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                const uint32_t* hit_prims, uint8_t* occ, uint32_t* pool, uint32_t refill,
                unsigned long long* ray_count, hipStream_t s);
void persist_batch(const TraceScene& S, const float* org, const float* dir, uint32_t n, int anyhit, float tmax,
                   float* t_out, uint32_t* prim_out, uint8_t* occ_out, uint32_t* pool, uint32_t refill,
                   unsigned long long* ray_count, unsigned long long* stats, hipStream_t s);

}  // namespace hr
