// hr_trace.h -- host interface of the traversal and integrator kernels.
#pragma once

#include "hr_device.h"

namespace hr {

// Everything a traversal reads, by device pointer.
struct TraceScene {
    const GridParams* gp;
    const uint2* cells;
    const uint2* leaves;       // two-level only
    const uint32_t* prims;
    const float4* rec;         // triangle records, 3 float4 each
    uint32_t nt;
    int two_level;
};

// Camera::init result (src/RT/Primitive/Camera.h:35-72)
struct CameraDev {
    float center[3], top_left[3], step_x[3], step_y[3];
    int res_x, res_y;
};

struct RenderArgs {
    uint32_t seed_lo, seed_hi;
    int jitter;
    int simple_sampler;          // 1: RTEngine::renderFrame mode 0's choice (RTEngine.cu:414-423):
                                 // RegularPixelSampler<2,2> with sample id = image for images
                                 // 0-3, the Gaussian pixel filter after; 0: `jitter` decides
    int ao_samples;
    float ao_dist;
    int pt_depth;
    int image;
    unsigned long long* stats;   // nullptr, or {rays, steps, tests} counters
    const FrameSel* frame;       // nullptr, or the frame's selection in device memory
                                 // (its image id then replaces `image`)
    int direct;                  // 1: each pixel's value is stored as it is (a frame lane's
                                 // own buffer, which k_frame_blend then blends into the
                                 // framebuffer); 0: blended in with weight 1/(image+1)
};

__device__ __forceinline__ int image_of(const RenderArgs& ra) { return ra.frame ? ra.frame->image : ra.image; }
// a captured frame whose rebuild overflowed (or that follows one that did
// before the host caught it) is skipped: each render kernel returns at entry,
// the same for the whole grid, so the frame leaves the framebuffer, the hit and
// path buffers and the ray and work counters as they were
__device__ __forceinline__ bool frame_skipped(const RenderArgs& ra) { return ra.frame && ra.frame->skip; }

// a pixel's value into the framebuffer: the running mean over images
// (AOIntegrator.h:181-187), or stored as it is for a frame lane
__device__ __forceinline__ void fb_put(const RenderArgs& ra, float* o, float x, float y, float z)
{
    if (ra.direct) {
        o[0] = x;
        o[1] = y;
        o[2] = z;
        return;
    }
    const float nw = 1.f / (float)(image_of(ra) + 1);
    const float ow = 1.f - nw;
    o[0] = o[0] * ow + x * nw;
    o[1] = o[1] * ow + y * nw;
    o[2] = o[2] * ow + z * nw;
}

struct Shading {
    const float* mats;          // 12 floats per PhongMaterial
    const uint32_t* mat_idx;    // per triangle, nullptr -> material 0
    const float* lights;        // 15 floats per AreaLightSource
    const float* light_cdf;     // inclusive sum of light weights
    int nlights;
    const float4* normals;      // current-frame normals (xyz, 0), nullptr -> geometric
    const uint32_t* nidx;       // 3 normal indices per triangle
};

// hit buffer: 7 SoA planes of npix floats (hit point xyz, dir xyz, t) + prims
// wave_steps > 0: k_primary_wave (batched candidate tests); the same hits.
// Otherwise k_primary: blocks_per_cu > 0 (with next, a device word) runs that
// many workgroups per CU taking 8x8 pixel tiles from a counter, 0 one
// workgroup per 256 pixels; the same hits either way
void launch_primary(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, float* hits,
                    uint32_t* hit_prims, unsigned long long* ray_count, hipStream_t s,
                    uint32_t wave_steps = 0, uint32_t blocks_per_cu = 0, uint32_t* next = nullptr);
void launch_ao_rays(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                    const uint32_t* hit_prims, uint8_t* occ, unsigned long long* ray_count, hipStream_t s);
void launch_ao_shade(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                     const uint32_t* hit_prims, const uint8_t* occ, const Shading& sh, float* fb,
                     hipStream_t s);
// wave_steps > 0: the wave-cooperative kernel (batched candidate tests, up to
// wave_steps grid steps per lane and batch); 0: one path per lane throughout.
// The same image either way.
void launch_pt(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
               const uint32_t* hit_prims, const Shading& sh, float* fb,
               unsigned long long* ray_count, hipStream_t s, uint32_t wave_steps = 0);
// current-frame normals: a * (1 - 0) + b * 0 per component, as the reference
// uploads normals with coefficient 0 (src/RT/RTEngine.cu:209-210); a, b are
// SoA key-frame blocks of nn normals
// zero n floats of fb when fs->image == 0 (read on the device)
void launch_clear_first(float* fb, size_t n, const FrameSel* fs, hipStream_t s);
// after a captured frame's rebuild: halt = min(halt, fs->seq) when the rebuild
// overflowed, and fs->skip = (halt <= fs->seq).  halt holds the least sequence
// number of an overflowed frame until the host resets it to 0xFFFFFFFF, so
// every frame queued behind an overflowed one is skipped too.  frame_rays (a
// frame lane's ray counter, or nullptr) is cleared for the frame.
void launch_frame_gate(const BuildStatus* st, uint32_t* halt, FrameSel* fs, unsigned long long* frame_rays,
                       hipStream_t s);
// a frame lane's values into the framebuffer, in frame order: unless the frame
// or an earlier one overflowed (halt <= fs->seq), fb = fb * (1 - w) + rad * w
// with w = 1 / (image + 1) (fb taken as 0 for image 0, as after its clear), and
// the frame's ray count is added to the context's
void launch_frame_blend(const FrameSel* fs, const uint32_t* halt, const float* rad, float* fb, uint32_t npix,
                        const unsigned long long* frame_rays, unsigned long long* ray_count, hipStream_t s);
// simpleShade (src/RT/Integrator/SimpleIntegrator.h:42-125): barycentric colour
// times |cos| at each primary hit, background (0.15, 0.2, 0.3), running mean
void launch_simple_shade(const TraceScene& S, const RenderArgs& ra, uint32_t npix, const float* hits,
                         const uint32_t* hit_prims, float* fb, hipStream_t s);

// 24-byte ray records (origin xyz, direction xyz) -> SoA planes `stride` floats apart
void launch_wire_to_soa(const float* rec, uint32_t n, float* org, float* dir, uint32_t stride, hipStream_t s);
// AORayExporter's ray dump (AORayExporter.h:50-68, :262-300) of the render whose
// primary hits are in hits / hit_prims: npix camera rays, then ao_samples rays
// per pixel as AmbientOcclusionRayGenerator makes them, as 24-byte records
void launch_ao_dump(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, const float* hits,
                    const uint32_t* hit_prims, float* out, hipStream_t s);

// Buffers of the multi-kernel path tracer (PathTracer.h:1038-1215), npix pixels.
// rgb buffers hold 3 floats per pixel; new_radiance and residue carry over from
// one image to the next and are cleared with the camera.
struct PathBuffers {
    float* end;                 // 7 SoA planes: end-of-path hit point, direction, t
    uint32_t* end_prims;
    float* importance;          // path throughput (ImportanceBuffer)
    float* ls;                  // light radiance of the shadow sample
    float* dir_ls;              // unit direction to the light, 0 if occluded or none
    float* new_radiance;        // running mean of throughput x direct light
    float* residue;             // what the blur held back (addIndirectIllumination)
    float* blended;             // fb * old_w + contribution * new_w (blur input)
    float* contrib;             // importance * new_radiance + residue
};
constexpr int kPathMaxBounces = 8;   // RFRACTION_PATH_MAX_DEPTH (PathTracer.h:23)
// tracePath, shadow rays, computeDirectIllumination and addIndirectIllumination;
// hits / hit_prims hold the primary hits
void launch_path_multi(const TraceScene& S, const CameraDev& cam, const RenderArgs& ra, const float* hits,
                       const uint32_t* hit_prims, const Shading& sh, const PathBuffers& pb, float* fb,
                       unsigned long long* ray_count, hipStream_t s, hipEvent_t after_trace = nullptr);
// fs != nullptr: a and b are blocks fs->k1 and fs->k2 of keys, chosen on the device
void launch_frame_normals(const float* a, const float* b, uint32_t nn, float4* out, hipStream_t s,
                          const FrameSel* fs = nullptr, const float* keys = nullptr);
void launch_trace_batch(const TraceScene& S, const float* org, const float* dir, uint32_t n,
                        int anyhit, float tmax, float* t_out, uint32_t* prim_out, uint8_t* occ_out,
                        unsigned long long* ray_count, unsigned long long* stats, hipStream_t s);

}  // namespace hr
