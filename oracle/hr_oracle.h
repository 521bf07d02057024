/*
 * hr_oracle.h -- scalar CPU restatement of happyray's per-frame hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and by nothing else.  The product (happyray_amd/,
 * include/happyray.h) never calls it and never falls back to it.
 *
 * It follows the *host* branches of the reference (IEEE division and
 * square root, no fused multiply-add; build with -ffp-contract=off), with
 * the documented departures:
 *   - canonical (cell, primitive id) order inside every cell and leaf
 *     (the reference's order comes from shared-memory atomics);
 *   - a Philox4x32-10 counter RNG keyed by (pixel, image, stream) instead
 *     of KISS / XorShift seeded by the CUDA thread id;
 *   - transcendental functions (pow, log, exp, sin/cos) evaluated by the
 *     double-precision routines in or_math.c (a written spec that the GPU
 *     kernels implement too) and rounded to float;
 *   - AO occlusion decided against a configurable distance instead of the
 *     bounding-box-derived alpha, any-hit termination at that distance;
 *   - a fixed maximum path depth for the path tracer;
 *   - exact-insertion pairs that fail the triangle/box test are dropped
 *     instead of being written with key = #cells and sorted last.
 * Parity status of every function is listed in the header of the file that
 * implements it.
 */
#ifndef HR_ORACLE_H
#define HR_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* grid kinds (src/RT/RTEngine.cu:64-89) */
#define OR_GRID_UNIFORM        0  /* uniform grid, bounding-box insertion   */
#define OR_GRID_UNIFORM_EXACT  1  /* uniform grid, exact triangle insertion */
#define OR_GRID_TWO_LEVEL      2  /* two-level grid                          */
#define OR_GRID_HIERARCHY      3  /* instanced two-level grid hierarchy      */

typedef struct {
    const float* vx;   /* SoA vertex coordinates, nv each */
    const float* vy;
    const float* vz;
    uint32_t nv;
    const uint32_t* idx; /* 3 * nt vertex indices */
    uint32_t nt;
    /* optional per-corner normals (NULL nidx: geometric normals) */
    const float* nx;     /* SoA normal coordinates, nn each */
    const float* ny;
    const float* nz;
    uint32_t nn;
    const uint32_t* nidx; /* 3 * nt normal indices */
} or_mesh;

typedef struct {
    int kind;
    float bmin[3], bmax[3];
    int res[3];
    float cell_size[3], cell_size_rcp[3];
    uint32_t num_cells;      /* top-level cells                                  */
    uint32_t num_leaves;     /* two-level only                                   */
    uint32_t num_refs;       /* length of prims[]                                */
    uint32_t num_top_pairs;  /* two-level only: (top cell, prim) pairs           */
    uint32_t num_bbox_pairs; /* exact uniform only: pairs before the exact test  */
    uint32_t* cells;         /* 2*num_cells: UG (begin,end); TLG (leafBegin, packed) */
    uint32_t* leaves;        /* 2*num_leaves (begin,end)                          */
    uint32_t* prims;         /* num_refs primitive ids                            */
} or_grid;

typedef struct {
    uint64_t rays;       /* rays traversed */
    uint64_t steps;      /* DDA cell steps (top + leaf) */
    uint64_t tests;      /* ray/triangle tests */
} or_stats;

typedef struct {
    float center[3], top_left[3], step_x[3], step_y[3];
    int res_x, res_y;
} or_camera;

/* PhongMaterial layout (src/RT/Primitive/Material.hpp:37-84): 12 floats */
typedef struct {
    float kd[3];  float ior_w;      /* diffuse, IOR (+10 when refractive) */
    float ks[3];  float spec_exp;   /* specular, exponent                 */
    float ke[3];  float pad;        /* emission                           */
} or_material;

/* AreaLightSource (src/RT/Primitive/LightSource.hpp:38-99) */
typedef struct {
    float position[3], normal[3], intensity[3], edge1[3], edge2[3];
} or_light;

typedef struct {
    uint32_t seed_lo, seed_hi;  /* Philox key */
    int jitter;                 /* 0: pixel centre, 1: Gaussian pixel filter */
    int ao_samples;             /* AO rays per primary hit */
    float ao_dist;              /* AO occlusion distance   */
    int pt_depth;               /* max path vertices        */
    int nthreads;               /* host threads (tiles of 64 rays) */
    int simple_sampler;         /* 1: renderFrame mode 0's pixel sampler (RTEngine.cu:414-423):
                                   RegularPixelSampler<2,2> for images 0-3, Gaussian after */
} or_render_params;

/* ---- spec math (or_math.c) ---- */
float or_powf(float x, float y);
float or_logf(float x);
float or_expf(float x);
void  or_sincosf(float a, float* s, float* c);
void  or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float or_u01(uint32_t x);

/* ---- key frames and bounds (or_build.c) ---- */
void or_lerp_bounds(const float* x1, const float* y1, const float* z1,
                    const float* x2, const float* y2, const float* z2,
                    float coeff, uint32_t nv,
                    float* ox, float* oy, float* oz,
                    float bmin[3], float bmax[3]);
void or_frame_ids(float current, float step, uint32_t num_frames,
                  uint32_t* k1, uint32_t* k2, float* coeff);

/* ---- grid builds (or_build.c) ---- */
int  or_grid_resolution(const float bmin[3], const float bmax[3], uint32_t nprims,
                        float lambda, int res[3]);
int  or_build(const or_mesh* mesh, const float bmin[3], const float bmax[3],
              int kind, float lambda_top, float lambda_leaf, or_grid* out);
void or_grid_free(or_grid* g);
/* per-primitive cell range and exact-test helpers, exposed for tests */
void or_prim_cell_range(const or_mesh* m, uint32_t prim, const or_grid* g,
                        int lo[3], int hi[3]);
int  or_tri_box_overlap(const float v0[3], const float v1[3], const float v2[3],
                        const float center[3], const float half[3]);

/* renderFrame mode 0: primary rays + simpleShade (SimpleIntegrator.h:42-125);
 * fb (3 floats per pixel) is the running mean, weight 1/(image_id+1) */
int  or_render_simple(const or_mesh* mesh, const or_grid* g, const or_camera* cam,
                      const or_render_params* p, int image_id, float* fb, or_stats* stats);
/* renderFrame mode 0 with input rays (SimpleRayTraverser, RTEngine.cu:410-413):
 * rays are n 24-byte records (origin xyz, direction xyz); rgb gets the simple
 * shading of each ray's closest hit at image 0 */
int  or_render_rays_simple(const or_mesh* mesh, const or_grid* g, const float* rays, uint32_t n,
                           int nthreads, float* rgb, or_stats* stats);
/* AORayExporter's dump (AORayExporter.h:50-68): the npix camera rays, then
 * ao_samples AO rays per pixel, as (npix * (ao_samples + 1)) 24-byte records */
int  or_dump_ao_rays(const or_mesh* mesh, const or_grid* g, const or_camera* cam,
                     const or_render_params* p, int image_id, float* out);
/* renderFrame mode 1: the multi-kernel path tracer (PathTracer.h:374-1215).
 * state holds 6 floats per pixel carried from image to image (new radiance,
 * residue; zero them with the camera); end / end_prims (may be NULL) receive
 * the end vertices (7 SoA planes) and importance (may be NULL) the path
 * throughput (3 floats per pixel). */
int  or_render_pt_multi(const or_mesh* mesh, const or_grid* g, const or_camera* cam,
                        const or_material* mats, const uint32_t* mat_idx,
                        const or_light* lights, const float* light_weights, int nlights,
                        const or_render_params* p, int image_id, float* fb, float* state,
                        float* end, uint32_t* end_prims, float* importance, or_stats* stats);

/* the camera ray directions of all pixels (before the traversal's reciprocal) */
void or_primary_dirs(const or_camera* cam, const or_render_params* p, int image_id,
                     float* dx, float* dy, float* dz);
/* paths from given rays: closest hit of each ray stored as a primary hit, then
 * the path tracer of or_render_pt from it (image 0, ray i = pixel i) */
int  or_path_trace_rays(const or_mesh* mesh, const or_grid* g,
                        const float* ox, const float* oy, const float* oz,
                        const float* dx, const float* dy, const float* dz, uint32_t n,
                        const or_material* mats, const uint32_t* mat_idx,
                        const or_light* lights, const float* light_weights, int nlights,
                        const or_render_params* p, float* rgb, or_stats* stats);

/* path-tracer shading normal of primitive prim seen from hit point org
 * (PathTracer.h:109-135), exposed for tests */
void or_shading_normal(const or_mesh* mesh, uint32_t prim, const float org[3], float out[3]);

/* The path tracers' sampling steps, exposed for tests (the renderers call the
 * same functions).  or_pt_glass: one vertex on a transparent material of index
 * ior (PathTracer.h:133-183; mode1 != 0: :556-622), updating org, dir and the
 * throughput att; returns 0 reflected, 1 refracted, 2 ended.
 * or_power_cosine_dir: PowerCosineHemisphereSampler about the unit axis upv
 * (HemisphereSamplers.hpp:78-94).  or_cosine_dir: CosineHemisphereSampler
 * about the unit normal n (:58-71) in getLocalCoordinates' frame.
 * or_pt_glossy: the megakernel's glossy bounce (PathTracer.h:284-305) for
 * cos_ne = -n.dir >= 0; returns 0 when the path ends.  or_pm_merge: mode 1's
 * addIndirectIllumination (PathTracer.h:889-990) on 3-float-per-pixel
 * buffers; updates residue and fb. */
int  or_pt_glass(float org[3], float dir[3], float att[3], const float n[3], float cos_ne, float ior,
                 float ray_t, const float kd[3], float w_fres, float w_tir, int mode1);
void or_power_cosine_dir(float w0, float w1, float se, const float upv[3], float out[3]);
void or_cosine_dir(float w0, float w1, const float n[3], float out[3]);
int  or_pt_glossy(float dir[3], float att[3], const float n[3], float cos_ne, const float ks[3], float se,
                  float w0, float w1);
void or_pm_merge(const float* imp, const float* nrad, float* residue, float* fb, int image_id, int rx, int ry);

/* ---- traversal (or_trace.c) ---- */
int  or_trace(const or_mesh* mesh, const or_grid* g,
              const float* ox, const float* oy, const float* oz,
              const float* dx, const float* dy, const float* dz,
              uint32_t n, int anyhit, float tmax, int nthreads,
              float* t_out, uint32_t* prim_out, or_stats* stats);
void or_camera_init(or_camera* cam, const float pos[3], const float dir[3],
                    const float up[3], float fov_deg, int res_x, int res_y);
/* primary hit buffer: hit point, stored direction, t, primitive id */
int  or_primary(const or_mesh* mesh, const or_grid* g, const or_camera* cam,
                const or_render_params* p, int image_id,
                float* hx, float* hy, float* hz, float* hdx, float* hdy, float* hdz,
                float* ht, uint32_t* hprim, or_stats* stats);
int  or_render_ao(const or_mesh* mesh, const or_grid* g, const or_camera* cam,
                  const or_material* mats, const uint32_t* mat_idx,
                  const or_render_params* p, int image_id,
                  float* fb_rgb, uint8_t* occ_out, or_stats* stats);
int  or_render_pt(const or_mesh* mesh, const or_grid* g, const or_camera* cam,
                  const or_material* mats, const uint32_t* mat_idx,
                  const or_light* lights, const float* light_weights, int nlights,
                  const or_render_params* p, int image_id,
                  float* fb_rgb, or_stats* stats);

/* ---- instanced hierarchy (or_hier.c) ---- */
/* one instance: world-to-object rows m_row_0..2 of the instances file and the
 * instance's world box (WFObjectReader.cpp:762-824) */
typedef struct {
    uint32_t object;
    float m[3][4];
    float aabb_min[3], aabb_max[3];
} or_instance;

typedef struct {
    float bmin[3], bmax[3];
    int res[3];
    float cell_size[3], cell_size_rcp[3];
    uint32_t first_cell, num_cells;
} or_object_grid;

typedef struct {
    or_grid top;             /* cells: (begin, end) into inst_refs; num_leaves = object
                                cells, num_refs = object references, num_top_pairs =
                                instance references */
    uint32_t nobj, ninst;
    or_object_grid* objs;
    uint32_t num_obj_cells;
    uint32_t* obj_cells;     /* 2 * num_obj_cells (begin, end) into prims */
    uint32_t num_obj_refs;
    uint32_t* prims;         /* triangle ids */
    uint32_t num_inst_refs;
    uint32_t* inst_refs;     /* instance ids */
    const or_instance* inst; /* the caller's instances (kept, not copied) */
} or_hier;

/* m: current-frame mesh; kx, ky, kz: the first key frame's vertices (object
 * bounds); bmin, bmax: the padded scene bounds of or_lerp_bounds; object k
 * owns triangles [obj_first[k], obj_first[k+1]).  Returns 0, or -2 when a
 * resolution or count exceeds 32 bits, -1 on allocation failure. */
int  or_build_hier(const or_mesh* m, const float* kx, const float* ky, const float* kz,
                   const float bmin[3], const float bmax[3], const uint32_t* obj_first, uint32_t nobj,
                   const or_instance* inst, uint32_t ninst, float lambda_top, float lambda_leaf,
                   or_hier* out);
void or_hier_free(or_hier* h);
/* closest hits (t, prim, instance; prim = nt and instance 0xFFFFFFFF on a
 * miss) or occlusion (distance below tmax after the 0.9999 shadow stop) */
int  or_trace_hier(const or_mesh* m, const or_hier* h, const float* ox, const float* oy, const float* oz,
                   const float* dx, const float* dy, const float* dz, uint32_t n, int anyhit, float tmax,
                   float* t_out, uint32_t* prim_out, uint32_t* inst_out, uint8_t* occ_out);
/* TLGHAOIntegrator (TLGridHierarchyAOIntegrator.h:421-573) at image image_id:
 * fb is the running mean; occ_out (ao_samples per pixel), hits7 (7 SoA
 * planes), hprim_out and hinst_out may be NULL */
int  or_render_ao_hier(const or_mesh* m, const or_hier* h, const or_camera* cam, const or_material* mats,
                       const uint32_t* mat_idx, const or_render_params* p, int image_id, float* fb,
                       uint8_t* occ_out, float* hits7, uint32_t* hprim_out, uint32_t* hinst_out);

#ifdef __cplusplus
}
#endif
#endif
