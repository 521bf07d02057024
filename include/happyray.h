/*
 * happyray.h -- C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * One opaque context replaces the static, single-instance RTEngine facade
 * (src/RT/RTEngine.h:36-68).  Every entry point returns an hr_status; the
 * message of the last failure on the calling thread is hr_last_error().
 * The reference prints and exit(1)s on CUDA errors
 * (src/Utils/CUDAUtil.h:199-206); here nothing exits the process.
 *
 * Threading: calls on one context must not overlap; different contexts are
 * independent (each owns its HIP stream and device memory).
 * Ownership: the caller owns every host buffer it passes; the context owns
 * all device memory, which grows and is reused across frames like the
 * reference's memory managers (src/RT/Structure/MemoryManager.h:73-139).
 *
 * Host buffers are plain C arrays: vertices are SoA (x[nv], y[nv], z[nv]),
 * indices are uint32 triplets, rays are SoA.  Entry points whose name ends
 * in _dev take device pointers instead (for callers that keep data in HBM).
 */
#ifndef HAPPYRAY_H
#define HAPPYRAY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HR_OK = 0,
    HR_ERR_ARG = 1,        /* invalid argument                              */
    HR_ERR_DEVICE = 2,     /* HIP runtime error / no device                 */
    HR_ERR_OOM = 3,        /* device allocation failed                      */
    HR_ERR_IO = 4,         /* file could not be read or written             */
    HR_ERR_STATE = 5,      /* call out of order (e.g. render before build)  */
    HR_ERR_CAPACITY = 6,   /* a fixed capacity was exceeded                 */
    HR_ERR_PARSE = 7,      /* malformed scene / camera / light file         */
    HR_ERR_BUDGET = 8      /* request refused by the memory budget          */
} hr_status;

typedef enum {
    HR_GRID_UNIFORM = 0,       /* uniform grid, bounding-box insertion          */
    HR_GRID_UNIFORM_EXACT = 1, /* uniform grid, exact triangle/box insertion    */
    HR_GRID_TWO_LEVEL = 2,     /* two-level grid                                */
    HR_GRID_HIERARCHY = 3      /* instanced two-level grid hierarchy: a uniform grid per
                                  object under a uniform grid of instance boxes
                                  (the reference's TLGRIDHIERARCHY build,
                                  src/RT/RTEngine.cu:260-349); renders AO only    */
} hr_grid_kind;

/* RTEngine::renderFrame modes (src/RT/RTEngine.cu:398-444) */
typedef enum {
    HR_MODE_SIMPLE = 0,        /* mode 0: primary rays + simpleShade; the 2x2 regular
                                  pixel sampler for images 0-3, the Gaussian filter after */
    HR_MODE_PATH_MULTI = 1,    /* mode 1: multi-kernel path tracer with the blur / residue
                                  merge (PathTracer.h:374-1215); up to 8 bounces; jitter
                                  and pt_depth are not used                             */
    HR_MODE_PATH = 2,          /* mode 2: path-tracing megakernel                       */
    HR_MODE_AO = 3             /* mode 3: ambient occlusion                             */
} hr_mode;

typedef struct hr_ctx hr_ctx;

typedef struct {
    int device;                /* HIP device ordinal                            */
    uint32_t seed_lo, seed_hi; /* key of the counter-based RNG                  */
} hr_options;

/* Grid description returned after a build (UniformGrid / TwoLevelGrid
 * parameters, src/RT/Structure/UniformGrid.h:37-138,
 * src/RT/Structure/TwoLevelGrid.h:163-230). */
typedef struct {
    int kind;
    float bmin[3], bmax[3];
    int res[3];
    float cell_size[3], cell_size_rcp[3];
    uint32_t num_cells;      /* (top-level) cells                         */
    uint32_t num_leaves;     /* two-level: leaf cells; hierarchy: object cells */
    uint32_t num_refs;       /* primitive references (length of prims)   */
    uint32_t num_top_pairs;  /* two-level: top-level (cell, prim) pairs;
                                hierarchy: (top cell, instance) pairs     */
    uint32_t num_triangles;
    uint32_t num_objects;    /* hierarchy: objects, else 0                */
    uint32_t num_instances;  /* hierarchy: instances, else 0              */
} hr_grid_info;

/* AreaLightSource (src/RT/Primitive/LightSource.hpp:38-99) */
typedef struct {
    float position[3], normal[3], intensity[3], edge1[3], edge2[3];
} hr_area_light;

typedef struct {
    int jitter;          /* 0: pixel centres; 1: Gaussian pixel filter (Camera.h:116-132) */
    int ao_samples;      /* AO rays per primary hit (AOIntegrator.h:45 uses 4)            */
    float ao_dist;       /* occlusion distance (the reference uses 0.05*|bbox diagonal|)  */
    int pt_depth;        /* path vertices                                                 */
    int collect_stats;   /* 1: count DDA steps and triangle tests (hr_get_work_stats)    */
} hr_render_params;

/* Device times from hipEvents recorded on the context's stream around each
 * phase of the last build / render. */
typedef struct {
    double build_ms;         /* last hr_build                                    */
    double render_ms;        /* last hr_render, all kernels                      */
    double primary_ms;       /* primary-ray kernel of the last render            */
    double secondary_ms;     /* AO-ray kernel or path-tracing kernel             */
    double shade_ms;         /* AO shading kernel                                */
    uint64_t rays_traced;    /* rays traversed since the context was created     */
    int primary_kernel;      /* camera-ray kernel of the last render: 0 one ray per
                                lane, 1 persistent, 3 one ray per lane with
                                wave-batched tests (-1 before the first render)   */
    uint64_t frames_replayed;  /* captured frames replayed and completed since the
                                  context was created, and the sums of their times
                                  (rebuild graph, render, and the render's camera-
                                  ray, secondary and shading parts); a replay whose
                                  events cannot be read is not counted             */
    double replay_build_ms, replay_render_ms, replay_primary_ms, replay_secondary_ms, replay_shade_ms;
} hr_timing;

/* Work counters of the renders run with collect_stats = 1 and of all ray
 * batches, split by kernel: primary rays (and ray batches), then secondary
 * rays (AO rays, or the path tracer's shadow and extension rays). */
typedef struct {
    uint64_t rays;           /* rays traversed              */
    uint64_t steps;          /* DDA cell steps (top + leaf) */
    uint64_t tests;          /* ray/triangle tests          */
} hr_work_counts;

typedef struct {
    hr_work_counts primary;
    hr_work_counts secondary;
} hr_work_stats;

/* ---- lifetime: RTEngine::init / cleanup (src/RT/RTEngine.cu:183-192, :455-472) ---- */
const char* hr_last_error(void);
hr_status hr_create(hr_ctx** out, const hr_options* opts);
hr_status hr_destroy(hr_ctx* ctx);
/* the context's HIP stream (hipStream_t), for callers that time or order work:
 * the stream of the lane that ran the last captured frame (hr_tuning.frame_lanes),
 * on which the context's other work runs until the next captured frame */
hr_status hr_get_stream(hr_ctx* ctx, void** stream);

/* ---- geometry: RTEngine::upload(frame1, frame2, coeff) (src/RT/RTEngine.cu:194-215,
 *      src/Application/WFObjectUploader.h:48-119).  v = v1*(1-c) + v2*c per vertex,
 *      bounds padded by 0.001*|diagonal|; interpolation and bounds run on the GPU. */
hr_status hr_upload_mesh(hr_ctx* ctx, const float* xyz1, const float* xyz2, float coeff,
                         uint32_t nv, const uint32_t* idx, uint32_t nt);
/* PhongMaterial table (12 floats each, src/RT/Primitive/Material.hpp:37-84) and one
 * material index per triangle (WFObjectUploader.h:254-330). */
hr_status hr_upload_materials(hr_ctx* ctx, const float* mats, uint32_t nmat,
                              const uint32_t* mat_idx, uint32_t nt);
/* Key frames resident in device memory (AnimationManager, src/Application/
 * AnimationManager.hpp:46-138): xyz holds nframes blocks of SoA vertices. */
hr_status hr_set_keyframes(hr_ctx* ctx, const float* xyz, uint32_t nframes, uint32_t nv,
                           const uint32_t* idx, uint32_t nt);
/* Interpolate resident key frames k1, k2 with coefficient c on the GPU
 * (AnimationManager::getFrameId / getNextFrameId / getInterpolationCoefficient). */
hr_status hr_set_frame(hr_ctx* ctx, uint32_t k1, uint32_t k2, float coeff);
/* Per-corner vertex normals (ObjUploader::uploadObjFrameNormalData /
 * uploadObjFrameNormalIndexData, src/Application/WFObjectUploader.h:121-252, called
 * from src/RT/RTEngine.cu:209-213): normals holds nframes blocks of SoA xyz (nn per
 * block), nidx three normal indices per triangle.  nframes is 1 (one block for every
 * frame) or the number of geometry key frames (2 after hr_upload_mesh with two
 * frames).  The path tracer interpolates them at the hit point
 * (src/RT/Integrator/PathTracer.h:109-135); the normals used are those of the first
 * key frame of the current frame, as the reference uploads them with coefficient 0.
 * Without normals the path tracer shades with the geometric normal.  Uploading new
 * geometry drops the normals. */
hr_status hr_set_normals(hr_ctx* ctx, const float* normals, uint32_t nframes, uint32_t nn,
                         const uint32_t* nidx, uint32_t nt);

/* ---- instanced hierarchy (HR_GRID_HIERARCHY) ----
 * One instance of an object: the world-to-object transform as the rows
 * m_row_0..2 of the instances file (WFObject::loadInstances,
 * src/Application/WFObjectReader.cpp:762-824; the fourth column is the
 * translation) and the instance's world-space box (AABB_min / AABB_max). */
typedef struct {
    uint32_t object;
    float m[3][4];
    float aabb_min[3], aabb_max[3];
} hr_instance;

/* Objects and instances of the current mesh (replaces RTEngine::upload's
 * TLGRIDHIERARCHY part, src/RT/RTEngine.cu:260-349).  Object k owns the
 * triangles [obj_first[k], obj_first[k+1]) (obj_first has nobj + 1 entries,
 * obj_first[0] == 0, non-decreasing; triangles past obj_first[nobj] are not
 * in any object and never hit).  Object grids are bounded by the object's
 * triangles on the first key frame of the current frame, padded by 1e-4
 * (RTEngine.cu:264-326); the top grid spans the padded scene bounds and holds
 * the instance boxes. */
hr_status hr_set_instances(hr_ctx* ctx, const uint32_t* obj_first, uint32_t nobj, const hr_instance* inst,
                           uint32_t ninst);

/* one object's uniform grid after an HR_GRID_HIERARCHY build; its cells are
 * entries [first_cell, first_cell + num_cells) of hr_get_grid's leaves array */
typedef struct {
    float bmin[3], bmax[3];
    int res[3];
    float cell_size[3], cell_size_rcp[3];
    uint32_t first_cell;
    uint32_t num_cells;
} hr_object_grid;

/* grids: nobj entries (may be NULL); inst_refs: the instance id of each top-cell
 * reference, hr_grid_info.num_top_pairs entries (may be NULL).  With this kind,
 * hr_get_grid returns the top cells (begin, end into inst_refs), the object
 * cells as leaves (begin, end into prims) and the object references as prims. */
hr_status hr_get_object_grids(hr_ctx* ctx, hr_object_grid* grids, uint32_t* inst_refs);

/* closest hits with the instance hit (TLGridHierarchyTraverser::getBestHitInstance,
 * src/RT/Algorithm/TLGridHierarchyTraverser.h:242; 0xFFFFFFFF on a miss).
 * hr_trace and hr_occluded work on the hierarchy too: an occlusion ray stops
 * once its distance is below 0.9999 (TLGridHierarchyTraverser.h:262-267,
 * :366-372) and is occluded when its final distance is below tmax. */
hr_status hr_trace_instanced(hr_ctx* ctx, const float* org, const float* dir, uint32_t n, float* t_out,
                             uint32_t* prim_out, uint32_t* inst_out);
/* instance hit by each pixel's primary ray in the last HR_GRID_HIERARCHY render */
hr_status hr_get_primary_instances(hr_ctx* ctx, uint32_t* inst);

/* ---- build: setGridDensities + buildAccStruct (src/RT/RTEngine.cu:354-368, :449-453) ---- */
hr_status hr_set_grid(hr_ctx* ctx, hr_grid_kind kind, float lambda_top, float lambda_leaf);
hr_status hr_build(hr_ctx* ctx);
hr_status hr_get_grid_info(hr_ctx* ctx, hr_grid_info* out);
/* Copy the grid to host memory.  cells: 2*num_cells uint32 (uniform: begin,end;
 * two-level: leafBegin, packed resolution|flags), leaves: 2*num_leaves (may be NULL
 * for uniform grids), prims: num_refs. */
hr_status hr_get_grid(hr_ctx* ctx, uint32_t* cells, uint32_t* leaves, uint32_t* prims);

/* ---- camera and lights: RTEngine::setCamera / setLights (src/RT/RTEngine.cu:370-396) ---- */
hr_status hr_set_camera(hr_ctx* ctx, const float pos[3], const float dir[3], const float up[3],
                        float fov_deg, int res_x, int res_y);
hr_status hr_set_lights(hr_ctx* ctx, const hr_area_light* lights, const float* weights, int n);

/* ---- render: RTEngine::renderFrame (src/RT/RTEngine.cu:398-442) ----
 * Accumulates into the context's float3 framebuffer with weight 1/(image_id+1)
 * (AOIntegrator.h:181-187); image_id 0 restarts the running mean.  rgb_out
 * (3*res_x*res_y floats) receives a copy when not NULL. */
hr_status hr_render(hr_ctx* ctx, hr_mode mode, int image_id, const hr_render_params* params,
                    float* rgb_out);
/* Captured frames (no reference counterpart: the reference runs upload, build and
 * render as synchronous phases, src/RT/RTEngine.cu:194-453).  One call renders a
 * frame of the resident key frames: interpolation of k1, k2 with coeff, a full
 * rebuild, primary rays and the integrator, with results equal to hr_set_frame +
 * hr_build + hr_render.  The rebuild is recorded once as a HIP graph (one per
 * frame slot) and replayed by later calls, followed by the render kernels; the
 * frame's (k1, k2, coeff, image_id) reaches the kernels through device memory.
 * The graphs are recorded again when the mode, render parameters, grid settings,
 * camera, light count, geometry, tuning or any buffer changes.  A replayed
 * rebuild that outgrows its buffers makes the device skip that frame's render
 * (every render kernel returns at entry, so the framebuffer and counters stay as
 * they were); the host finds it after the replay and renders the frame again
 * without graphs while the buffers grow.  hr_get_timing reports the rebuild graph
 * as build_ms and the render as render_ms, primary_ms, secondary_ms, shade_ms. */
hr_status hr_render_frame_graph(hr_ctx* ctx, uint32_t k1, uint32_t k2, float coeff, hr_mode mode,
                                int image_id, const hr_render_params* params, float* rgb_out);
/* Captured frames without waiting (no reference counterpart: the reference's
 * renderFrame returns after the frame, src/RT/RTEngine.cu:398-442).  The same
 * frame as hr_render_frame_graph with rgb_out NULL, but a replay is only
 * queued: the call returns once the frame submitted two calls earlier has
 * finished, so the next frame's rebuild follows this frame's render on the GPU
 * with no host round trip between them.  When the graphs must be recorded
 * (first frame, changed state) the call renders without graphs and waits, as
 * hr_render_frame_graph does.  Every other call on the context first completes
 * the frames in flight, in order: it takes their grid info and timing, and a
 * frame whose replayed rebuild overflowed (skipped on the device, together
 * with any frame queued behind it) is rendered again without graphs then, so
 * the running mean over images stays exactly that of hr_render_frame_graph;
 * errors of that work are returned by that call.  hr_synchronize completes
 * them all. */
hr_status hr_submit_frame_graph(hr_ctx* ctx, uint32_t k1, uint32_t k2, float coeff, hr_mode mode,
                                int image_id, const hr_render_params* params);
/* The framebuffer after the last render (3*res_x*res_y floats), as rgb_out of
 * hr_render receives it; completes the captured frames in flight first. */
hr_status hr_get_framebuffer(hr_ctx* ctx, float* rgb_out);
/* Primary-ray hit buffer of the last render (SimpleRayBuffer, RayBuffers.h:34-195):
 * 7 SoA float planes (hit point xyz, direction xyz, t) and primitive ids. */
hr_status hr_get_primary_hits(hr_ctx* ctx, float* planes7, uint32_t* prims);
/* AO occlusion flags of the last render, which must be an AO render (HR_ERR_STATE
 * otherwise): 1 = occluded, ao_samples per pixel. */
hr_status hr_get_ao_occlusion(hr_ctx* ctx, uint8_t* occ);

/* ---- ray batches: the trace<> kernel with a RayLoader generator
 * (src/RT/Algorithm/RayTracingKernels.h:47-152, RayGenerators.h:347-364).
 * Closest hit: t_out = distance (FLT_MAX on a miss), prim_out = primitive id or
 * num_triangles on a miss. */
hr_status hr_trace(hr_ctx* ctx, const float* org, const float* dir, uint32_t n,
                   float* t_out, uint32_t* prim_out);
/* Any hit within (1e-4, tmax): occ_out[i] = 1 when occluded. */
hr_status hr_occluded(hr_ctx* ctx, const float* org, const float* dir, uint32_t n, float tmax,
                      uint8_t* occ_out);
hr_status hr_trace_dev(hr_ctx* ctx, const float* d_org, const float* d_dir, uint32_t n,
                       float* d_t_out, uint32_t* d_prim_out);
hr_status hr_occluded_dev(hr_ctx* ctx, const float* d_org, const float* d_dir, uint32_t n, float tmax,
                          uint8_t* d_occ_out);
/* Paths from given rays: each ray's closest hit starts a path of `depth` vertices
 * shaded as the path-tracing render shades a primary hit (src/RT/Integrator/
 * PathTracer.h:41-372), ray i taking the random streams of pixel i of image 0;
 * rgb_out receives 3 floats of radiance per ray (0 for rays that miss). */
hr_status hr_path_trace(hr_ctx* ctx, const float* org, const float* dir, uint32_t n, int depth,
                        float* rgb_out);
/* the same with device buffers; returns before the work finishes (hr_synchronize) */
hr_status hr_path_trace_dev(hr_ctx* ctx, const float* d_org, const float* d_dir, uint32_t n, int depth,
                            float* d_rgb_out);
/* No counterpart in the reference.  hr_path_trace_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and results, but the hit
 * records and the ray-fetch counter live in a fixed scratch array in the
 * library's device code (about 72 MiB per device, held until the process
 * ends), so the call allocates nothing.  n is at most 2,097,152 rays
 * (HR_ERR_ARG above).  Every context on a device shares that device's array:
 * calls that overlap on one device (other contexts' streams, other threads)
 * race and must be ordered by the caller.  The context must have been built;
 * returns before the work finishes (hr_synchronize). */
hr_status hr_path_trace_dev_noalloc(hr_ctx* ctx, const float* d_org, const float* d_dir, uint32_t n, int depth,
                                    float* d_rgb_out);

/* Kernel choices of a context (no reference counterpart: the reference fixes
 * its launch shapes at compile time, src/RT/Algorithm/RayTracingKernels.h:34-43). */
typedef struct {
    int persistent_primary;    /* camera rays: 0 one ray per lane, 1 persistent traversal
                                  with dynamic ray fetch, 3 one ray per lane with the
                                  wave's candidate tests batched, 2 (default) whichever of
                                  0, 1 and 3 traced the camera rays fastest: the first
                                  six timed frames after the mesh, key frames, grid,
                                  camera or tuning change run 0, 1, 3, 0, 1, 3 (each waits
                                  for the one before it to finish), later frames the one
                                  with the least time
                                  (hr_timing.primary_kernel shows which ran).  Same results
                                  either way; the instanced hierarchy takes 2 and 3 as 0 */
    int persistent_secondary;  /* AO rays and ray batches: 1 (default) persistent, 0 not */
    int refill;                /* idle lanes of a 64-lane wave that trigger a fetch,
                                  1..64 (default 56)                                     */
    int hier_refill;           /* the same for the AO rays of the instanced hierarchy,
                                  whose lengths vary more: 1..64 (default 12)            */
    int hier_bounds_blocks;    /* workgroups of the hierarchy build's object-bounds pass:
                                  0 (default) one triangle per thread on up to 2048
                                  workgroups; 1..2048 caps the count, so that each
                                  thread walks a run of several triangles              */
    int persistent_path;       /* path tracing (mode 2, paths from given rays): 0 one
                                  path per lane, 1 persistent paths with dynamic pixel
                                  fetch and batched shading (same results)             */
    int path_refill;           /* idle lanes of a wave that trigger fetching new
                                  pixels in persistent paths, 1..64                    */
    int path_shade;            /* lanes waiting to shade that trigger a wave's shading
                                  step in persistent paths, 1..64                      */
    int fuse_units;            /* persistent traversal with per-lane tests: 1 (default) a
                                  lane that steps to a new cell or leaf also tests its first
                                  candidate in the same loop iteration; 0 one unit of work
                                  (a triangle test or a grid step) per lane and iteration
                                  (same results)                                         */
    int test_batch;            /* persistent traversal: 1 (default) each lane hands up to 12
                                  candidates of its cell or leaf per loop iteration to its
                                  wave, which tests them all at full width; 0 each lane
                                  tests its own candidates (fuse_units applies).  Same
                                  results and work statistics; with 1 an any-hit ray may
                                  test up to 11 candidates past its first hit, which the
                                  statistics leave out                                  */
    int advance_steps;         /* test_batch 1: grid steps a lane may take per loop
                                  iteration to reach its next candidates, 1..4 (default 2) */
    int frame_lanes;           /* captured frames: 2 (default) two lanes of per-frame
                                  buffers, each with its own stream, taken in turn, so
                                  that a frame's rebuild and render overlap the previous
                                  frame's render; each lane renders into a buffer of its
                                  own, blended into the framebuffer in frame order.  3
                                  two lanes where a frame's rebuild runs beside the
                                  previous frame's camera rays and each frame's
                                  secondary rays run alone.  1 one lane.  Same results;
                                  the multi-kernel path tracer (mode 1) and renders with
                                  collect_stats use one lane                            */
    int camera_blocks;         /* camera rays one per lane (persistent_primary 0, or
                                  that kernel chosen under 2):
                                  k in 1..16 (default 4) runs k workgroups per CU whose
                                  waves take 8x8 pixel tiles from a counter, leaving
                                  wave slots to a neighbouring frame's work (the
                                  instanced hierarchy's camera kernel takes its
                                  occupancy, six, for any k); 0 one workgroup per 256
                                  pixels.  Same results                                  */
} hr_tuning;
hr_status hr_set_tuning(hr_ctx* ctx, const hr_tuning* t);
hr_status hr_get_tuning(hr_ctx* ctx, hr_tuning* t);

hr_status hr_get_timing(hr_ctx* ctx, hr_timing* out);
/* Read and clear the work counters. */
hr_status hr_get_work_stats(hr_ctx* ctx, hr_work_stats* out);
hr_status hr_synchronize(hr_ctx* ctx);

/* ---- batches of scenes (no reference counterpart: the reference renders one scene
 * per process, src/RT/RTEngine.h:36-68) ----
 * A batch keeps many independent scenes resident on one device.  hr_batch_build
 * builds the two-level grid of every scene in one segmented pass (the kernels of
 * hr_build over all scenes at once); hr_batch_trace_dev traces ray batches of any
 * scenes in persistent launches that mix the rays of all scenes.  Each scene's grid
 * (read back scene-local by hr_batch_get_grid) equals the grid hr_build makes for
 * that scene alone with the same densities, and each ray's result equals hr_trace /
 * hr_occluded on that grid.  Calls on one batch must not overlap. */
typedef struct hr_batch hr_batch;
typedef struct {
    const float* xyz;          /* host: SoA vertices, 3 * nv floats                */
    uint32_t nv;
    const uint32_t* idx;       /* host: 3 * nt vertex indices (scene-local)        */
    uint32_t nt;
} hr_mesh_desc;
typedef struct {
    uint32_t scene;            /* index into the batch's scenes                    */
    uint32_t n;                /* rays                                            */
    int anyhit;                /* 0: closest hit; 1: occluded within (1e-4, tmax) */
    float tmax;                /* any hit only                                    */
    const float* org;          /* device: SoA origins, 3 * n                       */
    const float* dir;          /* device: SoA directions, 3 * n                    */
    float* t_out;              /* device: closest-hit distance (optional for any)  */
    uint32_t* prim_out;        /* device: scene-local triangle id, nt on a miss    */
    uint8_t* occ_out;          /* device: any hit: 1 when occluded                 */
} hr_ray_batch;
typedef struct {
    int refill;                /* idle lanes of a wave that trigger a fetch, 1..64 */
    int blocks_per_cu;         /* persistent workgroups per CU, 0: occupancy limit */
} hr_batch_tuning;
typedef struct {
    float lambda_top[4];       /* per size class: < 4K, < 16K, < 64K, >= 64K tris */
    float lambda_leaf[4];
    hr_batch_tuning tuning;
    int evaluations;           /* timed candidates                                 */
    float best_build_trace_ms; /* build + trace of the chosen densities            */
    float best_trace_ms;       /* trace with the chosen persistence settings        */
} hr_autotune_result;
hr_status hr_batch_create(hr_batch** out, const hr_options* opts);
hr_status hr_batch_destroy(hr_batch* batch);
hr_status hr_batch_get_stream(hr_batch* batch, void** stream);
/* upload the scenes (replacing any previous ones); densities reset to 1.2 / 2.2 */
hr_status hr_batch_set_scenes(hr_batch* batch, const hr_mesh_desc* scenes, uint32_t n);
/* per-scene top-level and leaf densities (NULL: 1.2 / 2.2, src/RT/RTEngine.cu:79-80) */
hr_status hr_batch_set_densities(hr_batch* batch, const float* lambda_top, const float* lambda_leaf);
hr_status hr_batch_build(hr_batch* batch);
hr_status hr_batch_get_grid_info(hr_batch* batch, uint32_t scene, hr_grid_info* out);
/* scene-local copy of one scene's grid, in the layout of hr_get_grid */
hr_status hr_batch_get_grid(hr_batch* batch, uint32_t scene, uint32_t* cells, uint32_t* leaves,
                            uint32_t* prims);
/* enqueue the traces of nb ray batches (one closest-hit and one any-hit persistent
 * launch); returns before they finish (hr_batch_synchronize) */
hr_status hr_batch_trace_dev(hr_batch* batch, const hr_ray_batch* batches, uint32_t nb);
hr_status hr_batch_set_tuning(hr_batch* batch, const hr_batch_tuning* t);
/* current tuning and per-scene densities (arrays of the scene count, may be NULL) */
hr_status hr_batch_get_tuning(hr_batch* batch, hr_batch_tuning* t, float* lambda_top,
                              float* lambda_leaf);
/* Choose densities per scene-size class and the persistence settings by timing
 * builds and the traces of the given ray batches on the device; the choice stays
 * set on the batch, which is left built with it. */
hr_status hr_batch_autotune(hr_batch* batch, const hr_ray_batch* batches, uint32_t nb,
                            hr_autotune_result* out);
/* device times of the last build and the last trace; rays traced so far */
hr_status hr_batch_get_timing(hr_batch* batch, double* build_ms, double* trace_ms,
                              uint64_t* rays_traced);
/* read and clear the work counters (primary: closest hit, secondary: any hit) */
hr_status hr_batch_get_work_stats(hr_batch* batch, hr_work_stats* out);
hr_status hr_batch_synchronize(hr_batch* batch);

/* ---- request server (no reference counterpart) ----
 * A server keeps scenes resident on one device and serves trace, occlusion, path and
 * rebuild requests submitted from any thread.  A worker thread batches continuously:
 * while one wave of requests runs on the GPU, requests that arrive are admitted into
 * the next wave as soon as a wave slot and device memory are free; a wave traces the
 * rays of all its requests, of any scenes, in persistent launches (hr_batch_trace_dev's
 * kernels).  A scene's grid is built on first use and evicted least-recently-used when
 * the memory budget requires; a rebuild request (new key-frame interpolation) builds
 * into a second grid while earlier waves still read the first.  Each request's result
 * equals the single-scene entry point's result for its rays (hr_trace, hr_occluded,
 * hr_path_trace) on its scene's current frame, alone or batched with others. */
typedef struct hr_server hr_server;
typedef struct {
    int device;
    uint64_t memory_budget;    /* device bytes for grids and wave scratch (0: 200 GiB) */
    uint32_t max_wave_rays;    /* rays admitted into one wave (0: 16M); a larger request
                                  forms a wave alone                                    */
    uint32_t seed_lo, seed_hi; /* RNG key of path requests                               */
} hr_server_options;
typedef enum {
    HR_REQ_CLOSEST = 0,        /* t_out, prim_out (either may be NULL)                  */
    HR_REQ_OCCLUDED = 1,       /* occ_out within (1e-4, tmax)                           */
    HR_REQ_PATH = 2,           /* rgb_out: 3 floats per ray, `depth` path vertices      */
    HR_REQ_REBUILD = 3         /* the scene's frame becomes (k1, k2, coeff)              */
} hr_request_kind;
typedef struct {
    int kind;                  /* hr_request_kind                                      */
    uint32_t scene;
    uint32_t n;                /* rays (at most 2^30)                                   */
    float tmax;
    int depth;
    int on_device;             /* 1: rays and outputs are device pointers; NULL outputs
                                  are computed and dropped                              */
    const float* org;          /* SoA: x, y and z planes of n floats, `stride` apart   */
    const float* dir;
    uint32_t stride;           /* floats between the planes (0: n)                     */
    float* t_out;
    uint32_t* prim_out;
    uint8_t* occ_out;
    float* rgb_out;
    uint32_t k1, k2;           /* rebuild                                               */
    float coeff;
    int layout;                /* HR_RAYS_SOA, or HR_RAYS_RECORDS: org points to n 24-byte
                                  ray records (dir and stride unused)                  */
} hr_request;
enum { HR_RAYS_SOA = 0, HR_RAYS_RECORDS = 1 };
typedef struct {
    int status;                /* hr_status of the request                              */
    double latency_ms;         /* submit to completion                                  */
    uint64_t wave;             /* wave that ran it (0: none)                            */
} hr_request_result;
typedef struct {
    const float* keys;         /* host: nframes blocks of SoA vertices (3 * nv each)    */
    uint32_t nframes, nv;
    const uint32_t* idx;
    uint32_t nt;
    const float* mats;         /* 12 floats each, may be NULL                           */
    uint32_t nmat;
    const uint32_t* mat_idx;
    const hr_area_light* lights;
    const float* light_weights;
    int nlights;
    float lambda_top, lambda_leaf;   /* two-level densities (0: 1.2 / 2.2)              */
} hr_server_scene;
typedef struct {
    uint64_t requests, rays, waves, builds, rebuilds, evictions;
    uint64_t resident_bytes, peak_resident_bytes;
    uint32_t resident_scenes;
} hr_server_stats;
hr_status hr_server_create(hr_server** out, const hr_server_options* opts);
/* finishes the requests already submitted, then releases everything */
hr_status hr_server_destroy(hr_server* server);
/* the scene's data is copied; it becomes resident on first use (frame (0, 0, 0)) */
hr_status hr_server_add_scene(hr_server* server, const hr_server_scene* scene, uint32_t* id);
/* the caller keeps the request's buffers alive until hr_server_wait returns */
hr_status hr_server_submit(hr_server* server, const hr_request* request, uint64_t* id);
/* blocks until the request is done; returns its status (result may be NULL) */
hr_status hr_server_wait(hr_server* server, uint64_t id, hr_request_result* result);
hr_status hr_server_poll(hr_server* server, uint64_t id, int* done);
hr_status hr_server_get_stats(hr_server* server, hr_server_stats* out);

/* ---- scene files: SceneLoader::loadScene (src/Application/SceneLoader.cpp:212-332) ----
 * hr_scene_load parses a scene.cfg (src/Application/SceneConfigurationReader.cpp:29-190)
 * and the files it names: OBJ/MTL key frames (src/Application/WFObjectReader.cpp), a
 * camera.cfg (src/Application/CameraManager.cpp:96-168) and a light.cfg
 * (src/Application/LightSourceLoader.cpp:29-100).  Without a light file, area lights
 * are derived from emissive quads of the first key frame, or one default light is
 * placed above the scene (SceneLoader.cpp:94-210).  Relative paths are tried against
 * the working directory first, then against the directory of the scene.cfg.
 * Everything stays in host memory until hr_scene_upload; no GPU is needed to parse.
 * All key frames must have the same vertex, normal and triangle counts. */
typedef struct hr_scene hr_scene;
typedef struct {
    uint32_t num_frames;          /* key frames                                         */
    uint32_t num_vertices, num_normals, num_triangles, num_materials;
    uint32_t num_lights;
    float frame_step;             /* AnimationManager step size (scene.cfg "frame step") */
    int has_camera;               /* 1 when the scene.cfg names a camera file            */
    float cam_pos[3], cam_dir[3], cam_up[3];   /* after the camera file's rotation      */
    float cam_fov;
    int cam_res_x, cam_res_y;
    uint32_t num_objects;         /* objects of the first key frame ('o' lines)          */
    uint32_t num_instances;       /* entries of the scene.cfg's instances file           */
} hr_scene_info;
hr_status hr_scene_load(const char* cfg_path, hr_scene** out);
/* one OBJ file as a static scene (AnimationManager::read(const char*) + light setup) */
hr_status hr_scene_load_obj(const char* obj_path, hr_scene** out);
hr_status hr_scene_free(hr_scene* scene);
hr_status hr_scene_get_info(const hr_scene* scene, hr_scene_info* out);
/* SoA xyz of key frame `frame`: 3*num_vertices floats (vertices) or 3*num_normals
 * floats (normals) */
hr_status hr_scene_get_vertices(const hr_scene* scene, uint32_t frame, float* xyz);
hr_status hr_scene_get_normals(const hr_scene* scene, uint32_t frame, float* xyz);
/* 3 vertex indices, 3 normal indices and 1 material index per triangle; any may be NULL */
hr_status hr_scene_get_triangles(const hr_scene* scene, uint32_t* idx, uint32_t* nidx,
                                 uint32_t* mat_idx);
/* PhongMaterial rows after Material::setupPhongCoefficients
 * (src/Application/WFObjectReader.cpp:49-57): 12 floats per material */
hr_status hr_scene_get_materials(const hr_scene* scene, float* mats);
/* area lights and their sampling weights (area * |intensity|) */
hr_status hr_scene_get_lights(const hr_scene* scene, hr_area_light* lights, float* weights);
/* objects as triangle ranges (num_objects + 1 entries, see hr_set_instances) and the
 * instances of the scene.cfg's "instances" file (WFObject::loadInstances,
 * src/Application/WFObjectReader.cpp:762-824: an obj_id line starts an instance with
 * the identity transform; m_row_0..2 set the world-to-object rows, AABB_min/AABB_max
 * its box, which is unbounded until given); either may be NULL.  hr_scene_upload
 * passes them to hr_set_instances when the scene has instances. */
hr_status hr_scene_get_instances(const hr_scene* scene, uint32_t* obj_first, hr_instance* inst);
/* Make the scene a context's geometry: key frames resident (frame 0 current),
 * normals, materials, lights and, when the scene has one, the camera. */
hr_status hr_scene_upload(hr_ctx* ctx, const hr_scene* scene);
/* AnimationManager frame arithmetic (src/Application/AnimationManager.hpp:68-112):
 * the key frames and coefficient of animation time `current`, and the next time. */
hr_status hr_scene_frame(const hr_scene* scene, float current, uint32_t* k1, uint32_t* k2,
                         float* coeff);
hr_status hr_scene_next_frame(const hr_scene* scene, float* current);

/* ---- ray records: the ray file of SimpleRayTraverser::loadRaysFromFile
 * (src/RT/Integrator/SimpleRayTraverser.h:200-262) and AORayExporter's dump
 * (src/RT/Integrator/AORayExporter.h:50-68, :262-310) ----
 * A ray record is 24 bytes: origin x, y, z then direction x, y, z as float32;
 * records follow each other with no header (a file of b bytes holds b / 24 rays). */
/* closest hits / occlusion of n host ray records (as hr_trace / hr_occluded) */
hr_status hr_trace_wire(hr_ctx* ctx, const float* rays, uint32_t n, float* t_out, uint32_t* prim_out);
hr_status hr_occluded_wire(hr_ctx* ctx, const float* rays, uint32_t n, float tmax, uint8_t* occ_out);
/* renderFrame mode 0 with input rays (RTEngine.cu:410-413, SimpleRayTraverser::integrate):
 * the closest hit of each record, then simpleShade at image 0: 3 floats per ray */
hr_status hr_render_rays(hr_ctx* ctx, const float* rays, uint32_t n, float* rgb_out);
/* The rays of the last AO render as AORayExporter dumps them: the npix camera rays
 * (camera position, unnormalized direction), then ao_samples AO rays per pixel
 * (pixel-major), a missed pixel's records holding its stored hit point and camera
 * direction; (npix * (ao_samples + 1)) records into rays_out, which holds
 * max_rays records (HR_ERR_ARG when fewer; rays_out NULL: *count only).  *count
 * (may be NULL) receives the number of records.  The last render must have
 * been an AO render of a grid (HR_ERR_STATE otherwise). */
hr_status hr_dump_ao_rays(hr_ctx* ctx, float* rays_out, uint64_t max_rays, uint64_t* count);

/* ---- image output (host only): Image::gammaCorrect + Image::writePNG
 * (src/Utils/ImagePNG.cpp:64-72, :268-330) ----
 * rgb: 3 floats per pixel, row-major from the top-left pixel, as the framebuffer.
 * hr_image_to_rgb8: powf(v, 1/gamma) per channel (gamma <= 0: none), clamped to
 * [0, 1], times 255, truncated to a byte (NaN -> 0).
 * hr_write_png: an 8-bit RGB PNG (no interlace, filter 0, zlib deflate).
 * hr_write_screenshot: both, with gamma 2.2, as SDLGLApplication::writeScreenShot
 * (src/Application/SDLGLApplication.cpp:120-143). */
hr_status hr_image_to_rgb8(const float* rgb, int width, int height, float gamma, uint8_t* out);
hr_status hr_write_png(const char* path, const uint8_t* rgb8, int width, int height);
hr_status hr_write_screenshot(const char* path, const float* rgb, int width, int height);

#ifdef __cplusplus
}
#endif
#endif
