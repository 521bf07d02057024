// hr_hier.h -- the instanced two-level grid hierarchy: uniform grids over the
// triangles of each object, a uniform grid over instance boxes on top, and a
// traversal that steps the top grid and, for each instance in a cell, the
// instance's object grid in the instance's own coordinates.
//
// Reference behaviour followed:
//   GeometryInstanceMatrix           src/RT/Structure/TwoLevelGridHierarchy.h:569-696
//   TwoLevelGridHierarchy            src/RT/Structure/TwoLevelGridHierarchy.h:750-977
//   object grids (bounds, res)       src/RT/RTEngine.cu:264-330,
//                                    src/RT/Algorithm/TLGridHierarchySortBuilder.cu:29-76
//   multi-grid pairs                 src/RT/Algorithm/TLGridHierarchySortBuilder.cu:109-255
//   top grid over instances          src/RT/Algorithm/TLGridHierarchySortBuilder.cu:259-520
//   traversal                        src/RT/Algorithm/TLGridHierarchyTraverser.h:45-400
#pragma once

#include "hr_build.h"
#include "hr_trace.h"

namespace hr {

// one instance: world-to-object transform (rows irot0..2 as the reference
// stores them, plus a translation) and the world-space box of the instance
struct InstanceDev {
    float bmin[3], bmax[3];
    float irot0[3], irot1[3], irot2[3], it[3];
    uint32_t object;
    uint32_t pad;
};

// Per instance: its object's grid and first object cell, gathered by the
// build so that the traversal loads them beside the instance rather than
// after it (the object id is no longer on the load path).
struct InstGrid {
    float bmin[3], bmax[3], cs[3], csr[3];
    int res[3];
    uint32_t base;
};

// Everything the hierarchy traversal reads, by device pointer.
struct HierScene {
    const GridParams* top;          // uniform grid over the instance boxes
    const uint2* top_cells;         // (begin, end) into inst_refs
    const uint32_t* inst_refs;      // instance index per top-level reference
    const InstanceDev* inst;
    const InstGrid* inst_grid;      // per instance: its object's grid
    const GridParams* obj_gp;       // per object: its uniform grid
    const uint32_t* obj_cell_base;  // per object: its first cell in obj_cells
    const uint2* obj_cells;         // every object's cells, (begin, end) into prims
    const uint32_t* prims;          // every object's primitive references
    const float4* rec;              // triangle records (object coordinates)
    uint32_t nt;
};

struct HierArgs {
    const uint32_t* idx;
    const float *vx, *vy, *vz;      // current-frame vertices (object coordinates)
    const float *kx, *ky, *kz;      // first key frame of the current frame: object bounds
    uint32_t nt;                    // triangles of the objects: [0, nt) (later ones are not built)
    uint32_t nobj;
    const uint32_t* tri_obj;        // per triangle: its object
    const uint32_t* obj_first;      // nobj + 1: first triangle of each object
    uint32_t ninst;
    const InstanceDev* inst;
    float lambda_top, lambda_leaf;
    uint32_t bounds_blocks;         // cap on the object-bounds workgroups (0: kObjBoundsBlocks)
};

struct HierBuffers {
    BuildStatus* status;            // bounds of all vertices (from the lerp), counts, overflow bits
    GridParams* top;
    GridParams* obj_gp;             // nobj
    uint32_t* obj_bounds_enc;       // 6 per object
    uint32_t* obj_part;             // 7 * kObjBoundsBlocks: per-block partial object bounds
    uint32_t* obj_cell_counts;      // nobj
    uint32_t* obj_cell_base;        // nobj: first cell of each object in obj_cells
    InstGrid* inst_grid;            // ninst: each instance's object grid
    float4* rec;
    uint32_t *wave_counts, *wave_offsets;
    uint32_t waves_cap;
    uint32_t *keys, *vals, *keys_sorted;
    uint32_t* prims;                // object references (leaf_pairs_cap)
    uint2* obj_cells;               // leaves_cap
    uint32_t *inst_counts, *inst_offsets;   // ninst
    uint2* top_cells;               // cells_cap
    uint32_t* inst_refs;            // pairs_cap
    uint64_t cells_cap;             // top cells
    uint32_t pairs_cap;             // top references
    uint32_t leaves_cap;            // object cells
    uint32_t leaf_pairs_cap;        // object references
    ScanTemp scan;
    SortTemp sort;
};

constexpr uint32_t kObjBoundsBlocks = 2048;   // workgroups of the object-bounds pass (at most)

// everything after the lerp / bounds: object grids, then the top grid
void run_build_hier(const HierArgs& a, HierBuffers& b, hipStream_t s);

// closest hits (prim_out, inst_out) or occlusion of n SoA rays
void launch_trace_hier(const HierScene& H, const float* org, const float* dir, uint32_t n, int anyhit,
                       float tmax, float* t_out, uint32_t* prim_out, uint32_t* inst_out, uint8_t* occ_out,
                       unsigned long long* ray_count, hipStream_t s);
// TLGHAOIntegrator (src/RT/Integrator/TLGridHierarchyAOIntegrator.h:421-573):
// primary rays with instance ids, AO rays from the instance's turned normal,
// the AO shading of computeAOIlluminationTLGH
// pers_primary / pers_ao: the camera / AO rays run in the persistent kernel,
// which fetches rays through the device word *pool (reset here) whenever
// `refill` lanes of a wave are idle; batch: with the triangle tests batched
// across the wave (the same outputs); after_primary / after_ao (or nullptr):
// events recorded after the camera rays and after the AO rays; camera_blocks
// > 0: the one-ray-per-lane camera kernel runs that many workgroups per CU,
// whose waves take 8x8 pixel tiles through *pool (the same outputs)
//
// The hierarchy's camera kernel runs alone, so it takes tiles from a counter
// at its occupancy (six workgroups of four waves per CU): 1024x1024 camera
// rays on 10,000 instances 4.71 -> 4.44 ms (four per CU: 4.92 ms)
constexpr uint32_t kHierCameraBlocks = 6;
void launch_render_ao_hier(const HierScene& H, const CameraDev& cam, const RenderArgs& ra, float* hits,
                           uint32_t* hit_prims, uint32_t* hit_inst, uint8_t* occ, const Shading& sh, float* fb,
                           bool pers_primary, bool pers_ao, uint32_t* pool, uint32_t refill,
                           unsigned long long* ray_count, hipStream_t s, bool batch = false,
                           hipEvent_t after_primary = nullptr, hipEvent_t after_ao = nullptr,
                           uint32_t camera_blocks = 0);

}  // namespace hr
