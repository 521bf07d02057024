// hr_build.h -- host interface of the grid build kernels (hr_build.hip).
#pragma once

#include "hr_device.h"

namespace hr {

struct BuildArgs {
    int kind;                   // 0 uniform, 1 uniform exact, 2 two-level
    float lambda_top, lambda_leaf;
    const uint32_t* idx;        // device: 3 * nt vertex indices
    const float* vx;            // device: SoA vertices
    const float* vy;
    const float* vz;
    uint32_t nt;
};

// Device buffers of one build.  Capacities are element counts; the host grows
// them after an overflow and repeats the build.
struct BuildBuffers {
    GridParams* params;
    BuildStatus* status;
    float4* rec;                // triangle records: 9 floats per triangle
    uint32_t* wave_counts;      // per-wave pair counts
    uint32_t* wave_offsets;     // their exclusive scan
    uint32_t* keys;             // unsorted keys / values (pairs, then leaf pairs)
    uint32_t* vals;
    uint32_t* keys_sorted;
    uint32_t* top_vals_sorted;  // two-level: sorted top-level primitive ids
    uint32_t* prims;            // final primitive references
    uint2* cells;               // cells_cap
    uint2* leaves;              // leaves_cap
    uint32_t* leaf_counts;      // cells_cap
    uint64_t cells_cap;
    uint32_t pairs_cap;
    uint32_t leaves_cap;
    uint32_t leaf_pairs_cap;
    uint32_t waves_cap;         // entries of wave_counts
    ScanTemp scan;
    SortTemp sort;
};

// The scenes of a batch, built together: the triangles, cells, leaves and
// references of all scenes live in one set of arrays, and the cell keys are
// global cell indices (scene s owns cells [cell_base[s], cell_base[s+1])).
struct SceneTable {
    uint32_t ns;
    const uint32_t* tri_scene;   // per triangle: its scene
    const uint32_t* tri_base;    // ns + 1 triangle offsets
    const uint32_t* vert_base;   // ns + 1 vertex offsets
    const float* lambda_top;     // per scene
    const float* lambda_leaf;    // per scene
    GridParams* gps;             // per scene
    uint32_t* bounds_enc;        // 6 per scene
    uint32_t* cell_counts;       // per scene
    uint32_t* cell_base;         // ns + 1
    uint32_t* top_base;          // ns + 1: first top-level pair of each scene
    uint32_t* leaf_base;         // ns + 1: first leaf of each scene
    uint32_t* pair_base;         // ns + 1: first leaf-level reference of each scene
    int object_grids;            // 1: the object grids of an instanced hierarchy, whose cell
                                 // ranges clamp the first cell to res - 1 and the end to >= 1
};

void launch_reset(BuildStatus* st, hipStream_t s);
// fs != nullptr: the key frames are blocks fs->k1, fs->k2 of keys and the
// coefficient is fs->coeff, all read on the device
// part: device scratch of 6 * kBoundsParts words for the per-block partial bounds
constexpr uint32_t kBoundsParts = 1024;
void launch_lerp_bounds(const float* k1, const float* k2, float c, uint32_t nv, float* out,
                        BuildStatus* st, uint32_t* part, hipStream_t s, const FrameSel* fs = nullptr,
                        const float* keys = nullptr);
void launch_grid_setup(GridParams* gp, BuildStatus* st, uint32_t nprims, float lambda,
                       uint64_t cells_cap, hipStream_t s);
// everything after the bounds: setup, pairs, sort, ranges (and the leaf level)
void run_build(const BuildArgs& a, BuildBuffers& b, hipStream_t s);
// two-level build of every scene of a batch in one pass: reset, per-scene
// bounds, per-scene resolution, then the same count / scan / write / sort /
// range kernels over all scenes at once.  a.idx holds global vertex indices,
// a.nt is the total triangle count; b.params receives the total cell count.
void run_build_batch(const BuildArgs& a, BuildBuffers& b, const SceneTable& T, uint32_t nv_total,
                     hipStream_t s);

}  // namespace hr
