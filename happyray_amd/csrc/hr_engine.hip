// hr_engine.hip -- host side of the library: the context that replaces the
// reference's static RTEngine (src/RT/RTEngine.h:36-68, src/RT/RTEngine.cu:91-474)
// and the C ABI declared in include/happyray.h.
//
// The context owns one HIP stream (default priority) and grow-only device
// buffers (MemoryManager::allocate*, src/RT/Structure/MemoryManager.h:73-139).
// A build enqueues interpolation, bounds and the whole grid build without a
// host round trip, then reads one status block; when a capacity was too
// small the buffers grow to the reported size and the build runs again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/happyray.h"
#include "hr_build.h"
#include "hr_ctx_internal.h"
#include "hr_error.h"
#include "hr_hier.h"
#include "hr_host.h"
#include "hr_persist.h"
#include "hr_trace.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

hr_status hr::set_error(hr_status s, const std::string& msg)
{
    g_last_error = msg;
    return s;
}

using hr::DevBuf;
using hr::fail;


struct hr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t seed_lo = 0, seed_hi = 0;

    // geometry sources: the frame is (src1, src2, coeff)
    uint32_t nv = 0, nt = 0;
    DevBuf<float> key_a, key_b;       // hr_upload_mesh sources
    DevBuf<float> keyframes;          // hr_set_keyframes
    uint32_t nframes = 0;
    uint32_t geo_frames = 0;          // key frames of the geometry (2 after hr_upload_mesh with two)
    uint32_t k1 = 0, k2 = 0;          // key frames of the current frame
    const float* src1 = nullptr;
    const float* src2 = nullptr;
    float coeff = 1.f;

    // per-corner normals
    DevBuf<float> normal_keys;        // nframes_n blocks of SoA xyz
    DevBuf<uint32_t> nidx;
    DevBuf<float4> normals;           // current frame
    uint32_t nn = 0, nframes_n = 0;
    bool has_normals = false, normals_dirty = false;
    DevBuf<float> verts;
    DevBuf<uint32_t> idx;

    DevBuf<float> mats;
    DevBuf<uint32_t> mat_idx;
    bool has_mat_idx = false;

    int kind = HR_GRID_TWO_LEVEL;
    float lambda_top = 1.2f, lambda_leaf = 2.2f;   // src/RT/RTEngine.cu:79-80

    // build
    DevBuf<hr::GridParams> params;
    DevBuf<hr::BuildStatus> status;
    DevBuf<uint32_t> bounds_part;     // per-block partial bounds of the interpolation
    hr::BuildStatus* h_status = nullptr;  // pinned
    DevBuf<float4> rec;
    DevBuf<uint32_t> wave_counts, wave_offsets, keys, vals, keys_sorted, top_vals_sorted, prims, leaf_counts;
    DevBuf<uint2> cells, leaves;
    DevBuf<uint32_t> scan_chain, sort_chain, sort_keys_tmp, sort_vals_tmp, sort_hist;
    uint64_t cells_cap = 0;
    uint32_t pairs_cap = 0, leaves_cap = 0, leaf_pairs_cap = 0;
    bool built = false;
    hr_grid_info info{};

    // instanced hierarchy (HR_GRID_HIERARCHY): objects are triangle ranges
    // [obj_first[k], obj_first[k+1]) starting at triangle 0; its build reuses
    // cells (top grid), top_vals_sorted (instance references), leaves (object
    // cells) and prims (object references)
    uint32_t nobj = 0, ninst = 0, obj_tris = 0;
    DevBuf<uint32_t> tri_obj, obj_first, obj_bounds_enc, obj_cell_counts, obj_cell_base, inst_counts, inst_offsets,
        obj_part;
    DevBuf<hr::InstanceDev> inst;
    DevBuf<hr::InstGrid> inst_grid;
    DevBuf<hr::GridParams> obj_gp;
    DevBuf<uint32_t> hit_inst;        // instance of each pixel's primary hit (hierarchy renders)
    DevBuf<uint32_t> ray_inst;        // instance of each traced ray's hit (hr_trace_instanced)
    uint32_t hier_npix = 0;           // pixels of the last render when it was a hierarchy render
    bool ao_dump_ok = false;          // the last render was an AO render of a grid (hits, last_ao_*)
    uint32_t occ_npix = 0;            // pixels of the last render when it was an AO render
    hr::GridParams h_params{};

    // camera / lights
    hr::CameraDev cam{};
    bool has_camera = false;
    DevBuf<float> lights, light_cdf;
    int nlights = 0;

    // render
    DevBuf<float> hits, fb;
    DevBuf<uint32_t> hit_prims;
    DevBuf<uint8_t> occ;
    uint32_t last_npix = 0;
    int last_ao_samples = 0;
    // multi-kernel path tracer: end vertices (7 planes), then 7 rgb buffers
    // (importance, light, light direction, new radiance, residue, blended,
    // contribution); cleared with the camera
    DevBuf<float> pm;
    DevBuf<uint32_t> pm_prims;
    uint32_t pm_npix = 0;

    // ray batches
    DevBuf<float> ray_org, ray_dir, ray_t, ray_hits, ray_rgb, ray_wire;
    hr::RenderArgs last_ao_ra{};          // the last AO render's arguments and camera (ray dump)
    hr::CameraDev last_ao_cam{};
    DevBuf<uint32_t> ray_prim;
    DevBuf<uint8_t> ray_occ;

    // timing / counters
    hipEvent_t ev_b0 = nullptr, ev_b1 = nullptr, ev_r0 = nullptr, ev_r1 = nullptr;
    hipEvent_t ev_p1 = nullptr, ev_p2 = nullptr;
    bool build_timed = false, render_timed = false;
    DevBuf<unsigned long long> ray_count;
    DevBuf<unsigned long long> work;     // {rays, steps, tests} x {primary, secondary}
    DevBuf<uint32_t> pool;               // persistent kernels' ray counter
    hr_tuning tuning{ 2, 1, 56, 12, 0, 0, 32, 16, 1, 1, 2, 2, 4 };
    // persistent_primary 2: the camera-ray kernel is chosen by timing two frames
    // of each (0 one ray per lane, 1 persistent, 3 wave-batched tests) after the
    // scene, grid or camera changes; results are the same either way
    float pp_ms[3] = { 0.f, 0.f, 0.f };  // least time of each kernel's trials
    int pp_measured = 0;                 // trial frames timed (0 .. 6)
    int pp_pending = -1;                 // kernel of an eager frame whose time is unread
    int pp_choice = -1;                  // the fastest kernel once all are timed
    int last_pk = -1;                    // camera-ray kernel of the last render

    // captured frames: per slot (0, 1) a rebuild graph, which reads the frame
    // selection from and writes its status into the slot's pinned memory,
    // replayed per frame and followed by the frame's render kernels, which
    // read the same selection from device memory.  Up to two submitted frames
    // are in flight, in alternate slots; pending[0] is the older.
    struct FramePin {                        // pinned, one per slot
        hr::FrameSel sel;                    // source of the frame selection
        hr::BuildStatus status;              // build status after the rebuild
        hr::GridParams params;               // grid parameters after the rebuild
        uint32_t halt;                       // the halt word after the rebuild
    };
    struct Submitted {
        uint32_t k1, k2;
        float coeff;
        int image;
        int slot;
        uint32_t seq;
    };
    hipGraphExec_t g_build[2] = { nullptr, nullptr };
    std::vector<unsigned char> graph_sig[2]; // per lane: state the graphs were recorded with
    hr_mode graph_mode = HR_MODE_AO;         // the graphs' mode and render parameters
    hr_render_params graph_params{};
    int graph_lanes = 1;                     // lanes of the recorded graphs (slot s runs on lane s
                                             // with two, on lane 0 with one)
    DevBuf<uint32_t> halt;                   // least sequence number of an overflowed replay
    FramePin* h_pin = nullptr;
    hipEvent_t ev_g[2][4] = {};              // per slot: rebuild start, rebuild end, render end,
                                             // frame end (after its blend)
    hipEvent_t ev_k[2][5] = {};              // per slot, in the render: camera rays start / end,
                                             // secondary rays start / end, shading end
    hipEvent_t ev_built[2] = {};             // per lane: its last rebuild done (ordered lanes)
    hipEvent_t ev_blend = nullptr;           // the last blend into the framebuffer (ordered lanes)
    bool graph_ordered = false;              // two lanes, each frame's secondary rays alone
    bool deferred = false;                   // the newest frame in flight waits for its secondary part
    hipEvent_t ev_switch = nullptr;          // the lane left at the last lane switch: its work so far
    Submitted pending[2] = {};
    int npending = 0, next_slot = 0;
    uint32_t frame_seq = 0;
    float graph_build_ms = 0.f, graph_render_ms = 0.f;   // the last completed replay
    float graph_primary_ms = 0.f, graph_secondary_ms = 0.f, graph_shade_ms = 0.f;
    uint64_t replays = 0;                    // completed replays and the sums of their times
    double replay_ms[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 };   // build, render, primary, secondary, shade
    bool graph_timed = false;

    // Frame lanes: with two, captured frames alternate between two sets of
    // per-frame buffers, each with its own stream, so that one frame's rebuild
    // and render overlap the other's; each lane renders into its own buffer
    // (rad) and k_frame_blend blends it into fb in frame order.  The fields
    // above and below named in LaneFields hold lane `lane`; `stash` holds the
    // other lane's.  Everything else (geometry, materials, lights, camera,
    // framebuffer, counters) is shared.
    DevBuf<hr::FrameSel> frame_sel;
    DevBuf<float> rad;                       // a lane's frame values (3 floats per pixel)
    DevBuf<unsigned long long> frame_rays;   // a lane's rays of its frame, added to ray_count by the blend
    struct LaneFields {
        hipStream_t stream = nullptr;
        DevBuf<hr::GridParams> params;
        DevBuf<hr::BuildStatus> status;
        DevBuf<uint32_t> bounds_part;
        DevBuf<float4> rec;
        DevBuf<uint32_t> wave_counts, wave_offsets, keys, vals, keys_sorted, top_vals_sorted, prims, leaf_counts;
        DevBuf<uint2> cells, leaves;
        DevBuf<uint32_t> scan_chain, sort_chain, sort_keys_tmp, sort_vals_tmp, sort_hist;
        uint64_t cells_cap = 0;
        uint32_t pairs_cap = 0, leaves_cap = 0, leaf_pairs_cap = 0;
        DevBuf<float> verts;
        DevBuf<float4> normals;
        DevBuf<float> hits;
        DevBuf<uint32_t> hit_prims;
        DevBuf<uint8_t> occ;
        DevBuf<uint32_t> pool;
        DevBuf<hr::FrameSel> frame_sel;
        DevBuf<float> rad;
        DevBuf<unsigned long long> frame_rays;
    };
    LaneFields stash;
    int lane = 0;

    // this device's copy of the allocation-free entry point's scratch array
    unsigned char* noalloc = nullptr;
};

// ---------------------------------------------------------------------------
// Scratch of hr_path_trace_dev_noalloc: a fixed array in the code object's
// device memory, so that the entry point needs no allocation.  Each device
// holds its own copy for the life of the process; its address comes from
// hipGetSymbolAddress once per device (the array's name is not a device
// address in host code).  Layout:
//   [0, 4)                    persistent kernel's ray-fetch counter
//   [64, 112)                 default Phong material (no uploaded materials)
//   [256, 256 + 4 M)          closest-hit distances
//   [.. + 4 M)                primitive ids
//   [.. + 28 M)               7-plane hit records (plane stride = the call's n)
// with M = kNoAllocMaxRays.  Every call on a device shares its copy.
// ---------------------------------------------------------------------------
namespace {
constexpr uint32_t kNoAllocMaxRays = 1u << 21;
constexpr size_t kNoAllocHead = 256;
constexpr size_t kNoAllocBytes = kNoAllocHead + 36 * (size_t)kNoAllocMaxRays;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kNoAllocDevices] = {};

// the scratch address on `device`, which must be current; looked up (and the
// default material written) on the device's first call
hr_status noalloc_scratch(int device, unsigned char** out)
{
    *out = nullptr;
    if (device < 0 || device >= kNoAllocDevices) return fail(HR_ERR_ARG, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(HR_ERR_DEVICE, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        const float m[12] = { 0.2f, 0.2f, 0.2f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f };
        const hipError_t e2 = hipMemcpy((unsigned char*)p + 64, m, sizeof m, hipMemcpyHostToDevice);
        if (e2 != hipSuccess) return fail(HR_ERR_DEVICE, hipGetErrorString(e2));
        g_noalloc_ptr[device] = (unsigned char*)p;
    }
    *out = g_noalloc_ptr[device];
    return HR_OK;
}
}  // namespace

namespace {

hr_status set_device(hr_ctx* c)
{
    HR_HIP(hipSetDevice(c->device));
    return HR_OK;
}

hr_status settle_frames(hr_ctx* c);

// every entry point but hr_submit_frame_graph: the context is valid, its
// device current, and the captured frames in flight completed
hr_status check_ctx(hr_ctx* c)
{
    if (!c) return fail(HR_ERR_ARG, "null context");
    HR_TRY(set_device(c));
    return settle_frames(c);
}

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

hr_status upload(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return HR_OK;
    HR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return HR_OK;
}

hr_status validate_indices(const uint32_t* idx, uint32_t nt, uint32_t nv)
{
    for (size_t i = 0; i < 3 * (size_t)nt; ++i)
        if (idx[i] >= nv)
            return fail(HR_ERR_ARG, "vertex index " + std::to_string(idx[i]) + " at position " +
                                        std::to_string(i) + " is out of range (nv = " +
                                        std::to_string(nv) + ")");
    return HR_OK;
}

// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------

hr_status reserve_build(hr_ctx* c)
{
    const uint32_t nt = c->nt;
    const bool hier = c->kind == HR_GRID_HIERARCHY;
    const bool two = c->kind == HR_GRID_TWO_LEVEL || hier;
    const uint32_t pair_max = std::max(c->pairs_cap, two ? c->leaf_pairs_cap : 0u);
    HR_TRY(c->rec.reserve((9 * (size_t)std::max(nt, 1u) + 3) / 4));   // 9 floats per triangle
    const uint32_t waves = cdiv(std::max<uint64_t>(std::max(nt, c->pairs_cap), 1), 64);
    const uint32_t waves_cap = (waves + 3) / 4 * 4 + 4;
    HR_TRY(c->wave_counts.reserve(waves_cap));
    HR_TRY(c->wave_offsets.reserve(waves_cap));
    HR_TRY(c->keys.reserve(pair_max));
    HR_TRY(c->vals.reserve(pair_max));
    HR_TRY(c->keys_sorted.reserve(pair_max));
    HR_TRY(c->prims.reserve(two ? c->leaf_pairs_cap : c->pairs_cap));
    HR_TRY(c->cells.reserve(c->cells_cap));
    if (two) {
        HR_TRY(c->top_vals_sorted.reserve(c->pairs_cap));
        if (!hier) HR_TRY(c->leaf_counts.reserve(c->cells_cap));
        HR_TRY(c->leaves.reserve(c->leaves_cap));
    }
    if (hier) {
        HR_TRY(c->obj_gp.reserve(c->nobj));
        HR_TRY(c->obj_bounds_enc.reserve(6 * (size_t)c->nobj));
        HR_TRY(c->obj_part.reserve(7 * (size_t)hr::kObjBoundsBlocks));
        HR_TRY(c->obj_cell_counts.reserve(c->nobj));
        HR_TRY(c->obj_cell_base.reserve(c->nobj));
        HR_TRY(c->inst_counts.reserve(c->ninst));
        HR_TRY(c->inst_offsets.reserve(c->ninst));
        HR_TRY(c->inst_grid.reserve(c->ninst));
    }
    const size_t scan_n = std::max<size_t>({ (size_t)waves_cap, (size_t)c->cells_cap, (size_t)c->nobj,
                                             (size_t)c->ninst });
    {
        // the single-pass scan's tile flags must start cleared
        const size_t before = c->scan_chain.cap;
        HR_TRY(c->scan_chain.reserve(hr::scan_chain_words((uint32_t)std::min<size_t>(scan_n, 0xFFFFFFFFu))));
        if (c->scan_chain.cap != before) hr::fill_u32(c->scan_chain.p, 0u, c->scan_chain.cap, nullptr, 1, c->stream);
    }
    HR_TRY(c->sort_keys_tmp.reserve(pair_max));
    HR_TRY(c->sort_vals_tmp.reserve(pair_max));
    HR_TRY(c->sort_hist.reserve(hr::sort_hist_words(pair_max)));
    {
        // the sort passes' tile flags must start cleared
        const size_t before = c->sort_chain.cap;
        HR_TRY(c->sort_chain.reserve(hr::sort_chain_words(pair_max)));
        if (c->sort_chain.cap != before) hr::fill_u32(c->sort_chain.p, 0u, c->sort_chain.cap, nullptr, 1, c->stream);
    }
    return HR_OK;
}

hr::PersistOpts persist_opts(const hr_ctx* c)
{
    hr::PersistOpts o;
    o.fuse = (uint32_t)c->tuning.fuse_units;
    o.batch = (uint32_t)c->tuning.test_batch;
    o.steps = (uint32_t)c->tuning.advance_steps;
    return o;
}

// the path tracer's kernel: wave-cooperative with test_batch (its grid steps
// per batch), else one path per lane
uint32_t pt_wave_steps(const hr_ctx* c)
{
    return c->tuning.test_batch ? (uint32_t)c->tuning.advance_steps : 0u;
}

hr::BuildBuffers build_buffers(hr_ctx* c)
{
    hr::BuildBuffers b;
    b.params = c->params.p;
    b.status = c->status.p;
    b.rec = c->rec.p;
    b.wave_counts = c->wave_counts.p;
    b.wave_offsets = c->wave_offsets.p;
    b.keys = c->keys.p;
    b.vals = c->vals.p;
    b.keys_sorted = c->keys_sorted.p;
    b.top_vals_sorted = c->top_vals_sorted.p;
    b.prims = c->prims.p;
    b.cells = c->cells.p;
    b.leaves = c->leaves.p;
    b.leaf_counts = c->leaf_counts.p;
    b.cells_cap = c->cells_cap;
    b.pairs_cap = c->pairs_cap;
    b.leaves_cap = c->leaves_cap;
    b.leaf_pairs_cap = c->leaf_pairs_cap;
    b.waves_cap = (uint32_t)c->wave_counts.cap;
    b.scan.chain = c->scan_chain.p;
    b.scan.chain_cap = c->scan_chain.cap;
    b.sort.keys_tmp = c->sort_keys_tmp.p;
    b.sort.vals_tmp = c->sort_vals_tmp.p;
    b.sort.hist = c->sort_hist.p;
    b.sort.chain = c->sort_chain.p;
    return b;
}

// interpolation, bounds and the whole grid build; with fs the key frames and
// coefficient are read on the device (captured frames)
void enqueue_build(hr_ctx* c, const hr::FrameSel* fs = nullptr)
{
    hr::launch_reset(c->status.p, c->stream);
    if (fs)
        hr::launch_lerp_bounds(nullptr, nullptr, 0.f, c->nv, c->verts.p, c->status.p, c->bounds_part.p, c->stream, fs,
                               c->keyframes.p);
    else
        hr::launch_lerp_bounds(c->src1, c->src2, c->coeff, c->nv, c->verts.p, c->status.p, c->bounds_part.p,
                               c->stream);
    if (c->kind == HR_GRID_HIERARCHY) {
        hr::HierArgs a;
        a.idx = c->idx.p;
        a.vx = c->verts.p;
        a.vy = c->verts.p + c->nv;
        a.vz = c->verts.p + 2 * (size_t)c->nv;
        // object bounds come from the first key frame of the pair (RTEngine.cu:270-289)
        a.kx = c->src1;
        a.ky = c->src1 + c->nv;
        a.kz = c->src1 + 2 * (size_t)c->nv;
        a.nt = c->obj_tris;
        a.nobj = c->nobj;
        a.tri_obj = c->tri_obj.p;
        a.obj_first = c->obj_first.p;
        a.ninst = c->ninst;
        a.inst = c->inst.p;
        a.lambda_top = c->lambda_top;
        a.lambda_leaf = c->lambda_leaf;
        a.bounds_blocks = (uint32_t)c->tuning.hier_bounds_blocks;
        hr::HierBuffers b;
        b.status = c->status.p;
        b.top = c->params.p;
        b.obj_gp = c->obj_gp.p;
        b.obj_bounds_enc = c->obj_bounds_enc.p;
        b.obj_part = c->obj_part.p;
        b.obj_cell_counts = c->obj_cell_counts.p;
        b.obj_cell_base = c->obj_cell_base.p;
        b.inst_grid = c->inst_grid.p;
        b.rec = c->rec.p;
        b.wave_counts = c->wave_counts.p;
        b.wave_offsets = c->wave_offsets.p;
        b.waves_cap = (uint32_t)c->wave_counts.cap;
        b.keys = c->keys.p;
        b.vals = c->vals.p;
        b.keys_sorted = c->keys_sorted.p;
        b.prims = c->prims.p;
        b.obj_cells = c->leaves.p;
        b.inst_counts = c->inst_counts.p;
        b.inst_offsets = c->inst_offsets.p;
        b.top_cells = c->cells.p;
        b.inst_refs = c->top_vals_sorted.p;
        b.cells_cap = c->cells_cap;
        b.pairs_cap = c->pairs_cap;
        b.leaves_cap = c->leaves_cap;
        b.leaf_pairs_cap = c->leaf_pairs_cap;
        b.scan.chain = c->scan_chain.p;
        b.scan.chain_cap = c->scan_chain.cap;
        b.sort.keys_tmp = c->sort_keys_tmp.p;
        b.sort.vals_tmp = c->sort_vals_tmp.p;
        b.sort.hist = c->sort_hist.p;
        b.sort.chain = c->sort_chain.p;
        hr::run_build_hier(a, b, c->stream);
        return;
    }
    hr::BuildArgs a;
    a.kind = c->kind;
    a.lambda_top = c->lambda_top;
    a.lambda_leaf = c->lambda_leaf;
    a.idx = c->idx.p;
    a.vx = c->verts.p;
    a.vy = c->verts.p + c->nv;
    a.vz = c->verts.p + 2 * (size_t)c->nv;
    a.nt = c->nt;
    hr::BuildBuffers b = build_buffers(c);
    hr::run_build(a, b, c->stream);
}

uint32_t grow(uint64_t cur, uint64_t need)
{
    uint64_t v = std::max<uint64_t>(need + need / 4 + 1024, cur + cur / 2);
    return (uint32_t)std::min<uint64_t>(v, 0xFFFFFFF0ull);
}

// grid description after a successful build (c->h_params holds the parameters)
void set_info(hr_ctx* c, const hr::BuildStatus& st)
{
    const bool two = c->kind == HR_GRID_TWO_LEVEL || c->kind == HR_GRID_HIERARCHY;
    hr_grid_info& in = c->info;
    in.kind = c->kind;
    for (int i = 0; i < 3; ++i) {
        in.bmin[i] = c->h_params.bmin[i];
        in.bmax[i] = c->h_params.bmax[i];
        in.res[i] = c->h_params.res[i];
        in.cell_size[i] = c->h_params.cs[i];
        in.cell_size_rcp[i] = c->h_params.csr[i];
    }
    in.num_cells = st.num_cells;
    in.num_leaves = two ? st.num_leaves : 0u;
    in.num_refs = two ? st.num_leaf_pairs : st.num_pairs;
    in.num_top_pairs = two ? st.num_pairs : 0u;
    in.num_triangles = c->nt;
    const bool hier = c->kind == HR_GRID_HIERARCHY;
    in.num_objects = hier ? c->nobj : 0u;
    in.num_instances = hier ? c->ninst : 0u;
    c->built = true;
}

void drop_graphs(hr_ctx* c)
{
    for (int s = 0; s < 2; ++s) {
        if (c->g_build[s]) (void)hipGraphExecDestroy(c->g_build[s]);
        c->g_build[s] = nullptr;
        c->graph_sig[s].clear();
    }
}

// exchanges the current lane's per-frame buffers and stream with the other's
void swap_lane(hr_ctx* c)
{
    using std::swap;
    hr_ctx::LaneFields& o = c->stash;
    swap(c->stream, o.stream);
    swap(c->params, o.params);
    swap(c->status, o.status);
    swap(c->bounds_part, o.bounds_part);
    swap(c->rec, o.rec);
    swap(c->wave_counts, o.wave_counts);
    swap(c->wave_offsets, o.wave_offsets);
    swap(c->keys, o.keys);
    swap(c->vals, o.vals);
    swap(c->keys_sorted, o.keys_sorted);
    swap(c->top_vals_sorted, o.top_vals_sorted);
    swap(c->prims, o.prims);
    swap(c->leaf_counts, o.leaf_counts);
    swap(c->cells, o.cells);
    swap(c->leaves, o.leaves);
    swap(c->scan_chain, o.scan_chain);
    swap(c->sort_chain, o.sort_chain);
    swap(c->sort_keys_tmp, o.sort_keys_tmp);
    swap(c->sort_vals_tmp, o.sort_vals_tmp);
    swap(c->sort_hist, o.sort_hist);
    swap(c->cells_cap, o.cells_cap);
    swap(c->pairs_cap, o.pairs_cap);
    swap(c->leaves_cap, o.leaves_cap);
    swap(c->leaf_pairs_cap, o.leaf_pairs_cap);
    swap(c->verts, o.verts);
    swap(c->normals, o.normals);
    swap(c->hits, o.hits);
    swap(c->hit_prims, o.hit_prims);
    swap(c->occ, o.occ);
    swap(c->pool, o.pool);
    swap(c->frame_sel, o.frame_sel);
    swap(c->rad, o.rad);
    swap(c->frame_rays, o.frame_rays);
    c->lane ^= 1;
}

hr_status use_lane(hr_ctx* c, int lane)
{
    if (c->lane == lane) return HR_OK;
    swap_lane(c);
    if (!c->stream) {
        // first use of the second lane: its stream and fixed-size buffers
        HR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HR_TRY(c->params.reserve(1));
        HR_TRY(c->status.reserve(1));
        HR_TRY(c->bounds_part.reserve(6 * (size_t)hr::kBoundsParts));
        HR_TRY(c->pool.reserve(1));
        HR_TRY(c->frame_sel.reserve(1));
    }
    return HR_OK;
}

void release_lane(hr_ctx::LaneFields& o)
{
    o.params.release();
    o.status.release();
    o.bounds_part.release();
    o.rec.release();
    DevBuf<uint32_t>* ubufs[] = { &o.wave_counts, &o.wave_offsets, &o.keys, &o.vals, &o.keys_sorted,
                                  &o.top_vals_sorted, &o.prims, &o.leaf_counts, &o.scan_chain, &o.sort_chain,
                                  &o.sort_keys_tmp, &o.sort_vals_tmp, &o.sort_hist, &o.hit_prims, &o.pool };
    for (auto* b : ubufs) b->release();
    o.cells.release();
    o.leaves.release();
    o.verts.release();
    o.normals.release();
    o.hits.release();
    o.occ.release();
    o.frame_sel.release();
    o.rad.release();
    o.frame_rays.release();
    o.cells_cap = 0;
    o.pairs_cap = o.leaves_cap = o.leaf_pairs_cap = 0;
    if (o.stream) (void)hipStreamDestroy(o.stream);
    o.stream = nullptr;
}

hr_status do_build(hr_ctx* c)
{
    if (!c->src1 || c->nv == 0) return fail(HR_ERR_STATE, "no geometry uploaded");
    const bool hier = c->kind == HR_GRID_HIERARCHY;
    const bool two = c->kind == HR_GRID_TWO_LEVEL || hier;
    if (hier && c->ninst == 0) return fail(HR_ERR_STATE, "the instanced hierarchy needs hr_set_instances");
    if (hier && c->obj_tris > c->nt) return fail(HR_ERR_STATE, "the objects cover more triangles than the mesh has");
    // resolution product <= lambda * N (up to rounding) or the 32^3 fallback
    const uint32_t top_items = hier ? c->ninst : c->nt;
    const uint64_t cells_bound =
        std::max<uint64_t>(32768, (uint64_t)((double)c->lambda_top * (double)top_items * 1.001) + 64);
    if (cells_bound > 0xFFFFFFF0ull) return fail(HR_ERR_CAPACITY, "grid density too high");
    if (c->cells_cap < cells_bound) c->cells_cap = cells_bound;
    if (c->pairs_cap == 0) c->pairs_cap = std::max<uint32_t>(4096, 4 * top_items);
    if (two) {
        if (c->leaves_cap == 0) c->leaves_cap = std::max<uint32_t>(65536, 8 * c->nt);
        if (c->leaf_pairs_cap == 0) c->leaf_pairs_cap = std::max<uint32_t>(65536, 8 * c->nt);
    }
    c->built = false;
    for (int attempt = 0; attempt < 16; ++attempt) {
        HR_TRY(reserve_build(c));
        HR_HIP(hipEventRecord(c->ev_b0, c->stream));
        enqueue_build(c);
        HR_HIP(hipEventRecord(c->ev_b1, c->stream));
        HR_HIP(hipMemcpyAsync(c->h_status, c->status.p, sizeof(hr::BuildStatus), hipMemcpyDeviceToHost,
                              c->stream));
        HR_HIP(hipStreamSynchronize(c->stream));
        HR_HIP(hipGetLastError());
        const hr::BuildStatus& st = *c->h_status;
        if (st.overflow == 0u) {
            c->build_timed = true;
            HR_HIP(hipMemcpy(&c->h_params, c->params.p, sizeof(hr::GridParams), hipMemcpyDeviceToHost));
            set_info(c, st);
            return HR_OK;
        }
        if (std::getenv("HR_DEBUG_BUILD"))
            std::fprintf(stderr, "hr_build: attempt %d overflow 0x%x cells %u/%llu pairs %u/%u leaves %u/%u leaf pairs %u/%u\n",
                         attempt, st.overflow, st.num_cells, (unsigned long long)c->cells_cap, st.num_pairs,
                         c->pairs_cap, st.num_leaves, c->leaves_cap, st.num_leaf_pairs, c->leaf_pairs_cap);
        // the counts are valid up to the first stage that overflowed; the
        // later stages ran on nothing (or on stale data), so only the first
        // overflowed buffer grows per attempt
        if (st.overflow & hr::kOverflowCells) {
            if (st.num_cells == 0xFFFFFFFFu) return fail(HR_ERR_CAPACITY, "grid resolution too large");
            c->cells_cap = grow(c->cells_cap, st.num_cells);
        } else if (st.overflow & hr::kOverflowPairs) {
            c->pairs_cap = grow(c->pairs_cap, st.num_pairs);
        } else if (st.overflow & hr::kOverflowLeaves) {
            c->leaves_cap = grow(c->leaves_cap, st.num_leaves);
        } else if (st.overflow & hr::kOverflowLeafPairs) {
            c->leaf_pairs_cap = grow(c->leaf_pairs_cap, st.num_leaf_pairs);
        }
    }
    return fail(HR_ERR_CAPACITY, "grid build did not fit after repeated growth");
}

hr::HierScene hier_scene(hr_ctx* c)
{
    hr::HierScene H;
    H.top = c->params.p;
    H.top_cells = c->cells.p;
    H.inst_refs = c->top_vals_sorted.p;
    H.inst = c->inst.p;
    H.inst_grid = c->inst_grid.p;
    H.obj_gp = c->obj_gp.p;
    H.obj_cell_base = c->obj_cell_base.p;
    H.obj_cells = c->leaves.p;
    H.prims = c->prims.p;
    H.rec = c->rec.p;
    H.nt = c->nt;
    return H;
}

// entry points that the instanced hierarchy does not serve: the reference's
// hierarchy build renders ambient occlusion only (RTEngine.cu:400-404)
hr_status no_hierarchy(hr_ctx* c, const char* what)
{
    if (c->kind == HR_GRID_HIERARCHY)
        return fail(HR_ERR_STATE, std::string(what) + " is not available on the instanced hierarchy");
    return HR_OK;
}

hr::TraceScene trace_scene(hr_ctx* c)
{
    hr::TraceScene S;
    S.gp = c->params.p;
    S.cells = c->cells.p;
    S.leaves = c->kind == HR_GRID_TWO_LEVEL ? c->leaves.p : nullptr;
    S.prims = c->prims.p;
    S.rec = c->rec.p;
    S.nt = c->nt;
    S.two_level = c->kind == HR_GRID_TWO_LEVEL ? 1 : 0;
    return S;
}

// Camera::init (src/RT/Primitive/Camera.h:42-61), evaluated on the host
void camera_init(hr::CameraDev& cam, const float pos[3], const float dir[3], const float up[3], float fov,
                 int rx, int ry)
{
    using namespace hr;
    const float aspect = (float)rx / (float)ry;
    const f3 fwd = normalize(mk(dir[0], dir[1], dir[2]));
    const f3 right = normalize(cross(fwd, mk(up[0], up[1], up[2])));
    const f3 upa = normalize(cross(fwd, right));
    const float ang = fov * (float)kPi / 180.f;
    const f3 row = ((2.f * right) * tanf(ang / 2.f)) * aspect;
    const f3 col = (2.f * upa) * tanf(ang / 2.f);
    const f3 sx = row / (float)rx;
    const f3 sy = col / (float)ry;
    const f3 tl = fwd - row / 2.f - col / 2.f;
    for (int i = 0; i < 3; ++i) cam.center[i] = pos[i];
    cam.top_left[0] = tl.x; cam.top_left[1] = tl.y; cam.top_left[2] = tl.z;
    cam.step_x[0] = sx.x; cam.step_x[1] = sx.y; cam.step_x[2] = sx.z;
    cam.step_y[0] = sy.x; cam.step_y[1] = sy.y; cam.step_y[2] = sy.z;
    cam.res_x = rx;
    cam.res_y = ry;
}

hr_status ensure_default_material(hr_ctx* c)
{
    if (c->mats.p) return HR_OK;
    const float m[12] = { 0.2f, 0.2f, 0.2f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f };
    HR_TRY(c->mats.reserve(12));
    HR_TRY(upload(c->mats.p, m, sizeof m, c->stream));
    return HR_OK;
}

hr_status check_render_args(hr_ctx* c, hr_mode mode, int image_id, const hr_render_params* p)
{
    if (!c->has_camera) return fail(HR_ERR_STATE, "render before setCamera");
    if (!p) return fail(HR_ERR_ARG, "null render parameters");
    if (mode != HR_MODE_AO && mode != HR_MODE_PATH && mode != HR_MODE_SIMPLE && mode != HR_MODE_PATH_MULTI)
        return fail(HR_ERR_ARG, "unknown render mode");
    if (image_id < 0) return fail(HR_ERR_ARG, "negative image id");
    if (c->kind == HR_GRID_HIERARCHY && mode != HR_MODE_AO)
        return fail(HR_ERR_STATE, "the instanced hierarchy renders ambient occlusion only");
    if (mode == HR_MODE_AO && (p->ao_samples < 1 || p->ao_samples > 256))
        return fail(HR_ERR_ARG, "ao_samples must be in [1, 256]");
    // AO ray ids are 32-bit (pixel * samples + sample, over the image rounded
    // up to whole 8x8 tiles where rays are taken in tiled order), with room
    // for the ray pool's overshoot
    if (mode == HR_MODE_AO &&
        (uint64_t)((c->cam.res_x + 7) / 8 * 8) * (uint64_t)((c->cam.res_y + 7) / 8 * 8) *
                (uint64_t)p->ao_samples > 0xFFFFFFFFull - (1ull << 20))
        return fail(HR_ERR_ARG, "pixels * ao_samples exceeds 2^32 - 2^20 AO rays (32-bit ray ids)");
    if (mode == HR_MODE_PATH && (p->pt_depth < 1 || p->pt_depth > 64))
        return fail(HR_ERR_ARG, "pt_depth must be in [1, 64]");
    return HR_OK;
}

// buffers a render writes
hr_status prepare_render(hr_ctx* c, hr_mode mode, const hr_render_params* p)
{
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    HR_TRY(c->hits.reserve(7 * (size_t)npix));
    HR_TRY(c->hit_prims.reserve(npix));
    HR_TRY(c->fb.reserve(3 * (size_t)npix));
    if (mode == HR_MODE_AO) HR_TRY(c->occ.reserve((size_t)npix * (size_t)p->ao_samples));
    if (c->kind == HR_GRID_HIERARCHY) HR_TRY(c->hit_inst.reserve(npix));
    if (mode == HR_MODE_PATH_MULTI) {
        const size_t before = c->pm.cap;
        HR_TRY(c->pm.reserve(28 * (size_t)npix));
        HR_TRY(c->pm_prims.reserve(npix));
        if (c->pm.cap != before || c->pm_npix != npix) {
            // PathTracer::setResolution clears the carried buffers (PathTracer.h:1028-1035)
            hr::fill_u32((uint32_t*)c->pm.p, 0u, c->pm.cap, nullptr, 1, c->stream);
            c->pm_npix = npix;
        }
    }
    HR_TRY(ensure_default_material(c));
    return HR_OK;
}

hr::PathBuffers path_buffers(hr_ctx* c)
{
    const size_t n = c->pm_npix;
    float* q = c->pm.p;
    hr::PathBuffers pb;
    pb.end = q;
    pb.end_prims = c->pm_prims.p;
    pb.importance = q + 7 * n;
    pb.ls = q + 10 * n;
    pb.dir_ls = q + 13 * n;
    pb.new_radiance = q + 16 * n;
    pb.residue = q + 19 * n;
    pb.blended = q + 22 * n;
    pb.contrib = q + 25 * n;
    return pb;
}

bool uses_normals(hr_mode mode) { return mode == HR_MODE_PATH || mode == HR_MODE_PATH_MULTI; }

// current-frame normals from the normal key frames; with fs the key frames
// are chosen on the device
void enqueue_normals(hr_ctx* c, const hr::FrameSel* fs)
{
    const size_t blk = 3 * (size_t)c->nn;
    if (c->nframes_n == 1) {
        hr::launch_frame_normals(c->normal_keys.p, c->normal_keys.p, c->nn, c->normals.p, c->stream);
    } else if (fs) {
        hr::launch_frame_normals(nullptr, nullptr, c->nn, c->normals.p, c->stream, fs, c->normal_keys.p);
    } else {
        hr::launch_frame_normals(c->normal_keys.p + (size_t)c->k1 * blk, c->normal_keys.p + (size_t)c->k2 * blk,
                                 c->nn, c->normals.p, c->stream);
    }
}

void reset_primary_choice(hr_ctx* c)
{
    c->pp_measured = 0;
    c->pp_pending = -1;
    c->pp_choice = -1;
}

bool primary_decided(const hr_ctx* c) { return c->tuning.persistent_primary != 2 || c->pp_choice >= 0; }

// reads the camera-ray time of the last trial frame (its events must have
// been recorded by that frame; waits for it)
// the camera-ray kernels that persistent_primary 2 chooses from, in trial order:
// one ray per lane, persistent, one ray per lane with wave-batched tests
constexpr int kPrimaryKernels[3] = { 0, 1, 3 };
// two rounds of trials: a kernel's first launch in a process, and the first
// frame after a change, can take more than twice its steady time
constexpr int kPrimaryTrials = 6;

void resolve_primary_trial(hr_ctx* c)
{
    if (c->pp_pending < 0) return;
    float ms = 0.f;
    const bool timed =
        hipEventSynchronize(c->ev_p1) == hipSuccess && hipEventElapsedTime(&ms, c->ev_r0, c->ev_p1) == hipSuccess;
    if (!timed) (void)hipGetLastError();
    if (timed) {
        const int k = c->pp_pending % 3;
        c->pp_ms[k] = c->pp_pending < 3 ? ms : std::min(c->pp_ms[k], ms);
        c->pp_measured = c->pp_pending + 1;
    }
    c->pp_pending = -1;
    if (c->pp_measured == kPrimaryTrials) {
        int b = 0;
        for (int k = 1; k < 3; ++k)
            if (c->pp_ms[k] < c->pp_ms[b]) b = k;
        c->pp_choice = kPrimaryKernels[b];
    }
}

// the camera-ray kernel of the next frame: 0 one ray per lane, 1 persistent,
// 3 one ray per lane with wave-batched tests.  Under persistent_primary 2 the
// first six timed frames after a reset run kernels 0, 1, 3, 0, 1, 3, each timed
// by the frame's events (read by resolve_primary_trial at the start of the next
// frame), and the kernel with the least time is kept.  timed = false (a frame
// recorded into a graph) starts no trial.
int primary_kernel(hr_ctx* c, bool timed)
{
    if (c->tuning.persistent_primary != 2) return c->tuning.persistent_primary;
    if (c->pp_choice >= 0) return c->pp_choice;
    if (!timed) return 0;
    c->pp_pending = c->pp_measured;
    return kPrimaryKernels[c->pp_pending % 3];
}

// Primary rays and the integrator.  fs != nullptr: the image id is read on the
// device (captured frames); events = false while a graph is being recorded.
// phases: 1 the camera rays, 2 the secondary rays and the shading, 3 both
// (the instanced hierarchy and the path-multi mode take 3 only)
constexpr int kPhaseCamera = 1, kPhaseSecondary = 2, kPhaseAll = 3;

hr_status enqueue_render(hr_ctx* c, hr_mode mode, int image_id, const hr_render_params* p,
                         const hr::FrameSel* fs, bool events, bool lane_out = false,
                         const hipEvent_t* marks = nullptr, int phases = kPhaseAll)
{
    // events: the context's (timed uncaptured frames: after the camera rays,
    // after the secondary rays, at the end) or a captured frame's slot marks
    // (camera start, camera end, secondary start, secondary end, shading end)
    const hipEvent_t e_c0 = marks ? marks[0] : nullptr;
    const hipEvent_t e_p1 = events ? c->ev_p1 : marks ? marks[1] : nullptr;
    const hipEvent_t e_c1 = marks ? marks[2] : nullptr;
    const hipEvent_t e_p2 = events ? c->ev_p2 : marks ? marks[3] : nullptr;
    const hipEvent_t e_r1 = events ? c->ev_r1 : marks ? marks[4] : nullptr;
    // lane_out: a frame lane's render (captured), into the lane's own value
    // buffer and ray counter, which k_frame_blend later blends in and adds
    float* const fb = lane_out ? c->rad.p : c->fb.p;
    unsigned long long* const rays = lane_out ? c->frame_rays.p : c->ray_count.p;
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    hr::RenderArgs ra{};
    ra.seed_lo = c->seed_lo;
    ra.seed_hi = c->seed_hi;
    ra.jitter = mode == HR_MODE_PATH_MULTI ? 1 : p->jitter;   // PathTracer's RandomPrimaryRayGenerator
    ra.simple_sampler = mode == HR_MODE_SIMPLE ? 1 : 0;
    ra.ao_samples = p->ao_samples;
    ra.ao_dist = p->ao_dist;
    ra.pt_depth = p->pt_depth;
    ra.image = image_id;
    ra.stats = p->collect_stats ? c->work.p : nullptr;
    ra.frame = fs;
    ra.direct = lane_out ? 1 : 0;
    const hr::TraceScene S = trace_scene(c);
    hr::Shading sh;
    sh.mats = c->mats.p;
    sh.mat_idx = c->has_mat_idx ? c->mat_idx.p : nullptr;
    sh.lights = c->lights.p;
    sh.light_cdf = c->light_cdf.p;
    sh.nlights = c->nlights;
    sh.normals = nullptr;
    sh.nidx = nullptr;
    if (uses_normals(mode) && c->has_normals) {
        sh.normals = c->normals.p;
        sh.nidx = c->nidx.p;
    }
    // the last trial frame's camera-ray time, read before its events are recorded again
    if (events) resolve_primary_trial(c);
    if (events) HR_HIP(hipEventRecord(c->ev_r0, c->stream));
    c->hier_npix = c->kind == HR_GRID_HIERARCHY ? npix : 0u;
    c->ao_dump_ok = false;
    c->occ_npix = mode == HR_MODE_AO ? npix : 0u;
    if (c->kind == HR_GRID_HIERARCHY) {
        // TLGHAOIntegrator::integrate (TLGridHierarchyAOIntegrator.h:421-573)
        c->last_pk = c->tuning.persistent_primary == 1 ? 1 : 0;
        hr::launch_render_ao_hier(hier_scene(c), c->cam, ra, c->hits.p, c->hit_prims.p, c->hit_inst.p, c->occ.p, sh,
                                  fb, c->tuning.persistent_primary == 1, c->tuning.persistent_secondary != 0,
                                  c->pool.p, (uint32_t)c->tuning.hier_refill, rays, c->stream,
                                  c->tuning.test_batch != 0, e_p1, e_p2,
                                  c->tuning.camera_blocks > 0 ? hr::kHierCameraBlocks : 0u);
        HR_HIP(hipGetLastError());
        if (e_r1) HR_HIP(hipEventRecord(e_r1, c->stream));
        c->last_ao_samples = p->ao_samples;
        HR_HIP(hipGetLastError());
        return HR_OK;
    }
    if (mode == HR_MODE_PATH_MULTI && c->nlights == 0) {
        // PathTracer::integrate returns at once without lights (PathTracer.h:1049-1053)
        for (hipEvent_t e : { e_c0, e_p1, e_c1, e_p2, e_r1 })
            if (e) HR_HIP(hipEventRecord(e, c->stream));
        return HR_OK;
    }
    const bool pers = c->tuning.persistent_secondary != 0;
    const uint32_t refill = (uint32_t)c->tuning.refill;
    if (phases & kPhaseCamera) {
        const int pk = primary_kernel(c, events);
        c->last_pk = pk;
        if (e_c0) HR_HIP(hipEventRecord(e_c0, c->stream));
        if (pk == 1)
            hr::persist_primary(S, c->cam, ra, c->hits.p, c->hit_prims.p, c->pool.p, refill, rays,
                                c->stream, persist_opts(c));
        else
            hr::launch_primary(S, c->cam, ra, c->hits.p, c->hit_prims.p, rays, c->stream,
                               pk == 3 ? (uint32_t)c->tuning.advance_steps : 0u,
                               (uint32_t)c->tuning.camera_blocks, c->pool.p);
        if (e_p1) HR_HIP(hipEventRecord(e_p1, c->stream));
    }
    if (!(phases & kPhaseSecondary)) {
        HR_HIP(hipGetLastError());
        return HR_OK;
    }
    if (e_c1) HR_HIP(hipEventRecord(e_c1, c->stream));
    if (ra.stats) ra.stats = c->work.p + 3;
    if (mode == HR_MODE_AO) {
        if (pers)
            hr::persist_ao(S, ra, npix, c->hits.p, c->hit_prims.p, c->occ.p, c->pool.p, refill, rays,
                           c->stream, persist_opts(c));
        else
            hr::launch_ao_rays(S, ra, npix, c->hits.p, c->hit_prims.p, c->occ.p, rays, c->stream);
        if (e_p2) HR_HIP(hipEventRecord(e_p2, c->stream));
        hr::launch_ao_shade(S, ra, npix, c->hits.p, c->hit_prims.p, c->occ.p, sh, fb, c->stream);
        c->last_ao_samples = p->ao_samples;
        c->last_ao_ra = ra;
        c->last_ao_ra.stats = nullptr;
        c->last_ao_cam = c->cam;
        c->ao_dump_ok = !fs;
    } else if (mode == HR_MODE_PATH) {
        if (c->tuning.persistent_path)
            hr::persist_pt(S, ra, npix, c->hits.p, c->hit_prims.p, sh, fb, c->pool.p,
                           (uint32_t)c->tuning.path_refill, (uint32_t)c->tuning.path_shade, rays,
                           c->stream);
        else
            hr::launch_pt(S, ra, npix, c->hits.p, c->hit_prims.p, sh, fb, rays, c->stream,
                          pt_wave_steps(c));
        if (e_p2) HR_HIP(hipEventRecord(e_p2, c->stream));
    } else if (mode == HR_MODE_SIMPLE) {
        if (e_p2) HR_HIP(hipEventRecord(e_p2, c->stream));
        hr::launch_simple_shade(S, ra, npix, c->hits.p, c->hit_prims.p, fb, c->stream);
    } else {
        hr::launch_path_multi(S, c->cam, ra, c->hits.p, c->hit_prims.p, sh, path_buffers(c), fb,
                              rays, c->stream, e_p2);
    }
    if (e_r1) HR_HIP(hipEventRecord(e_r1, c->stream));
    HR_HIP(hipGetLastError());
    return HR_OK;
}

template <class T>
void put(std::vector<unsigned char>& v, const T& x)
{
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&x);
    v.insert(v.end(), b, b + sizeof(T));
}

// everything a recorded frame bakes into its kernel arguments
std::vector<unsigned char> graph_signature(hr_ctx* c, hr_mode mode, const hr_render_params* p)
{
    std::vector<unsigned char> v;
    put(v, (int)mode);
    put(v, p->jitter);
    put(v, p->ao_samples);
    put(v, p->ao_dist);
    put(v, p->pt_depth);
    put(v, p->collect_stats);
    put(v, c->seed_lo);
    put(v, c->seed_hi);
    put(v, c->kind);
    put(v, c->lambda_top);
    put(v, c->lambda_leaf);
    put(v, c->nt);
    put(v, c->nv);
    put(v, c->nframes);
    const void* ptrs[] = { c->bounds_part.p, c->keyframes.p, c->verts.p, c->idx.p, c->rec.p, c->wave_counts.p, c->keys.p,
                           c->vals.p, c->keys_sorted.p, c->top_vals_sorted.p, c->prims.p, c->cells.p,
                           c->leaves.p, c->leaf_counts.p, c->wave_offsets.p, c->scan_chain.p,
                           c->sort_chain.p, c->sort_keys_tmp.p,
                           c->sort_vals_tmp.p, c->sort_hist.p, c->mats.p, c->mat_idx.p, c->lights.p,
                           c->light_cdf.p, c->normal_keys.p, c->normals.p, c->nidx.p, c->hits.p,
                           c->hit_prims.p, c->fb.p, c->occ.p, c->pm.p, c->pm_prims.p, c->pool.p,
                           c->frame_sel.p, c->rad.p, c->frame_rays.p, c->status.p, c->params.p };
    for (const void* q : ptrs) put(v, q);
    put(v, c->cells_cap);
    put(v, c->pairs_cap);
    put(v, c->leaves_cap);
    put(v, c->leaf_pairs_cap);
    put(v, c->wave_counts.cap);
    put(v, c->wave_offsets.cap);
    put(v, c->pm_npix);
    put(v, c->cam);
    put(v, c->nlights);
    put(v, c->has_mat_idx);
    put(v, c->has_normals);
    put(v, c->nn);
    put(v, c->nframes_n);
    put(v, c->tuning);
    put(v, c->pp_choice);
    return v;
}

hr_status end_capture(hr_ctx* c, hipGraphExec_t* out)
{
    hipGraph_t g = nullptr;
    HR_HIP(hipStreamEndCapture(c->stream, &g));
    const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(HR_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return HR_OK;
}

// lanes for captured frames of this mode and these parameters: the
// multi-kernel path tracer carries per-image state of its own, and work
// statistics are not split per frame, so both keep to one lane
int frame_lanes(const hr_ctx* c, hr_mode mode, const hr_render_params* p)
{
    return c->tuning.frame_lanes >= 2 && mode != HR_MODE_PATH_MULTI && !p->collect_stats ? 2 : 1;
}

int slot_lane(const hr_ctx* c, int slot) { return c->graph_lanes == 2 ? slot : 0; }

// Gives the current lane buffers for the current state by a build of the
// current frame selection (no render): the capacities start from the other
// lane's, so the build seldom regrows.
hr_status prime_lane(hr_ctx* c, const hr_ctx::LaneFields& other, hr_mode mode, const hr_render_params* p)
{
    c->cells_cap = std::max(c->cells_cap, other.cells_cap);
    c->pairs_cap = std::max(c->pairs_cap, other.pairs_cap);
    c->leaves_cap = std::max(c->leaves_cap, other.leaves_cap);
    c->leaf_pairs_cap = std::max(c->leaf_pairs_cap, other.leaf_pairs_cap);
    HR_TRY(c->verts.reserve(3 * (size_t)c->nv));
    if (c->has_normals) HR_TRY(c->normals.reserve(c->nn));
    HR_TRY(do_build(c));
    return prepare_render(c, mode, p);
}

// Record, for each slot, the rebuild graph: frame selection upload from the
// slot's pinned memory, interpolation, build, normals, the frame gate, and
// read-back of the status, grid parameters and halt word into the slot.
// With two lanes slot s is recorded on lane s, whose buffers a build of the
// current frame selection sizes first, and each lane gets its value buffer.
// No memset node is recorded.
hr_status record_graphs(hr_ctx* c, hr_mode mode, const hr_render_params* p)
{
    drop_graphs(c);
    // nothing is in flight here and the halt word is clear: frames number from 0 again
    c->frame_seq = 0;
    const int lanes = frame_lanes(c, mode, p);
    c->graph_lanes = lanes;
    c->graph_ordered = lanes == 2 && c->tuning.frame_lanes == 3;
    const int home = c->lane;
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    hr_status st = HR_OK;
    for (int s = 0; s < 2 && st == HR_OK; ++s) {
        const int L = slot_lane(c, s);
        if (L != c->lane) {
            if ((st = use_lane(c, L)) != HR_OK) break;
            if ((st = prime_lane(c, c->stash, mode, p)) != HR_OK) break;
        }
        if (lanes == 2) {
            if ((st = c->rad.reserve(3 * (size_t)npix)) != HR_OK) break;
            if ((st = c->frame_rays.reserve(1)) != HR_OK) break;
        }
        hr::FrameSel* fs = c->frame_sel.p;
        hr_ctx::FramePin* pin = c->h_pin + s;
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
            st = fail(HR_ERR_DEVICE, "hipStreamBeginCapture failed");
            break;
        }
        hipError_t e = hipMemcpyAsync(fs, &pin->sel, sizeof(hr::FrameSel), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            enqueue_build(c, fs);
            if (c->has_normals) enqueue_normals(c, fs);
            hr::launch_frame_gate(c->status.p, c->halt.p, fs, lanes == 2 ? c->frame_rays.p : nullptr, c->stream);
            e = hipMemcpyAsync(&pin->status, c->status.p, sizeof(hr::BuildStatus), hipMemcpyDeviceToHost,
                               c->stream);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(&pin->params, c->params.p, sizeof(hr::GridParams), hipMemcpyDeviceToHost,
                               c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&pin->halt, c->halt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
        st = end_capture(c, &c->g_build[s]);
        if (e != hipSuccess) st = fail(HR_ERR_DEVICE, std::string("capture: ") + hipGetErrorString(e));
        if (st != HR_OK) break;
        c->graph_sig[c->lane] = graph_signature(c, mode, p);
    }
    const hr_status back = use_lane(c, home);
    if (st == HR_OK) st = back;
    if (st != HR_OK) {
        drop_graphs(c);
        return st;
    }
    c->graph_mode = mode;
    c->graph_params = *p;
    return HR_OK;
}

// ---- captured frames in flight ----

bool graphs_ready(hr_ctx* c, hr_mode mode, const hr_render_params* p)
{
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    return c->g_build[0] && c->g_build[1] && npix == c->last_npix &&
           primary_decided(c) && c->graph_lanes == frame_lanes(c, mode, p) &&
           c->graph_ordered == (c->graph_lanes == 2 && c->tuning.frame_lanes == 3) &&
           c->graph_sig[c->lane] == graph_signature(c, mode, p);
}

void select_frame(hr_ctx* c, const hr_ctx::Submitted& f)
{
    c->src1 = c->keyframes.p + (size_t)f.k1 * 3 * c->nv;
    c->src2 = c->keyframes.p + (size_t)f.k2 * 3 * c->nv;
    c->coeff = f.coeff;
    c->k1 = f.k1;
    c->k2 = f.k2;
    c->normals_dirty = true;
}

// one frame without graphs on the current lane (the build grows its buffers
// as needed); waits for the build, not for the render
hr_status frame_uncaptured(hr_ctx* c, const hr_ctx::Submitted& f, hr_mode mode, const hr_render_params* p)
{
    select_frame(c, f);
    c->built = false;
    HR_TRY(do_build(c));
    HR_TRY(prepare_render(c, mode, p));
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    if (f.image == 0 || npix != c->last_npix) {
        hr::fill_u32((uint32_t*)c->fb.p, 0u, 3 * (size_t)npix, nullptr, 1, c->stream);
        c->last_npix = npix;
    }
    if (uses_normals(mode) && c->has_normals) {
        enqueue_normals(c, nullptr);
        c->normals_dirty = false;
    }
    HR_TRY(enqueue_render(c, mode, f.image, p, nullptr, true));
    c->render_timed = true;
    c->graph_timed = false;
    return HR_OK;
}

hr_status sync_lanes(hr_ctx* c)
{
    HR_HIP(hipStreamSynchronize(c->stream));
    if (c->stash.stream) HR_HIP(hipStreamSynchronize(c->stash.stream));
    return HR_OK;
}

// A replayed rebuild outgrew its buffers: that frame and every frame queued
// behind it were skipped on the device (their render kernels returned at
// entry, or their values were not blended in).  Render them again in order
// without graphs on the current lane, which grows its buffers, and record
// the graphs anew.
hr_status recover_frames(hr_ctx* c)
{
    HR_TRY(sync_lanes(c));
    hr_ctx::Submitted redo[2];
    const int n = c->npending;
    for (int i = 0; i < n; ++i) redo[i] = c->pending[i];
    c->npending = 0;
    c->deferred = false;
    drop_graphs(c);
    hr::fill_u32(c->halt.p, 0xFFFFFFFFu, 1, nullptr, 1, c->stream);
    const hr_mode mode = c->graph_mode;
    const hr_render_params p = c->graph_params;
    for (int i = 0; i < n; ++i) HR_TRY(frame_uncaptured(c, redo[i], mode, &p));
    HR_HIP(hipStreamSynchronize(c->stream));
    resolve_primary_trial(c);
    return record_graphs(c, mode, &p);
}

// waits for the older frame in flight and takes its results (grid info,
// timing), or recovers it and the frames behind it
hr_status finish_deferred(hr_ctx* c, hipEvent_t after);

hr_status complete_oldest(hr_ctx* c)
{
    if (c->deferred && c->npending == 1) HR_TRY(finish_deferred(c, nullptr));
    const hr_ctx::Submitted f = c->pending[0];
    HR_HIP(hipEventSynchronize(c->ev_g[f.slot][3]));
    const hr_ctx::FramePin& pin = c->h_pin[f.slot];
    if (pin.halt <= f.seq) return recover_frames(c);
    c->h_params = pin.params;
    set_info(c, pin.status);
    const hipEvent_t* g = c->ev_g[f.slot];
    const hipEvent_t* k = c->ev_k[f.slot];
    float* outs[5] = { &c->graph_build_ms, &c->graph_render_ms, &c->graph_primary_ms, &c->graph_secondary_ms,
                       &c->graph_shade_ms };
    const hipEvent_t from[5] = { g[0], k[0], k[0], k[2], k[3] };
    const hipEvent_t to[5] = { g[1], k[4], k[1], k[3], k[4] };
    // a frame with an interval that cannot be read stays out of the sums
    bool all = true;
    for (int i = 0; i < 5; ++i) {
        float ms = 0.f;
        const hipError_t te = hipEventElapsedTime(&ms, from[i], to[i]);
        if (te != hipSuccess) {
            if (std::getenv("HR_DEBUG_TIMING"))
                std::fprintf(stderr, "replay timing: seq %u slot %d interval %d: %s\n", f.seq, f.slot, i,
                             hipGetErrorString(te));
            (void)hipGetLastError();   // the frame's results stand; leave no error behind
            ms = 0.f;
            all = false;
        }
        *outs[i] = ms;
    }
    if (all) {
        for (int i = 0; i < 5; ++i) c->replay_ms[i] += *outs[i];
        ++c->replays;
    }
    c->build_timed = c->render_timed = c->graph_timed = true;
    c->built = true;
    c->normals_dirty = false;
    if (c->graph_mode == HR_MODE_AO) c->last_ao_samples = c->graph_params.ao_samples;
    c->pending[0] = c->pending[1];
    --c->npending;
    return HR_OK;
}

// completes every frame in flight (called by every entry point but the
// frame submission itself); the current lane then holds the last frame's
// grid and hit buffers
hr_status settle_frames(hr_ctx* c)
{
    if (c->deferred) HR_TRY(finish_deferred(c, nullptr));
    while (c->npending > 0) HR_TRY(complete_oldest(c));
    return HR_OK;
}

// Ordered lanes: the secondary rays, shading and blend of the newest frame in
// flight, on its own lane, after `after` (the next frame's rebuild, so that
// the rebuild runs beside this frame's camera rays and not beside its
// secondary rays), and its blend after the previous frame's
hr_status finish_deferred(hr_ctx* c, hipEvent_t after)
{
    const hr_ctx::Submitted d = c->pending[c->npending - 1];
    const int back = c->lane;
    HR_TRY(use_lane(c, slot_lane(c, d.slot)));
    if (after) HR_HIP(hipStreamWaitEvent(c->stream, after, 0));
    HR_TRY(enqueue_render(c, c->graph_mode, 0, &c->graph_params, c->frame_sel.p, false, true, c->ev_k[d.slot],
                          kPhaseSecondary));
    HR_HIP(hipEventRecord(c->ev_g[d.slot][2], c->stream));
    HR_HIP(hipStreamWaitEvent(c->stream, c->ev_blend, 0));
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    hr::launch_frame_blend(c->frame_sel.p, c->halt.p, c->rad.p, c->fb.p, npix, c->frame_rays.p, c->ray_count.p,
                           c->stream);
    HR_HIP(hipGetLastError());
    HR_HIP(hipEventRecord(c->ev_blend, c->stream));
    HR_HIP(hipEventRecord(c->ev_g[d.slot][3], c->stream));
    c->deferred = false;
    return use_lane(c, back);
}

// Queues one captured frame.  With the graphs recorded for this state the
// frame is replayed into the free slot (after the frame two submissions back
// has completed) on that slot's lane, and the call returns without waiting
// for it; otherwise the frames in flight complete, this frame is rendered
// without graphs and the graphs are recorded.
hr_status submit_frame(hr_ctx* c, const hr_ctx::Submitted& f, hr_mode mode, const hr_render_params* p)
{
    if (c->npending == 2) HR_TRY(complete_oldest(c));
    if (!graphs_ready(c, mode, p)) {
        HR_TRY(settle_frames(c));
        if (frame_lanes(c, mode, p) == 1 && c->lane != 0) {
            // one lane runs on lane 0; the other lane's stream may still hold
            // uncaptured work that this frame must follow
            HR_TRY(sync_lanes(c));
            HR_TRY(use_lane(c, 0));
        }
        HR_TRY(frame_uncaptured(c, f, mode, p));
        HR_HIP(hipStreamSynchronize(c->stream));
        resolve_primary_trial(c);   // before the capture, which records no trial
        return record_graphs(c, mode, p);
    }
    const int s = c->next_slot;
    const int L = slot_lane(c, s);
    if (L != c->lane) {
        // everything queued so far on the lane left behind (its last blend
        // above all) comes before this frame's blend
        HR_HIP(hipEventRecord(c->ev_switch, c->stream));
        HR_TRY(use_lane(c, L));
    }
    select_frame(c, f);
    hr_ctx::Submitted q = f;
    q.slot = s;
    q.seq = c->frame_seq++;
    c->h_pin[s].sel = hr::FrameSel{ f.k1, f.k2, f.coeff, f.image, 0u, q.seq };
    HR_HIP(hipEventRecord(c->ev_g[s][0], c->stream));
    HR_HIP(hipGraphLaunch(c->g_build[s], c->stream));
    HR_HIP(hipEventRecord(c->ev_g[s][1], c->stream));
    if (c->graph_ordered) {
        // this rebuild runs beside the previous frame's camera rays; that
        // frame's secondary rays follow it, and this frame's camera rays follow
        // those, so the secondary rays have the GPU to themselves
        HR_HIP(hipEventRecord(c->ev_built[L], c->stream));
        if (c->deferred) {
            const hr_ctx::Submitted prev = c->pending[c->npending - 1];
            HR_TRY(finish_deferred(c, c->ev_built[L]));
            HR_HIP(hipStreamWaitEvent(c->stream, c->ev_k[prev.slot][3], 0));
        }
        HR_TRY(enqueue_render(c, mode, 0, p, c->frame_sel.p, false, true, c->ev_k[s], kPhaseCamera));
        c->pending[c->npending++] = q;
        c->deferred = true;
        c->next_slot = s ^ 1;
        return HR_OK;
    }
    // the render: launched directly (its few launches cost the host little, and
    // frames run ahead of the host), timed by ordinary events
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    if (c->graph_lanes == 1) hr::launch_clear_first(c->fb.p, 3 * (size_t)npix, c->frame_sel.p, c->stream);
    HR_TRY(enqueue_render(c, mode, 0, p, c->frame_sel.p, false, c->graph_lanes == 2, c->ev_k[s]));
    HR_HIP(hipEventRecord(c->ev_g[s][2], c->stream));
    if (c->graph_lanes == 2) {
        HR_HIP(hipStreamWaitEvent(c->stream, c->ev_switch, 0));
        hr::launch_frame_blend(c->frame_sel.p, c->halt.p, c->rad.p, c->fb.p, npix, c->frame_rays.p,
                               c->ray_count.p, c->stream);
        HR_HIP(hipGetLastError());
    }
    HR_HIP(hipEventRecord(c->ev_g[s][3], c->stream));
    c->pending[c->npending++] = q;
    c->next_slot = s ^ 1;
    return HR_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// internal interface (hr_ctx_internal.h)
// ---------------------------------------------------------------------------

hr::TraceScene hr::ctx_trace_scene(hr_ctx* c) { return trace_scene(c); }

size_t hr::ctx_device_bytes(hr_ctx* c)
{
    size_t b = 0;
    b += c->key_a.bytes() + c->key_b.bytes() + c->keyframes.bytes() + c->verts.bytes() + c->idx.bytes();
    b += c->normal_keys.bytes() + c->nidx.bytes() + c->normals.bytes() + c->mats.bytes() + c->mat_idx.bytes();
    b += c->rec.bytes() + c->wave_counts.bytes() + c->wave_offsets.bytes() + c->scan_chain.bytes();
    b += c->keys.bytes() + c->vals.bytes() + c->keys_sorted.bytes();
    b += c->top_vals_sorted.bytes() + c->prims.bytes() + c->leaf_counts.bytes() + c->cells.bytes();
    b += c->leaves.bytes() + c->sort_chain.bytes() + c->sort_keys_tmp.bytes() + c->sort_vals_tmp.bytes();
    b += c->sort_hist.bytes() + c->lights.bytes() + c->light_cdf.bytes() + c->hits.bytes() + c->fb.bytes();
    b += c->hit_prims.bytes() + c->occ.bytes() + c->ray_org.bytes() + c->ray_dir.bytes() + c->ray_t.bytes();
    b += c->ray_prim.bytes() + c->ray_occ.bytes() + c->ray_hits.bytes() + c->pm.bytes() + c->pm_prims.bytes();
    b += c->ray_wire.bytes() + c->tri_obj.bytes() + c->obj_first.bytes() + c->obj_bounds_enc.bytes();
    b += c->obj_cell_counts.bytes() + c->obj_cell_base.bytes() + c->inst_counts.bytes() + c->inst_offsets.bytes();
    b += c->inst.bytes() + c->inst_grid.bytes() + c->obj_gp.bytes() + c->hit_inst.bytes() + c->ray_inst.bytes() + c->obj_part.bytes();
    b += c->bounds_part.bytes() + c->frame_sel.bytes() + c->rad.bytes() + c->frame_rays.bytes();
    // the other frame lane's buffers
    const hr_ctx::LaneFields& o = c->stash;
    b += o.params.bytes() + o.status.bytes() + o.bounds_part.bytes() + o.rec.bytes() + o.wave_counts.bytes();
    b += o.wave_offsets.bytes() + o.keys.bytes() + o.vals.bytes() + o.keys_sorted.bytes() + o.top_vals_sorted.bytes();
    b += o.prims.bytes() + o.leaf_counts.bytes() + o.cells.bytes() + o.leaves.bytes() + o.scan_chain.bytes();
    b += o.sort_chain.bytes() + o.sort_keys_tmp.bytes() + o.sort_vals_tmp.bytes() + o.sort_hist.bytes();
    b += o.verts.bytes() + o.normals.bytes() + o.hits.bytes() + o.hit_prims.bytes() + o.occ.bytes() + o.pool.bytes();
    b += o.frame_sel.bytes() + o.rad.bytes() + o.frame_rays.bytes();
    return b;
}

namespace {
// paths from a hit buffer; default_mat: where the default material lies when
// the context has none (nullptr: allocate it in the context)
hr_status path_trace_hits(hr_ctx* c, const float* hits, const uint32_t* prims, uint32_t n, int depth, float* rgb,
                          hipStream_t s, const float* default_mat, uint32_t* pool = nullptr)
{
    if (!c->built) return fail(HR_ERR_STATE, "path trace before build");
    HR_TRY(no_hierarchy(c, "path tracing"));
    if (depth < 1 || depth > 64) return fail(HR_ERR_ARG, "depth must be in [1, 64]");
    if (n == 0) return HR_OK;
    if (!default_mat || c->mats.p) HR_TRY(ensure_default_material(c));
    if (c->has_normals && c->normals_dirty) {
        hipStream_t keep = c->stream;
        c->stream = s;
        enqueue_normals(c, nullptr);
        c->stream = keep;
        c->normals_dirty = false;
    }
    hr::RenderArgs ra{};
    ra.seed_lo = c->seed_lo;
    ra.seed_hi = c->seed_hi;
    ra.jitter = 0;
    ra.ao_samples = 1;
    ra.ao_dist = 0.f;
    ra.pt_depth = depth;
    ra.image = 0;
    ra.stats = nullptr;
    ra.frame = nullptr;
    hr::Shading sh;
    sh.mats = c->mats.p ? c->mats.p : default_mat;
    sh.mat_idx = c->has_mat_idx ? c->mat_idx.p : nullptr;
    sh.lights = c->lights.p;
    sh.light_cdf = c->light_cdf.p;
    sh.nlights = c->nlights;
    sh.normals = c->has_normals ? c->normals.p : nullptr;
    sh.nidx = c->has_normals ? c->nidx.p : nullptr;
    hr::fill_u32((uint32_t*)rgb, 0u, 3 * (size_t)n, nullptr, 1, s);
    if (c->tuning.persistent_path)
        hr::persist_pt(trace_scene(c), ra, n, hits, prims, sh, rgb, pool ? pool : c->pool.p,
                       (uint32_t)c->tuning.path_refill, (uint32_t)c->tuning.path_shade, c->ray_count.p, s);
    else
        hr::launch_pt(trace_scene(c), ra, n, hits, prims, sh, rgb, c->ray_count.p, s, pt_wave_steps(c));
    HR_HIP(hipGetLastError());
    return HR_OK;
}
}  // namespace

hr_status hr::ctx_path_trace(hr_ctx* c, const float* hits, const uint32_t* prims, uint32_t n, int depth, float* rgb,
                             hipStream_t s)
{
    return path_trace_hits(c, hits, prims, n, depth, rgb, s, nullptr);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace {
hr_status check_frame_args(hr_ctx* c, uint32_t k1, uint32_t k2, hr_mode mode, int image_id,
                           const hr_render_params* p)
{
    HR_TRY(no_hierarchy(c, "captured frames"));
    if (c->nframes == 0) return fail(HR_ERR_STATE, "captured frames need resident key frames");
    if (k1 >= c->nframes || k2 >= c->nframes) return fail(HR_ERR_ARG, "key frame index out of range");
    return check_render_args(c, mode, image_id, p);
}
}  // namespace

extern "C" {

const char* hr_last_error(void) { return g_last_error.c_str(); }

hr_status hr_create(hr_ctx** out, const hr_options* opts)
{
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(HR_ERR_DEVICE, "no HIP device available");
    hr_ctx* c = new hr_ctx();
    c->device = opts ? opts->device : 0;
    c->seed_lo = opts ? opts->seed_lo : 0u;
    c->seed_hi = opts ? opts->seed_hi : 0u;
    if (c->device < 0 || c->device >= n) {
        delete c;
        return fail(HR_ERR_ARG, "device ordinal out of range");
    }
    hr_status s = HR_OK;
    do {
        if ((s = set_device(c)) != HR_OK) break;
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
        if ((s = c->params.reserve(1)) != HR_OK) break;
        if ((s = c->status.reserve(1)) != HR_OK) break;
        if ((s = c->bounds_part.reserve(6 * (size_t)hr::kBoundsParts)) != HR_OK) break;
        if ((s = c->ray_count.reserve(1)) != HR_OK) break;
        if ((s = c->work.reserve(6)) != HR_OK) break;
        if ((s = c->pool.reserve(1)) != HR_OK) break;
        e = hipHostMalloc((void**)&c->h_status, sizeof(hr::BuildStatus), hipHostMallocDefault);
        if (e != hipSuccess) { s = fail(HR_ERR_OOM, hipGetErrorString(e)); break; }
        e = hipHostMalloc((void**)&c->h_pin, 2 * sizeof(hr_ctx::FramePin), hipHostMallocDefault);
        if (e != hipSuccess) { s = fail(HR_ERR_OOM, hipGetErrorString(e)); break; }
        std::memset((void*)c->h_pin, 0, 2 * sizeof(hr_ctx::FramePin));
        if ((s = c->frame_sel.reserve(1)) != HR_OK) break;
        if ((s = c->halt.reserve(1)) != HR_OK) break;
        for (auto& slot : c->ev_g)
            for (hipEvent_t& ev : slot) {
                e = hipEventCreate(&ev);
                if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
            }
        for (auto& slot : c->ev_k)
            for (hipEvent_t& ev : slot) {
                e = hipEventCreate(&ev);
                if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
            }
        for (hipEvent_t* ev : { &c->ev_built[0], &c->ev_built[1], &c->ev_blend }) {
            e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
            if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
        }
        if (s != HR_OK) break;
        if ((s = noalloc_scratch(c->device, &c->noalloc)) != HR_OK) break;
        hipEvent_t* evs[6] = { &c->ev_b0, &c->ev_b1, &c->ev_r0, &c->ev_r1, &c->ev_p1, &c->ev_p2 };
        for (hipEvent_t* ev : evs) {
            e = hipEventCreate(ev);
            if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
        }
        if (s != HR_OK) break;
        e = hipMemsetAsync(c->ray_count.p, 0, sizeof(unsigned long long), c->stream);
        if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
        e = hipMemsetAsync(c->work.p, 0, 6 * sizeof(unsigned long long), c->stream);
        if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
        e = hipMemsetAsync(c->halt.p, 0xFF, sizeof(uint32_t), c->stream);
        if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
        e = hipEventCreateWithFlags(&c->ev_switch, hipEventDisableTiming);
        if (e != hipSuccess) { s = fail(HR_ERR_DEVICE, hipGetErrorString(e)); break; }
    } while (0);
    if (s != HR_OK) {
        hr_destroy(c);
        return s;
    }
    *out = c;
    return HR_OK;
}

hr_status hr_destroy(hr_ctx* c)
{
    if (!c) return HR_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stash.stream) (void)hipStreamSynchronize(c->stash.stream);
    c->normals.release();
    c->nidx.release();
    c->normal_keys.release();
    DevBuf<float>* fbufs[] = { &c->key_a, &c->key_b, &c->keyframes, &c->verts, &c->mats, &c->lights,
                               &c->light_cdf, &c->hits, &c->fb, &c->ray_org, &c->ray_dir, &c->ray_t,
                               &c->ray_hits, &c->ray_rgb };
    for (auto* b : fbufs) b->release();
    c->pm.release();
    c->pm_prims.release();
    c->ray_wire.release();
    DevBuf<uint32_t>* ubufs[] = { &c->idx, &c->mat_idx, &c->wave_counts, &c->wave_offsets, &c->scan_chain,
                                  &c->keys, &c->vals,
                                  &c->keys_sorted, &c->top_vals_sorted, &c->prims, &c->leaf_counts,
                                  &c->sort_chain, &c->sort_keys_tmp, &c->sort_vals_tmp, &c->sort_hist,
                                  &c->hit_prims, &c->ray_prim, &c->tri_obj, &c->obj_first,
                                  &c->obj_bounds_enc, &c->obj_cell_counts, &c->obj_cell_base, &c->inst_counts,
                                  &c->inst_offsets, &c->hit_inst, &c->ray_inst, &c->obj_part, &c->bounds_part };
    for (auto* b : ubufs) b->release();
    c->occ.release();
    c->ray_occ.release();
    c->inst.release();
    c->inst_grid.release();
    c->obj_gp.release();
    c->rec.release();
    c->cells.release();
    c->leaves.release();
    c->params.release();
    c->status.release();
    c->ray_count.release();
    c->work.release();
    c->pool.release();
    drop_graphs(c);
    c->frame_sel.release();
    c->halt.release();
    c->rad.release();
    c->frame_rays.release();
    release_lane(c->stash);
    if (c->ev_switch) (void)hipEventDestroy(c->ev_switch);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    for (auto& slot : c->ev_g)
        for (hipEvent_t ev : slot)
            if (ev) (void)hipEventDestroy(ev);
    for (auto& slot : c->ev_k)
        for (hipEvent_t ev : slot)
            if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : { c->ev_built[0], c->ev_built[1], c->ev_blend })
        if (ev) (void)hipEventDestroy(ev);
    hipEvent_t evs[6] = { c->ev_b0, c->ev_b1, c->ev_r0, c->ev_r1, c->ev_p1, c->ev_p2 };
    for (hipEvent_t ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HR_OK;
}

hr_status hr_get_stream(hr_ctx* c, void** stream)
{
    HR_TRY(check_ctx(c));
    if (!stream) return fail(HR_ERR_ARG, "null output pointer");
    *stream = (void*)c->stream;
    return HR_OK;
}

hr_status hr_upload_mesh(hr_ctx* c, const float* xyz1, const float* xyz2, float coeff, uint32_t nv,
                         const uint32_t* idx, uint32_t nt)
{
    HR_TRY(check_ctx(c));
    reset_primary_choice(c);
    if (!xyz1 || nv == 0) return fail(HR_ERR_ARG, "a mesh needs at least one vertex");
    if (nt > 0 && !idx) return fail(HR_ERR_ARG, "null index buffer");
    if (nt > 0x3FFFFFFFu) return fail(HR_ERR_ARG, "too many triangles");
    HR_TRY(validate_indices(idx, nt, nv));
    const size_t vb = 3 * (size_t)nv * sizeof(float);
    HR_TRY(c->key_a.reserve(3 * (size_t)nv));
    HR_TRY(upload(c->key_a.p, xyz1, vb, c->stream));
    if (xyz2 && xyz2 != xyz1) {
        HR_TRY(c->key_b.reserve(3 * (size_t)nv));
        HR_TRY(upload(c->key_b.p, xyz2, vb, c->stream));
        c->src2 = c->key_b.p;
    } else {
        c->src2 = c->key_a.p;
    }
    c->src1 = c->key_a.p;
    c->coeff = coeff;
    // the resident key frames (and graphs recorded over them) describe the
    // previous mesh: frame selection needs hr_set_keyframes again
    c->nframes = 0;
    drop_graphs(c);
    c->geo_frames = c->src2 == c->src1 ? 1u : 2u;
    c->k1 = 0;
    c->k2 = c->geo_frames - 1u;
    c->has_normals = false;
    HR_TRY(c->verts.reserve(3 * (size_t)nv));
    HR_TRY(c->idx.reserve(3 * (size_t)std::max(nt, 1u)));
    HR_TRY(upload(c->idx.p, idx, 3 * (size_t)nt * sizeof(uint32_t), c->stream));
    c->nv = nv;
    c->nt = nt;
    c->built = false;
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_upload_materials(hr_ctx* c, const float* mats, uint32_t nmat, const uint32_t* mat_idx,
                              uint32_t nt)
{
    HR_TRY(check_ctx(c));
    if (!mats || nmat == 0) return fail(HR_ERR_ARG, "empty material table");
    HR_TRY(c->mats.reserve(12 * (size_t)nmat));
    HR_TRY(upload(c->mats.p, mats, 12 * (size_t)nmat * sizeof(float), c->stream));
    c->has_mat_idx = false;
    if (mat_idx) {
        if (nt != c->nt) return fail(HR_ERR_ARG, "material index count differs from the triangle count");
        for (uint32_t i = 0; i < nt; ++i)
            if (mat_idx[i] >= nmat) return fail(HR_ERR_ARG, "material index out of range");
        HR_TRY(c->mat_idx.reserve(std::max(nt, 1u)));
        HR_TRY(upload(c->mat_idx.p, mat_idx, (size_t)nt * sizeof(uint32_t), c->stream));
        c->has_mat_idx = true;
    }
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_set_keyframes(hr_ctx* c, const float* xyz, uint32_t nframes, uint32_t nv, const uint32_t* idx,
                           uint32_t nt)
{
    HR_TRY(check_ctx(c));
    reset_primary_choice(c);
    if (!xyz || nframes == 0 || nv == 0) return fail(HR_ERR_ARG, "empty key frames");
    if (nt > 0 && !idx) return fail(HR_ERR_ARG, "null index buffer");
    HR_TRY(validate_indices(idx, nt, nv));
    const size_t n = (size_t)nframes * 3 * nv;
    HR_TRY(c->keyframes.reserve(n));
    HR_TRY(upload(c->keyframes.p, xyz, n * sizeof(float), c->stream));
    HR_TRY(c->verts.reserve(3 * (size_t)nv));
    HR_TRY(c->idx.reserve(3 * (size_t)std::max(nt, 1u)));
    HR_TRY(upload(c->idx.p, idx, 3 * (size_t)nt * sizeof(uint32_t), c->stream));
    c->nframes = nframes;
    c->geo_frames = nframes;
    c->k1 = c->k2 = 0;
    c->has_normals = false;
    c->nv = nv;
    c->nt = nt;
    c->src1 = c->src2 = c->keyframes.p;
    c->coeff = 1.f;
    c->built = false;
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_set_frame(hr_ctx* c, uint32_t k1, uint32_t k2, float coeff)
{
    HR_TRY(check_ctx(c));
    if (c->nframes == 0) return fail(HR_ERR_STATE, "no key frames resident");
    if (k1 >= c->nframes || k2 >= c->nframes) return fail(HR_ERR_ARG, "key frame index out of range");
    c->src1 = c->keyframes.p + (size_t)k1 * 3 * c->nv;
    c->src2 = c->keyframes.p + (size_t)k2 * 3 * c->nv;
    c->coeff = coeff;
    c->k1 = k1;
    c->k2 = k2;
    c->normals_dirty = true;
    c->built = false;
    return HR_OK;
}

hr_status hr_set_normals(hr_ctx* c, const float* normals, uint32_t nframes, uint32_t nn, const uint32_t* nidx,
                         uint32_t nt)
{
    HR_TRY(check_ctx(c));
    if (!c->src1) return fail(HR_ERR_STATE, "normals before geometry");
    if (!normals || !nidx || nn == 0) return fail(HR_ERR_ARG, "empty normals");
    if (nt != c->nt) return fail(HR_ERR_ARG, "normal index count differs from the triangle count");
    if (nframes != 1 && nframes != c->geo_frames)
        return fail(HR_ERR_ARG, "normals need 1 block or one per geometry key frame");
    for (size_t i = 0; i < 3 * (size_t)nt; ++i)
        if (nidx[i] >= nn) return fail(HR_ERR_ARG, "normal index out of range");
    const size_t n = (size_t)nframes * 3 * nn;
    HR_TRY(c->normal_keys.reserve(n));
    HR_TRY(c->nidx.reserve(3 * (size_t)std::max(nt, 1u)));
    HR_TRY(c->normals.reserve(nn));
    HR_TRY(upload(c->normal_keys.p, normals, n * sizeof(float), c->stream));
    HR_TRY(upload(c->nidx.p, nidx, 3 * (size_t)nt * sizeof(uint32_t), c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    c->nn = nn;
    c->nframes_n = nframes;
    c->has_normals = true;
    c->normals_dirty = true;
    return HR_OK;
}

hr_status hr_set_grid(hr_ctx* c, hr_grid_kind kind, float lambda_top, float lambda_leaf)
{
    HR_TRY(check_ctx(c));
    reset_primary_choice(c);
    if (kind != HR_GRID_UNIFORM && kind != HR_GRID_UNIFORM_EXACT && kind != HR_GRID_TWO_LEVEL &&
        kind != HR_GRID_HIERARCHY)
        return fail(HR_ERR_ARG, "unknown grid kind");
    if (!(lambda_top > 0.f) || !(lambda_top < 1e6f)) return fail(HR_ERR_ARG, "lambda_top out of range");
    if ((kind == HR_GRID_TWO_LEVEL || kind == HR_GRID_HIERARCHY) && (!(lambda_leaf > 0.f) || !(lambda_leaf < 1e6f)))
        return fail(HR_ERR_ARG, "lambda_leaf out of range");
    if (kind != c->kind) {
        c->pairs_cap = c->leaves_cap = c->leaf_pairs_cap = 0;
        drop_graphs(c);
    }
    c->kind = kind;
    c->lambda_top = lambda_top;
    c->lambda_leaf = lambda_leaf;
    c->built = false;
    return HR_OK;
}

hr_status hr_build(hr_ctx* c)
{
    HR_TRY(check_ctx(c));
    c->graph_timed = false;   // hr_get_timing reports this build, not the last replay
    return do_build(c);
}

hr_status hr_get_grid_info(hr_ctx* c, hr_grid_info* out)
{
    HR_TRY(check_ctx(c));
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    if (!c->built) return fail(HR_ERR_STATE, "no grid built");
    *out = c->info;
    return HR_OK;
}

hr_status hr_get_grid(hr_ctx* c, uint32_t* cells, uint32_t* leaves, uint32_t* prims)
{
    HR_TRY(check_ctx(c));
    if (!c->built) return fail(HR_ERR_STATE, "no grid built");
    HR_HIP(hipStreamSynchronize(c->stream));
    if (cells)
        HR_HIP(hipMemcpy(cells, c->cells.p, 2 * (size_t)c->info.num_cells * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    if (leaves && (c->kind == HR_GRID_TWO_LEVEL || c->kind == HR_GRID_HIERARCHY) && c->info.num_leaves)
        HR_HIP(hipMemcpy(leaves, c->leaves.p, 2 * (size_t)c->info.num_leaves * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    if (prims && c->info.num_refs)
        HR_HIP(hipMemcpy(prims, c->prims.p, (size_t)c->info.num_refs * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    return HR_OK;
}

hr_status hr_set_camera(hr_ctx* c, const float pos[3], const float dir[3], const float up[3], float fov_deg,
                        int res_x, int res_y)
{
    HR_TRY(check_ctx(c));
    reset_primary_choice(c);
    if (!pos || !dir || !up) return fail(HR_ERR_ARG, "null camera vector");
    if (res_x <= 0 || res_y <= 0 || (int64_t)res_x * res_y > (1ll << 28))
        return fail(HR_ERR_ARG, "invalid resolution");
    camera_init(c->cam, pos, dir, up, fov_deg, res_x, res_y);
    c->has_camera = true;
    // RTEngine::setCamera -> PathTracer::setResolution (RTEngine.cu:389)
    if (c->pm.p) hr::fill_u32((uint32_t*)c->pm.p, 0u, c->pm.cap, nullptr, 1, c->stream);
    return HR_OK;
}

hr_status hr_set_lights(hr_ctx* c, const hr_area_light* lights, const float* weights, int n)
{
    HR_TRY(check_ctx(c));
    if (n < 0 || (n > 0 && !lights)) return fail(HR_ERR_ARG, "invalid light list");
    c->nlights = n;
    if (n == 0) return HR_OK;
    std::vector<float> cdf((size_t)n);
    float acc = 0.f;
    for (int i = 0; i < n; ++i) {
        acc += weights ? weights[i] : 1.f;
        cdf[(size_t)i] = acc;
    }
    HR_TRY(c->lights.reserve(15 * (size_t)n));
    HR_TRY(c->light_cdf.reserve((size_t)n));
    HR_TRY(upload(c->lights.p, lights, 15 * (size_t)n * sizeof(float), c->stream));
    HR_TRY(upload(c->light_cdf.p, cdf.data(), (size_t)n * sizeof(float), c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_render(hr_ctx* c, hr_mode mode, int image_id, const hr_render_params* p, float* rgb_out)
{
    HR_TRY(check_ctx(c));
    if (!c->built) return fail(HR_ERR_STATE, "render before build");
    HR_TRY(check_render_args(c, mode, image_id, p));
    HR_TRY(prepare_render(c, mode, p));
    const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
    if (image_id == 0 || npix != c->last_npix) {
        hr::fill_u32((uint32_t*)c->fb.p, 0u, 3 * (size_t)npix, nullptr, 1, c->stream);
        c->last_npix = npix;
    }
    if (uses_normals(mode) && c->has_normals && c->normals_dirty) {
        enqueue_normals(c, nullptr);
        c->normals_dirty = false;
    }
    HR_TRY(enqueue_render(c, mode, image_id, p, nullptr, true));
    c->render_timed = true;
    c->graph_timed = false;
    if (rgb_out) {
        HR_HIP(hipMemcpyAsync(rgb_out, c->fb.p, 3 * (size_t)npix * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
        HR_HIP(hipStreamSynchronize(c->stream));
    }
    return HR_OK;
}

hr_status hr_render_frame_graph(hr_ctx* c, uint32_t k1, uint32_t k2, float coeff, hr_mode mode, int image_id,
                                const hr_render_params* p, float* rgb_out)
{
    HR_TRY(check_ctx(c));
    HR_TRY(check_frame_args(c, k1, k2, mode, image_id, p));
    HR_TRY(submit_frame(c, hr_ctx::Submitted{ k1, k2, coeff, image_id, 0 }, mode, p));
    HR_TRY(settle_frames(c));
    if (rgb_out) {
        const uint32_t npix = (uint32_t)c->cam.res_x * (uint32_t)c->cam.res_y;
        HR_HIP(hipMemcpyAsync(rgb_out, c->fb.p, 3 * (size_t)npix * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
        HR_HIP(hipStreamSynchronize(c->stream));
    }
    return HR_OK;
}

hr_status hr_submit_frame_graph(hr_ctx* c, uint32_t k1, uint32_t k2, float coeff, hr_mode mode, int image_id,
                                const hr_render_params* p)
{
    if (!c) return fail(HR_ERR_ARG, "null context");
    HR_TRY(set_device(c));
    HR_TRY(check_frame_args(c, k1, k2, mode, image_id, p));
    return submit_frame(c, hr_ctx::Submitted{ k1, k2, coeff, image_id, 0 }, mode, p);
}

hr_status hr_get_framebuffer(hr_ctx* c, float* rgb_out)
{
    HR_TRY(check_ctx(c));
    if (!c->last_npix) return fail(HR_ERR_STATE, "nothing rendered");
    if (!rgb_out) return fail(HR_ERR_ARG, "null output pointer");
    HR_HIP(hipMemcpyAsync(rgb_out, c->fb.p, 3 * (size_t)c->last_npix * sizeof(float), hipMemcpyDeviceToHost,
                          c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_get_primary_hits(hr_ctx* c, float* planes7, uint32_t* prims)
{
    HR_TRY(check_ctx(c));
    if (!c->last_npix) return fail(HR_ERR_STATE, "nothing rendered");
    HR_HIP(hipStreamSynchronize(c->stream));
    if (planes7)
        HR_HIP(hipMemcpy(planes7, c->hits.p, 7 * (size_t)c->last_npix * sizeof(float), hipMemcpyDeviceToHost));
    if (prims)
        HR_HIP(hipMemcpy(prims, c->hit_prims.p, (size_t)c->last_npix * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    return HR_OK;
}

hr_status hr_get_ao_occlusion(hr_ctx* c, uint8_t* occ)
{
    HR_TRY(check_ctx(c));
    if (!c->occ_npix || c->occ_npix != c->last_npix || !c->last_ao_samples)
        return fail(HR_ERR_STATE, "the last render was not an AO render");
    if (!occ) return fail(HR_ERR_ARG, "null output pointer");
    HR_HIP(hipStreamSynchronize(c->stream));
    HR_HIP(hipMemcpy(occ, c->occ.p, (size_t)c->occ_npix * (size_t)c->last_ao_samples, hipMemcpyDeviceToHost));
    return HR_OK;
}

// 24-byte ray records on the host -> the context's SoA ray planes
static hr_status upload_wire(hr_ctx* c, const float* rays, uint32_t n)
{
    HR_TRY(c->ray_org.reserve(3 * (size_t)n));
    HR_TRY(c->ray_dir.reserve(3 * (size_t)n));
    HR_TRY(c->ray_wire.reserve(6 * (size_t)n));
    HR_TRY(upload(c->ray_wire.p, rays, 6 * (size_t)n * sizeof(float), c->stream));
    hr::launch_wire_to_soa(c->ray_wire.p, n, c->ray_org.p, c->ray_dir.p, n, c->stream);
    return HR_OK;
}

// wire: org holds n 24-byte ray records in host memory and dir is unused
static hr_status trace_common(hr_ctx* c, const float* org, const float* dir, uint32_t n, int anyhit,
                              float tmax, float* t_out, uint32_t* prim_out, uint8_t* occ_out, bool device,
                              bool wire = false, uint32_t* inst_out = nullptr)
{
    HR_TRY(check_ctx(c));
    if (!c->built) return fail(HR_ERR_STATE, "trace before build");
    if (n == 0) return HR_OK;
    if (!org || (!dir && !wire)) return fail(HR_ERR_ARG, "null ray buffer");
    const hr::TraceScene S = trace_scene(c);
    const bool hier = c->kind == HR_GRID_HIERARCHY;
    const bool pers = c->tuning.persistent_secondary != 0;
    const uint32_t refill = (uint32_t)c->tuning.refill;
    if (device) {
        if (hier)
            hr::launch_trace_hier(hier_scene(c), org, dir, n, anyhit, tmax, t_out, prim_out, nullptr, occ_out,
                                  c->ray_count.p, c->stream);
        else if (pers)
            hr::persist_batch(S, org, dir, n, anyhit, tmax, t_out, prim_out, occ_out, c->pool.p, refill,
                              c->ray_count.p, c->work.p, c->stream, nullptr, persist_opts(c));
        else
            hr::launch_trace_batch(S, org, dir, n, anyhit, tmax, t_out, prim_out, occ_out, c->ray_count.p,
                                   c->work.p, c->stream);
        HR_HIP(hipGetLastError());
        return HR_OK;
    }
    HR_TRY(c->ray_org.reserve(3 * (size_t)n));
    HR_TRY(c->ray_dir.reserve(3 * (size_t)n));
    HR_TRY(c->ray_t.reserve(n));
    HR_TRY(c->ray_prim.reserve(n));
    HR_TRY(c->ray_occ.reserve(n));
    if (wire) {
        HR_TRY(upload_wire(c, org, n));
    } else {
        HR_TRY(upload(c->ray_org.p, org, 3 * (size_t)n * sizeof(float), c->stream));
        HR_TRY(upload(c->ray_dir.p, dir, 3 * (size_t)n * sizeof(float), c->stream));
    }
    if (inst_out) HR_TRY(c->ray_inst.reserve(n));
    if (hier)
        hr::launch_trace_hier(hier_scene(c), c->ray_org.p, c->ray_dir.p, n, anyhit, tmax, c->ray_t.p, c->ray_prim.p,
                              inst_out ? c->ray_inst.p : nullptr, c->ray_occ.p, c->ray_count.p, c->stream);
    else if (pers)
        hr::persist_batch(S, c->ray_org.p, c->ray_dir.p, n, anyhit, tmax, c->ray_t.p, c->ray_prim.p,
                          c->ray_occ.p, c->pool.p, refill, c->ray_count.p, c->work.p, c->stream, nullptr,
                          persist_opts(c));
    else
        hr::launch_trace_batch(S, c->ray_org.p, c->ray_dir.p, n, anyhit, tmax, c->ray_t.p, c->ray_prim.p,
                               c->ray_occ.p, c->ray_count.p, c->work.p, c->stream);
    HR_HIP(hipGetLastError());
    if (t_out) HR_HIP(hipMemcpyAsync(t_out, c->ray_t.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (prim_out)
        HR_HIP(hipMemcpyAsync(prim_out, c->ray_prim.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
    if (occ_out) HR_HIP(hipMemcpyAsync(occ_out, c->ray_occ.p, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (inst_out)
        HR_HIP(hipMemcpyAsync(inst_out, c->ray_inst.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_trace(hr_ctx* c, const float* org, const float* dir, uint32_t n, float* t_out, uint32_t* prim_out)
{
    if (!t_out || !prim_out) return fail(HR_ERR_ARG, "null output buffer");
    return trace_common(c, org, dir, n, 0, 0.f, t_out, prim_out, nullptr, false);
}

hr_status hr_occluded(hr_ctx* c, const float* org, const float* dir, uint32_t n, float tmax, uint8_t* occ_out)
{
    if (!occ_out) return fail(HR_ERR_ARG, "null output buffer");
    if (!(tmax > 0.f)) return fail(HR_ERR_ARG, "tmax must be positive");
    return trace_common(c, org, dir, n, 1, tmax, nullptr, nullptr, occ_out, false);
}

hr_status hr_trace_wire(hr_ctx* c, const float* rays, uint32_t n, float* t_out, uint32_t* prim_out)
{
    if (!t_out || !prim_out) return fail(HR_ERR_ARG, "null output buffer");
    return trace_common(c, rays, nullptr, n, 0, 0.f, t_out, prim_out, nullptr, false, true);
}

hr_status hr_occluded_wire(hr_ctx* c, const float* rays, uint32_t n, float tmax, uint8_t* occ_out)
{
    if (!occ_out) return fail(HR_ERR_ARG, "null output buffer");
    if (!(tmax > 0.f)) return fail(HR_ERR_ARG, "tmax must be positive");
    return trace_common(c, rays, nullptr, n, 1, tmax, nullptr, nullptr, occ_out, false, true);
}

hr_status hr_render_rays(hr_ctx* c, const float* rays, uint32_t n, float* rgb_out)
{
    HR_TRY(check_ctx(c));
    HR_TRY(no_hierarchy(c, "input-ray rendering"));
    if (!c->built) return fail(HR_ERR_STATE, "render before build");
    if (!rgb_out) return fail(HR_ERR_ARG, "null output buffer");
    if (n == 0) return HR_OK;
    if (!rays) return fail(HR_ERR_ARG, "null ray buffer");
    HR_TRY(upload_wire(c, rays, n));
    HR_TRY(c->ray_t.reserve(n));
    HR_TRY(c->ray_prim.reserve(n));
    HR_TRY(c->ray_hits.reserve(7 * (size_t)n));
    HR_TRY(c->ray_rgb.reserve(3 * (size_t)n));
    // SimpleRayTraverser: trace, then simpleShade at image 0
    hr::persist_batch(trace_scene(c), c->ray_org.p, c->ray_dir.p, n, 0, 0.f, c->ray_t.p, c->ray_prim.p, nullptr,
                      c->pool.p, (uint32_t)c->tuning.refill, c->ray_count.p, c->work.p, c->stream, c->ray_hits.p,
                      persist_opts(c));
    hr::fill_u32((uint32_t*)c->ray_rgb.p, 0u, 3 * (size_t)n, nullptr, 1, c->stream);
    hr::RenderArgs ra{};
    ra.seed_lo = c->seed_lo;
    ra.seed_hi = c->seed_hi;
    hr::launch_simple_shade(trace_scene(c), ra, n, c->ray_hits.p, c->ray_prim.p, c->ray_rgb.p, c->stream);
    HR_HIP(hipGetLastError());
    HR_HIP(hipMemcpyAsync(rgb_out, c->ray_rgb.p, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_dump_ao_rays(hr_ctx* c, float* rays_out, uint64_t max_rays, uint64_t* count)
{
    HR_TRY(check_ctx(c));
    HR_TRY(no_hierarchy(c, "the AO ray dump"));
    if (!c->ao_dump_ok) return fail(HR_ERR_STATE, "the last render was not an AO render");
    const size_t npix = (size_t)c->last_ao_cam.res_x * (size_t)c->last_ao_cam.res_y;
    const size_t n = npix * (size_t)(c->last_ao_ra.ao_samples + 1);
    if (count) *count = n;
    if (!rays_out) return HR_OK;
    if (max_rays < n) return fail(HR_ERR_ARG, "output holds fewer records than the dump");
    HR_TRY(c->ray_wire.reserve(6 * n));
    hr::launch_ao_dump(trace_scene(c), c->last_ao_cam, c->last_ao_ra, c->hits.p, c->hit_prims.p, c->ray_wire.p, c->stream);
    HR_HIP(hipGetLastError());
    HR_HIP(hipMemcpyAsync(rays_out, c->ray_wire.p, 6 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_trace_dev(hr_ctx* c, const float* d_org, const float* d_dir, uint32_t n, float* d_t_out,
                       uint32_t* d_prim_out)
{
    if (!d_t_out || !d_prim_out) return fail(HR_ERR_ARG, "null output buffer");
    return trace_common(c, d_org, d_dir, n, 0, 0.f, d_t_out, d_prim_out, nullptr, true);
}

hr_status hr_occluded_dev(hr_ctx* c, const float* d_org, const float* d_dir, uint32_t n, float tmax,
                          uint8_t* d_occ_out)
{
    if (!d_occ_out) return fail(HR_ERR_ARG, "null output buffer");
    if (!(tmax > 0.f)) return fail(HR_ERR_ARG, "tmax must be positive");
    return trace_common(c, d_org, d_dir, n, 1, tmax, nullptr, nullptr, d_occ_out, true);
}

hr_status hr_path_trace(hr_ctx* c, const float* org, const float* dir, uint32_t n, int depth, float* rgb_out)
{
    HR_TRY(check_ctx(c));
    HR_TRY(no_hierarchy(c, "path tracing"));
    if (!c->built) return fail(HR_ERR_STATE, "path trace before build");
    if (!rgb_out) return fail(HR_ERR_ARG, "null output buffer");
    if (depth < 1 || depth > 64) return fail(HR_ERR_ARG, "depth must be in [1, 64]");
    if (n == 0) return HR_OK;
    if (!org || !dir) return fail(HR_ERR_ARG, "null ray buffer");
    HR_TRY(c->ray_org.reserve(3 * (size_t)n));
    HR_TRY(c->ray_dir.reserve(3 * (size_t)n));
    HR_TRY(c->ray_t.reserve(n));
    HR_TRY(c->ray_prim.reserve(n));
    HR_TRY(c->ray_hits.reserve(7 * (size_t)n));
    HR_TRY(c->ray_rgb.reserve(3 * (size_t)n));
    HR_TRY(upload(c->ray_org.p, org, 3 * (size_t)n * sizeof(float), c->stream));
    HR_TRY(upload(c->ray_dir.p, dir, 3 * (size_t)n * sizeof(float), c->stream));
    hr::persist_batch(trace_scene(c), c->ray_org.p, c->ray_dir.p, n, 0, 0.f, c->ray_t.p, c->ray_prim.p, nullptr,
                      c->pool.p, (uint32_t)c->tuning.refill, c->ray_count.p, c->work.p, c->stream, c->ray_hits.p,
                      persist_opts(c));
    HR_TRY(hr::ctx_path_trace(c, c->ray_hits.p, c->ray_prim.p, n, depth, c->ray_rgb.p, c->stream));
    HR_HIP(hipMemcpyAsync(rgb_out, c->ray_rgb.p, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

hr_status hr_path_trace_dev(hr_ctx* c, const float* d_org, const float* d_dir, uint32_t n, int depth,
                            float* d_rgb_out)
{
    HR_TRY(check_ctx(c));
    HR_TRY(no_hierarchy(c, "path tracing"));
    if (!c->built) return fail(HR_ERR_STATE, "path trace before build");
    if (!d_rgb_out) return fail(HR_ERR_ARG, "null output buffer");
    if (depth < 1 || depth > 64) return fail(HR_ERR_ARG, "depth must be in [1, 64]");
    if (n == 0) return HR_OK;
    if (!d_org || !d_dir) return fail(HR_ERR_ARG, "null ray buffer");
    HR_TRY(c->ray_t.reserve(n));
    HR_TRY(c->ray_prim.reserve(n));
    HR_TRY(c->ray_hits.reserve(7 * (size_t)n));
    hr::persist_batch(trace_scene(c), d_org, d_dir, n, 0, 0.f, c->ray_t.p, c->ray_prim.p, nullptr, c->pool.p,
                      (uint32_t)c->tuning.refill, c->ray_count.p, c->work.p, c->stream, c->ray_hits.p,
                      persist_opts(c));
    return hr::ctx_path_trace(c, c->ray_hits.p, c->ray_prim.p, n, depth, d_rgb_out, c->stream);
}

hr_status hr_path_trace_dev_noalloc(hr_ctx* c, const float* d_org, const float* d_dir, uint32_t n, int depth,
                                    float* d_rgb_out)
{
    HR_TRY(check_ctx(c));
    HR_TRY(no_hierarchy(c, "path tracing"));
    if (!c->built) return fail(HR_ERR_STATE, "path trace before build");
    if (!d_rgb_out) return fail(HR_ERR_ARG, "null output buffer");
    if (depth < 1 || depth > 64) return fail(HR_ERR_ARG, "depth must be in [1, 64]");
    if (n > kNoAllocMaxRays)
        return fail(HR_ERR_ARG, "hr_path_trace_dev_noalloc takes at most 2097152 rays per call");
    if (n == 0) return HR_OK;
    if (!d_org || !d_dir) return fail(HR_ERR_ARG, "null ray buffer");
    if (!c->noalloc) return fail(HR_ERR_STATE, "no scratch address for this device");
    unsigned char* base = c->noalloc;
    uint32_t* pool = (uint32_t*)base;                       // cleared by persist_batch's fill kernel
    const float* default_mat = (const float*)(base + 64);
    float* t = (float*)(base + kNoAllocHead);
    uint32_t* prim = (uint32_t*)(t + kNoAllocMaxRays);
    float* hits = (float*)(prim + kNoAllocMaxRays);         // 7 planes of stride n
    hr::persist_batch(trace_scene(c), d_org, d_dir, n, 0, 0.f, t, prim, nullptr, pool, (uint32_t)c->tuning.refill,
                      c->ray_count.p, c->work.p, c->stream, hits, persist_opts(c));
    return path_trace_hits(c, hits, prim, n, depth, d_rgb_out, c->stream, default_mat, pool);
}

hr_status hr_get_timing(hr_ctx* c, hr_timing* out)
{
    HR_TRY(check_ctx(c));
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    HR_HIP(hipStreamSynchronize(c->stream));
    out->build_ms = out->render_ms = out->primary_ms = out->secondary_ms = out->shade_ms = 0.0;
    float ms = 0.f;
    if (c->graph_timed) {
        // the last completed replay: rebuild graph, render and its parts
        out->build_ms = c->graph_build_ms;
        out->render_ms = c->graph_render_ms;
        out->primary_ms = c->graph_primary_ms;
        out->secondary_ms = c->graph_secondary_ms;
        out->shade_ms = c->graph_shade_ms;
    } else {
        if (c->build_timed && hipEventElapsedTime(&ms, c->ev_b0, c->ev_b1) == hipSuccess) out->build_ms = ms;
        if (c->render_timed) {
            if (hipEventElapsedTime(&ms, c->ev_r0, c->ev_r1) == hipSuccess) out->render_ms = ms;
            if (hipEventElapsedTime(&ms, c->ev_r0, c->ev_p1) == hipSuccess) out->primary_ms = ms;
            if (hipEventElapsedTime(&ms, c->ev_p1, c->ev_p2) == hipSuccess) out->secondary_ms = ms;
            if (hipEventElapsedTime(&ms, c->ev_p2, c->ev_r1) == hipSuccess) out->shade_ms = ms;
        }
        (void)hipGetLastError();   // an interval not recorded reads as 0; leave no error behind
    }
    unsigned long long rays = 0;
    HR_HIP(hipMemcpy(&rays, c->ray_count.p, sizeof rays, hipMemcpyDeviceToHost));
    out->rays_traced = rays;
    out->primary_kernel = c->last_pk;
    out->frames_replayed = c->replays;
    out->replay_build_ms = c->replay_ms[0];
    out->replay_render_ms = c->replay_ms[1];
    out->replay_primary_ms = c->replay_ms[2];
    out->replay_secondary_ms = c->replay_ms[3];
    out->replay_shade_ms = c->replay_ms[4];
    return HR_OK;
}

hr_status hr_get_work_stats(hr_ctx* c, hr_work_stats* out)
{
    HR_TRY(check_ctx(c));
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    unsigned long long w[6] = { 0, 0, 0, 0, 0, 0 };
    HR_HIP(hipStreamSynchronize(c->stream));
    HR_HIP(hipMemcpy(w, c->work.p, sizeof w, hipMemcpyDeviceToHost));
    hr::fill_u32((uint32_t*)c->work.p, 0u, 12, nullptr, 1, c->stream);
    HR_HIP(hipStreamSynchronize(c->stream));
    out->primary.rays = w[0];
    out->primary.steps = w[1];
    out->primary.tests = w[2];
    out->secondary.rays = w[3];
    out->secondary.steps = w[4];
    out->secondary.tests = w[5];
    return HR_OK;
}

hr_status hr_set_tuning(hr_ctx* c, const hr_tuning* t)
{
    HR_TRY(check_ctx(c));
    if (!t) return fail(HR_ERR_ARG, "null tuning");
    if (t->persistent_primary < 0 || t->persistent_primary > 3)
        return fail(HR_ERR_ARG, "persistent_primary must be 0, 1, 2 or 3");
    if (t->persistent_secondary != 0 && t->persistent_secondary != 1)
        return fail(HR_ERR_ARG, "persistent_secondary must be 0 or 1");
    if (t->refill < 1 || t->refill > 64) return fail(HR_ERR_ARG, "refill must be in [1, 64]");
    if (t->hier_refill < 1 || t->hier_refill > 64) return fail(HR_ERR_ARG, "hier_refill must be in [1, 64]");
    if (t->hier_bounds_blocks < 0 || t->hier_bounds_blocks > (int)hr::kObjBoundsBlocks)
        return fail(HR_ERR_ARG, "hier_bounds_blocks must be in [0, 2048]");
    if (t->persistent_path != 0 && t->persistent_path != 1) return fail(HR_ERR_ARG, "persistent_path must be 0 or 1");
    if (t->path_refill < 1 || t->path_refill > 64) return fail(HR_ERR_ARG, "path_refill must be in [1, 64]");
    if (t->path_shade < 1 || t->path_shade > 64) return fail(HR_ERR_ARG, "path_shade must be in [1, 64]");
    if (t->fuse_units != 0 && t->fuse_units != 1) return fail(HR_ERR_ARG, "fuse_units must be 0 or 1");
    if (t->test_batch != 0 && t->test_batch != 1) return fail(HR_ERR_ARG, "test_batch must be 0 or 1");
    if (t->advance_steps < 1 || t->advance_steps > 16) return fail(HR_ERR_ARG, "advance_steps must be in [1, 16]");
    if (t->frame_lanes < 1 || t->frame_lanes > 3) return fail(HR_ERR_ARG, "frame_lanes must be 1, 2 or 3");
    if (t->camera_blocks < 0 || t->camera_blocks > 16) return fail(HR_ERR_ARG, "camera_blocks must be in [0, 16]");
    c->tuning = *t;
    reset_primary_choice(c);
    return HR_OK;
}

hr_status hr_get_tuning(hr_ctx* c, hr_tuning* t)
{
    HR_TRY(check_ctx(c));
    if (!t) return fail(HR_ERR_ARG, "null output pointer");
    *t = c->tuning;
    return HR_OK;
}

hr_status hr_synchronize(hr_ctx* c)
{
    HR_TRY(check_ctx(c));
    HR_HIP(hipStreamSynchronize(c->stream));
    return HR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// instanced hierarchy
// ---------------------------------------------------------------------------

hr_status hr_set_instances(hr_ctx* c, const uint32_t* obj_first, uint32_t nobj, const hr_instance* inst,
                           uint32_t ninst)
{
    HR_TRY(check_ctx(c));
    if (nobj == 0 || !obj_first) return fail(HR_ERR_ARG, "at least one object is needed");
    if (ninst == 0 || !inst) return fail(HR_ERR_ARG, "at least one instance is needed");
    if (obj_first[0] != 0) return fail(HR_ERR_ARG, "the first object starts at triangle 0");
    for (uint32_t k = 0; k < nobj; ++k)
        if (obj_first[k + 1] < obj_first[k]) return fail(HR_ERR_ARG, "object triangle ranges must not decrease");
    if (obj_first[nobj] > 0x3FFFFFFFu) return fail(HR_ERR_CAPACITY, "too many triangles");
    std::vector<hr::InstanceDev> h(ninst);
    for (uint32_t i = 0; i < ninst; ++i) {
        const hr_instance& a = inst[i];
        if (a.object >= nobj) return fail(HR_ERR_ARG, "instance object id out of range");
        for (int r = 0; r < 3; ++r) {
            bool ok = std::isfinite(a.aabb_min[r]) && std::isfinite(a.aabb_max[r]);
            for (int k = 0; k < 4; ++k) ok = ok && std::isfinite(a.m[r][k]);
            if (!ok) return fail(HR_ERR_ARG, "instance transform and box must be finite");
        }
        hr::InstanceDev& d = h[i];
        for (int k = 0; k < 3; ++k) {
            d.bmin[k] = a.aabb_min[k];
            d.bmax[k] = a.aabb_max[k];
            // GeometryInstanceMatrix::setTransformation (TwoLevelGridHierarchy.h:618-632)
            d.irot0[k] = a.m[k][0];
            d.irot1[k] = a.m[k][1];
            d.irot2[k] = a.m[k][2];
            d.it[k] = a.m[k][3];
        }
        d.object = a.object;
        d.pad = 0;
    }
    const uint32_t ntri = obj_first[nobj];
    std::vector<uint32_t> tri_obj(std::max(ntri, 1u), 0u);
    for (uint32_t k = 0; k < nobj; ++k)
        for (uint32_t t = obj_first[k]; t < obj_first[k + 1]; ++t) tri_obj[t] = k;
    HR_TRY(c->inst.reserve(ninst));
    HR_TRY(c->obj_first.reserve((size_t)nobj + 1));
    HR_TRY(c->tri_obj.reserve(tri_obj.size()));
    HR_TRY(upload(c->inst.p, h.data(), (size_t)ninst * sizeof(hr::InstanceDev), c->stream));
    HR_TRY(upload(c->obj_first.p, obj_first, ((size_t)nobj + 1) * sizeof(uint32_t), c->stream));
    HR_TRY(upload(c->tri_obj.p, tri_obj.data(), tri_obj.size() * sizeof(uint32_t), c->stream));
    HR_HIP(hipStreamSynchronize(c->stream));
    c->nobj = nobj;
    c->ninst = ninst;
    c->obj_tris = ntri;
    if (c->kind == HR_GRID_HIERARCHY) c->built = false;
    return HR_OK;
}

hr_status hr_get_object_grids(hr_ctx* c, hr_object_grid* grids, uint32_t* inst_refs)
{
    HR_TRY(check_ctx(c));
    if (c->kind != HR_GRID_HIERARCHY) return fail(HR_ERR_STATE, "the grid is not an instanced hierarchy");
    if (!c->built) return fail(HR_ERR_STATE, "no grid built");
    HR_HIP(hipStreamSynchronize(c->stream));
    if (grids) {
        std::vector<hr::GridParams> gp(c->nobj);
        std::vector<uint32_t> base(c->nobj);
        HR_HIP(hipMemcpy(gp.data(), c->obj_gp.p, (size_t)c->nobj * sizeof(hr::GridParams), hipMemcpyDeviceToHost));
        HR_HIP(hipMemcpy(base.data(), c->obj_cell_base.p, (size_t)c->nobj * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < c->nobj; ++k) {
            hr_object_grid& o = grids[k];
            for (int i = 0; i < 3; ++i) {
                o.bmin[i] = gp[k].bmin[i];
                o.bmax[i] = gp[k].bmax[i];
                o.res[i] = gp[k].res[i];
                o.cell_size[i] = gp[k].cs[i];
                o.cell_size_rcp[i] = gp[k].csr[i];
            }
            o.first_cell = base[k];
            o.num_cells = gp[k].num_cells;
        }
    }
    if (inst_refs && c->info.num_top_pairs)
        HR_HIP(hipMemcpy(inst_refs, c->top_vals_sorted.p, (size_t)c->info.num_top_pairs * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    return HR_OK;
}

hr_status hr_trace_instanced(hr_ctx* c, const float* org, const float* dir, uint32_t n, float* t_out,
                             uint32_t* prim_out, uint32_t* inst_out)
{
    HR_TRY(check_ctx(c));
    if (c->kind != HR_GRID_HIERARCHY) return fail(HR_ERR_STATE, "the grid is not an instanced hierarchy");
    if (!t_out || !prim_out || !inst_out) return fail(HR_ERR_ARG, "null output buffer");
    return trace_common(c, org, dir, n, 0, 0.f, t_out, prim_out, nullptr, false, false, inst_out);
}

hr_status hr_get_primary_instances(hr_ctx* c, uint32_t* inst)
{
    HR_TRY(check_ctx(c));
    if (c->kind != HR_GRID_HIERARCHY) return fail(HR_ERR_STATE, "the grid is not an instanced hierarchy");
    if (!c->last_npix || c->hier_npix != c->last_npix || c->hit_inst.cap < c->last_npix)
        return fail(HR_ERR_STATE, "the last render was not a hierarchy render");
    if (!inst) return fail(HR_ERR_ARG, "null output pointer");
    HR_HIP(hipStreamSynchronize(c->stream));
    HR_HIP(hipMemcpy(inst, c->hit_inst.p, (size_t)c->last_npix * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return HR_OK;
}
