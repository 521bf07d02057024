// hr_batch.hip -- batches of scenes (no reference counterpart; the reference
// renders one scene per process, src/RT/RTEngine.h:36-68).
//
// Many independent scenes stay resident in one set of device arrays.  One
// segmented two-level build constructs every scene's grid with the kernels
// of the single-scene build (global cell keys: scene s owns the cells
// [cell_base[s], cell_base[s+1])), and persistent launches trace the rays of
// all scenes together, each lane taking the next ray from a global counter
// whatever its scene.  Each scene's grid, read back scene-local, is the grid
// a single-scene build of that scene produces, and each ray's result is the
// single-scene result.
//
// The autotuner times candidate grid densities per scene-size class and
// candidate persistence settings on the caller's own rays, and keeps the
// fastest (build + trace) choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/happyray.h"
#include "hr_build.h"
#include "hr_host.h"
#include "hr_persist.h"

using hr::DevBuf;
using hr::fail;

struct hr_batch {
    int device = 0;
    hipStream_t stream = nullptr;

    // scenes
    uint32_t ns = 0, nv = 0, nt = 0;
    std::vector<uint32_t> h_vbase, h_tbase;       // ns + 1
    std::vector<float> h_lt, h_ll;                // per-scene densities
    DevBuf<float> verts;                          // SoA, 3 * nv
    DevBuf<uint32_t> idx, tri_scene, vbase, tbase;
    DevBuf<float> lam_top, lam_leaf;

    // per-scene grid data
    DevBuf<hr::GridParams> gps;
    DevBuf<uint32_t> bounds, cell_counts, cell_base, top_base, leaf_base, pair_base;

    // build buffers (all scenes)
    DevBuf<hr::GridParams> agg;
    DevBuf<hr::BuildStatus> status;
    hr::BuildStatus* h_status = nullptr;          // pinned
    DevBuf<float4> rec;
    DevBuf<uint32_t> wave_counts, wave_offsets, keys, vals, keys_sorted, top_vals_sorted, prims, leaf_counts;
    DevBuf<uint2> cells, leaves;
    DevBuf<uint32_t> scan_chain, sort_chain, sort_keys_tmp, sort_vals_tmp, sort_hist;
    uint64_t cells_cap = 0;
    uint32_t pairs_cap = 0, leaves_cap = 0, leaf_pairs_cap = 0;
    bool built = false;
    std::vector<hr::GridParams> h_gps;
    std::vector<uint32_t> h_cell_base, h_top_base, h_leaf_base, h_pair_base;
    DevBuf<hr::SceneDesc> descs;

    // traces
    DevBuf<hr::RayBatchDev> rb_dev[2];
    DevBuf<uint32_t> prefix_dev[2];
    hr::RayBatchDev* h_rb = nullptr;              // pinned staging
    uint32_t* h_prefix = nullptr;
    size_t h_rb_cap = 0, h_prefix_cap = 0;
    DevBuf<uint32_t> pool;
    DevBuf<unsigned long long> ray_count, work;
    hr_batch_tuning tuning{ 48, 0 };

    hipEvent_t ev_b0 = nullptr, ev_b1 = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
    bool build_timed = false, trace_timed = false;
};

namespace {

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t grow(uint64_t cur, uint64_t need)
{
    const uint64_t v = std::max<uint64_t>(need + need / 4 + 1024, cur + cur / 2);
    return (uint32_t)std::min<uint64_t>(v, 0xFFFFFFF0ull);
}

hr_status check_batch(hr_batch* b)
{
    if (!b) return fail(HR_ERR_ARG, "null batch");
    HR_HIP(hipSetDevice(b->device));
    return HR_OK;
}

hr_status reserve_build(hr_batch* b)
{
    const uint32_t pair_max = std::max(b->pairs_cap, b->leaf_pairs_cap);
    HR_TRY(b->rec.reserve((9 * (size_t)std::max(b->nt, 1u) + 3) / 4));   // 9 floats per triangle
    const uint32_t waves = cdiv(std::max<uint64_t>(std::max(b->nt, b->pairs_cap), 1), 64);
    const uint32_t waves_cap = (waves + 3) / 4 * 4 + 4;
    HR_TRY(b->wave_counts.reserve(waves_cap));
    HR_TRY(b->wave_offsets.reserve(waves_cap));
    HR_TRY(b->keys.reserve(pair_max));
    HR_TRY(b->vals.reserve(pair_max));
    HR_TRY(b->keys_sorted.reserve(pair_max));
    HR_TRY(b->prims.reserve(b->leaf_pairs_cap));
    HR_TRY(b->cells.reserve(b->cells_cap));
    HR_TRY(b->top_vals_sorted.reserve(b->pairs_cap));
    HR_TRY(b->leaf_counts.reserve(b->cells_cap));
    HR_TRY(b->leaves.reserve(b->leaves_cap));
    const size_t scan_n = std::max<size_t>({ (size_t)waves_cap, (size_t)b->cells_cap, (size_t)b->ns + 1});
    {
        // the single-pass scan's tile flags must start cleared
        const size_t before = b->scan_chain.cap;
        HR_TRY(b->scan_chain.reserve(hr::scan_chain_words((uint32_t)std::min<size_t>(scan_n, 0xFFFFFFFFu))));
        if (b->scan_chain.cap != before) hr::fill_u32(b->scan_chain.p, 0u, b->scan_chain.cap, nullptr, 1, b->stream);
    }
    HR_TRY(b->sort_keys_tmp.reserve(pair_max));
    HR_TRY(b->sort_vals_tmp.reserve(pair_max));
    HR_TRY(b->sort_hist.reserve(hr::sort_hist_words(pair_max)));
    {
        // the sort passes' tile flags must start cleared
        const size_t before = b->sort_chain.cap;
        HR_TRY(b->sort_chain.reserve(hr::sort_chain_words(pair_max)));
        if (b->sort_chain.cap != before) hr::fill_u32(b->sort_chain.p, 0u, b->sort_chain.cap, nullptr, 1, b->stream);
    }
    return HR_OK;
}

hr::BuildBuffers buffers(hr_batch* b)
{
    hr::BuildBuffers B;
    B.params = b->agg.p;
    B.status = b->status.p;
    B.rec = b->rec.p;
    B.wave_counts = b->wave_counts.p;
    B.wave_offsets = b->wave_offsets.p;
    B.keys = b->keys.p;
    B.vals = b->vals.p;
    B.keys_sorted = b->keys_sorted.p;
    B.top_vals_sorted = b->top_vals_sorted.p;
    B.prims = b->prims.p;
    B.cells = b->cells.p;
    B.leaves = b->leaves.p;
    B.leaf_counts = b->leaf_counts.p;
    B.cells_cap = b->cells_cap;
    B.pairs_cap = b->pairs_cap;
    B.leaves_cap = b->leaves_cap;
    B.leaf_pairs_cap = b->leaf_pairs_cap;
    B.waves_cap = (uint32_t)b->wave_counts.cap;
    B.scan.chain = b->scan_chain.p;
    B.scan.chain_cap = b->scan_chain.cap;
    B.sort.keys_tmp = b->sort_keys_tmp.p;
    B.sort.vals_tmp = b->sort_vals_tmp.p;
    B.sort.hist = b->sort_hist.p;
    B.sort.chain = b->sort_chain.p;
    return B;
}

hr::SceneTable table(hr_batch* b)
{
    hr::SceneTable T;
    T.ns = b->ns;
    T.tri_scene = b->tri_scene.p;
    T.tri_base = b->tbase.p;
    T.vert_base = b->vbase.p;
    T.lambda_top = b->lam_top.p;
    T.lambda_leaf = b->lam_leaf.p;
    T.gps = b->gps.p;
    T.bounds_enc = b->bounds.p;
    T.cell_counts = b->cell_counts.p;
    T.cell_base = b->cell_base.p;
    T.top_base = b->top_base.p;
    T.leaf_base = b->leaf_base.p;
    T.pair_base = b->pair_base.p;
    return T;
}

hr_status upload(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return HR_OK;
    HR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return HR_OK;
}

hr_status download(void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return HR_OK;
    HR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return HR_OK;
}

hr_status upload_densities(hr_batch* b)
{
    HR_TRY(upload(b->lam_top.p, b->h_lt.data(), b->ns * sizeof(float), b->stream));
    HR_TRY(upload(b->lam_leaf.p, b->h_ll.data(), b->ns * sizeof(float), b->stream));
    HR_HIP(hipStreamSynchronize(b->stream));
    b->built = false;
    return HR_OK;
}

hr_status do_build(hr_batch* b)
{
    if (b->ns == 0) return fail(HR_ERR_STATE, "no scenes");
    uint64_t cells_bound = 64;
    for (uint32_t s = 0; s < b->ns; ++s) {
        const double n = (double)(b->h_tbase[s + 1] - b->h_tbase[s]);
        cells_bound += std::max<uint64_t>(32768, (uint64_t)((double)b->h_lt[s] * n * 1.001) + 64);
    }
    if (cells_bound > 0xFFFFFFF0ull) return fail(HR_ERR_CAPACITY, "grid densities too high for one batch");
    if (b->cells_cap < cells_bound) b->cells_cap = cells_bound;
    if (b->pairs_cap == 0) b->pairs_cap = std::max<uint32_t>(4096, 4 * b->nt);
    if (b->leaves_cap == 0) b->leaves_cap = std::max<uint32_t>(65536, 8 * b->nt);
    if (b->leaf_pairs_cap == 0) b->leaf_pairs_cap = std::max<uint32_t>(65536, 8 * b->nt);
    b->built = false;
    hr::BuildArgs a;
    a.kind = HR_GRID_TWO_LEVEL;
    a.lambda_top = 0.f;
    a.lambda_leaf = 0.f;
    a.idx = b->idx.p;
    a.vx = b->verts.p;
    a.vy = b->verts.p + b->nv;
    a.vz = b->verts.p + 2 * (size_t)b->nv;
    a.nt = b->nt;
    const hr::SceneTable T = table(b);
    for (int attempt = 0; attempt < 16; ++attempt) {
        HR_TRY(reserve_build(b));
        hr::BuildBuffers B = buffers(b);
        HR_HIP(hipEventRecord(b->ev_b0, b->stream));
        hr::run_build_batch(a, B, T, b->nv, b->stream);
        HR_HIP(hipEventRecord(b->ev_b1, b->stream));
        HR_HIP(hipMemcpyAsync(b->h_status, b->status.p, sizeof(hr::BuildStatus), hipMemcpyDeviceToHost,
                              b->stream));
        HR_HIP(hipStreamSynchronize(b->stream));
        HR_HIP(hipGetLastError());
        const hr::BuildStatus st = *b->h_status;
        if (st.overflow == 0u) {
            b->build_timed = true;
            const size_t n1 = (size_t)b->ns + 1;
            b->h_gps.resize(b->ns);
            b->h_cell_base.resize(n1);
            b->h_top_base.resize(n1);
            b->h_leaf_base.resize(n1);
            b->h_pair_base.resize(n1);
            HR_TRY(download(b->h_gps.data(), b->gps.p, b->ns * sizeof(hr::GridParams)));
            HR_TRY(download(b->h_cell_base.data(), b->cell_base.p, n1 * sizeof(uint32_t)));
            HR_TRY(download(b->h_top_base.data(), b->top_base.p, n1 * sizeof(uint32_t)));
            HR_TRY(download(b->h_leaf_base.data(), b->leaf_base.p, n1 * sizeof(uint32_t)));
            HR_TRY(download(b->h_pair_base.data(), b->pair_base.p, n1 * sizeof(uint32_t)));
            std::vector<hr::SceneDesc> d(b->ns);
            for (uint32_t s = 0; s < b->ns; ++s) {
                d[s].S.gp = b->gps.p + s;
                d[s].S.cells = b->cells.p + b->h_cell_base[s];
                d[s].S.leaves = b->leaves.p;
                d[s].S.prims = b->prims.p;
                d[s].S.rec = b->rec.p;
                d[s].S.nt = b->nt;
                d[s].S.two_level = 1;
                d[s].tri_base = b->h_tbase[s];
                d[s].nt = b->h_tbase[s + 1] - b->h_tbase[s];
            }
            HR_TRY(b->descs.reserve(b->ns));
            HR_HIP(hipMemcpy(b->descs.p, d.data(), b->ns * sizeof(hr::SceneDesc), hipMemcpyHostToDevice));
            b->built = true;
            return HR_OK;
        }
        if (std::getenv("HR_DEBUG_BUILD"))
            std::fprintf(stderr, "hr_batch_build: attempt %d overflow 0x%x cells %u/%llu pairs %u/%u leaves %u/%u leaf pairs %u/%u\n",
                         attempt, st.overflow, st.num_cells, (unsigned long long)b->cells_cap, st.num_pairs,
                         b->pairs_cap, st.num_leaves, b->leaves_cap, st.num_leaf_pairs, b->leaf_pairs_cap);
        // counts are valid up to the first stage that overflowed: grow that one
        if (st.overflow & hr::kOverflowCells) {
            if (st.num_cells == 0xFFFFFFFFu || st.num_cells == 0u)
                return fail(HR_ERR_CAPACITY, "grid resolution too large");
            b->cells_cap = grow(b->cells_cap, st.num_cells);
        } else if (st.overflow & hr::kOverflowPairs) {
            b->pairs_cap = grow(b->pairs_cap, st.num_pairs);
        } else if (st.overflow & hr::kOverflowLeaves) {
            b->leaves_cap = grow(b->leaves_cap, st.num_leaves);
        } else if (st.overflow & hr::kOverflowLeafPairs) {
            b->leaf_pairs_cap = grow(b->leaf_pairs_cap, st.num_leaf_pairs);
        }
    }
    return fail(HR_ERR_CAPACITY, "batch build did not fit after repeated growth");
}

hr_status grow_pinned(hr_batch* b, size_t nrb, size_t npre)
{
    if (nrb > b->h_rb_cap) {
        if (b->h_rb) (void)hipHostFree(b->h_rb);
        b->h_rb = nullptr;
        b->h_rb_cap = 0;
        HR_HIP(hipHostMalloc((void**)&b->h_rb, nrb * sizeof(hr::RayBatchDev), hipHostMallocDefault));
        b->h_rb_cap = nrb;
    }
    if (npre > b->h_prefix_cap) {
        if (b->h_prefix) (void)hipHostFree(b->h_prefix);
        b->h_prefix = nullptr;
        b->h_prefix_cap = 0;
        HR_HIP(hipHostMalloc((void**)&b->h_prefix, npre * sizeof(uint32_t), hipHostMallocDefault));
        b->h_prefix_cap = npre;
    }
    return HR_OK;
}

// enqueue the closest-hit and the any-hit launch of a set of ray batches
hr_status enqueue_trace(hr_batch* b, const hr_ray_batch* rb, uint32_t nb, bool timed)
{
    if (!b->built) return fail(HR_ERR_STATE, "trace before hr_batch_build");
    if (nb > 0 && !rb) return fail(HR_ERR_ARG, "null ray batches");
    for (uint32_t i = 0; i < nb; ++i) {
        const hr_ray_batch& r = rb[i];
        if (r.scene >= b->ns) return fail(HR_ERR_ARG, "ray batch " + std::to_string(i) + ": scene out of range");
        if (r.n == 0) continue;
        if (!r.org || !r.dir) return fail(HR_ERR_ARG, "ray batch " + std::to_string(i) + ": null rays");
        if (r.anyhit ? (!r.occ_out || !(r.tmax > 0.f)) : (!r.t_out || !r.prim_out))
            return fail(HR_ERR_ARG, "ray batch " + std::to_string(i) + ": missing outputs");
    }
    // previous launches may still read the pinned tables
    HR_HIP(hipStreamSynchronize(b->stream));
    HR_TRY(grow_pinned(b, (size_t)nb, 2 * ((size_t)nb + 1)));
    uint32_t off[2] = { 0, 0 };
    uint64_t total[2] = { 0, 0 };
    // closest-hit batches first, then any-hit batches, each with its prefix
    std::vector<uint32_t> order[2];
    for (uint32_t i = 0; i < nb; ++i)
        if (rb[i].n) order[rb[i].anyhit ? 1 : 0].push_back(i);
    uint32_t* pre[2] = { b->h_prefix, b->h_prefix + order[0].size() + 1 };
    hr::RayBatchDev* dst[2] = { b->h_rb, b->h_rb + order[0].size() };
    for (int k = 0; k < 2; ++k) {
        pre[k][0] = 0;
        for (size_t q = 0; q < order[k].size(); ++q) {
            const hr_ray_batch& r = rb[order[k][q]];
            hr::RayBatchDev& d = dst[k][q];
            d.scene = r.scene;
            d.n = r.n;
            d.tmax = r.anyhit ? r.tmax : 0.f;
            d.stride = 0;
            d.org = r.org;
            d.dir = r.dir;
            d.t_out = r.t_out;
            d.prim_out = r.prim_out;
            d.occ_out = r.occ_out;
            d.hits = nullptr;
            total[k] += r.n;
            if (total[k] > 0xFFFFFFFFull) return fail(HR_ERR_CAPACITY, "more than 2^32 rays in one launch");
            pre[k][q + 1] = (uint32_t)total[k];
        }
        off[k] = (uint32_t)order[k].size();
        HR_TRY(b->rb_dev[k].reserve(std::max<size_t>(off[k], 1)));
        HR_TRY(b->prefix_dev[k].reserve((size_t)off[k] + 1));
        HR_TRY(upload(b->rb_dev[k].p, dst[k], off[k] * sizeof(hr::RayBatchDev), b->stream));
        HR_TRY(upload(b->prefix_dev[k].p, pre[k], ((size_t)off[k] + 1) * sizeof(uint32_t), b->stream));
    }
    if (timed) HR_HIP(hipEventRecord(b->ev_t0, b->stream));
    for (int k = 0; k < 2; ++k)
        hr::persist_multi(b->descs.p, b->rb_dev[k].p, b->prefix_dev[k].p, off[k], (uint32_t)total[k], k,
                          b->pool.p, (uint32_t)b->tuning.refill, (uint32_t)b->tuning.blocks_per_cu,
                          b->ray_count.p, b->work.p + 3 * k, b->stream);
    if (timed) {
        HR_HIP(hipEventRecord(b->ev_t1, b->stream));
        b->trace_timed = true;
    }
    HR_HIP(hipGetLastError());
    return HR_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.f;
}

}  // namespace

extern "C" {

hr_status hr_batch_create(hr_batch** out, const hr_options* opts)
{
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(HR_ERR_DEVICE, "no HIP device available");
    hr_batch* b = new hr_batch();
    b->device = opts ? opts->device : 0;
    if (b->device < 0 || b->device >= n) {
        delete b;
        return fail(HR_ERR_ARG, "device ordinal out of range");
    }
    hr_status s = HR_OK;
    do {
        if (hipSetDevice(b->device) != hipSuccess) { s = fail(HR_ERR_DEVICE, "hipSetDevice failed"); break; }
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            s = fail(HR_ERR_DEVICE, "hipStreamCreate failed");
            break;
        }
        if ((s = b->agg.reserve(1)) != HR_OK) break;
        if ((s = b->status.reserve(1)) != HR_OK) break;
        if ((s = b->pool.reserve(1)) != HR_OK) break;
        if ((s = b->ray_count.reserve(1)) != HR_OK) break;
        if ((s = b->work.reserve(6)) != HR_OK) break;
        if (hipHostMalloc((void**)&b->h_status, sizeof(hr::BuildStatus), hipHostMallocDefault) != hipSuccess) {
            s = fail(HR_ERR_OOM, "pinned allocation failed");
            break;
        }
        hipEvent_t* evs[4] = { &b->ev_b0, &b->ev_b1, &b->ev_t0, &b->ev_t1 };
        for (hipEvent_t* ev : evs)
            if (hipEventCreate(ev) != hipSuccess) s = fail(HR_ERR_DEVICE, "hipEventCreate failed");
        if (s != HR_OK) break;
        hr::fill_u32((uint32_t*)b->ray_count.p, 0u, 2, nullptr, 1, b->stream);
        hr::fill_u32((uint32_t*)b->work.p, 0u, 12, nullptr, 1, b->stream);
        if (hipStreamSynchronize(b->stream) != hipSuccess) s = fail(HR_ERR_DEVICE, "stream synchronize failed");
    } while (0);
    if (s != HR_OK) {
        hr_batch_destroy(b);
        return s;
    }
    *out = b;
    return HR_OK;
}

hr_status hr_batch_destroy(hr_batch* b)
{
    if (!b) return HR_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    DevBuf<float>* f[] = { &b->verts, &b->lam_top, &b->lam_leaf };
    for (auto* x : f) x->release();
    DevBuf<uint32_t>* u[] = { &b->idx, &b->tri_scene, &b->vbase, &b->tbase, &b->bounds, &b->cell_counts,
                              &b->cell_base, &b->top_base, &b->leaf_base, &b->pair_base, &b->wave_counts,
                              &b->wave_offsets, &b->scan_chain,
                              &b->keys, &b->vals, &b->keys_sorted, &b->top_vals_sorted, &b->prims,
                              &b->leaf_counts, &b->sort_chain, &b->sort_keys_tmp, &b->sort_vals_tmp,
                              &b->sort_hist, &b->prefix_dev[0], &b->prefix_dev[1], &b->pool };
    for (auto* x : u) x->release();
    b->gps.release();
    b->agg.release();
    b->status.release();
    b->rec.release();
    b->cells.release();
    b->leaves.release();
    b->descs.release();
    b->rb_dev[0].release();
    b->rb_dev[1].release();
    b->ray_count.release();
    b->work.release();
    if (b->h_status) (void)hipHostFree(b->h_status);
    if (b->h_rb) (void)hipHostFree(b->h_rb);
    if (b->h_prefix) (void)hipHostFree(b->h_prefix);
    hipEvent_t evs[4] = { b->ev_b0, b->ev_b1, b->ev_t0, b->ev_t1 };
    for (hipEvent_t ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return HR_OK;
}

hr_status hr_batch_get_stream(hr_batch* b, void** stream)
{
    HR_TRY(check_batch(b));
    if (!stream) return fail(HR_ERR_ARG, "null output pointer");
    *stream = (void*)b->stream;
    return HR_OK;
}

hr_status hr_batch_set_scenes(hr_batch* b, const hr_mesh_desc* scenes, uint32_t n)
{
    HR_TRY(check_batch(b));
    if (n == 0 || !scenes) return fail(HR_ERR_ARG, "a batch needs at least one scene");
    uint64_t nv = 0, nt = 0;
    for (uint32_t s = 0; s < n; ++s) {
        const hr_mesh_desc& m = scenes[s];
        if (m.nv == 0 || !m.xyz) return fail(HR_ERR_ARG, "scene " + std::to_string(s) + " has no vertices");
        if (m.nt > 0 && !m.idx) return fail(HR_ERR_ARG, "scene " + std::to_string(s) + ": null indices");
        for (size_t i = 0; i < 3 * (size_t)m.nt; ++i)
            if (m.idx[i] >= m.nv)
                return fail(HR_ERR_ARG, "scene " + std::to_string(s) + ": vertex index out of range");
        nv += m.nv;
        nt += m.nt;
    }
    if (nv > 0xFFFFFFF0ull || nt > 0x3FFFFFFFull) return fail(HR_ERR_CAPACITY, "batch too large");
    b->ns = n;
    b->nv = (uint32_t)nv;
    b->nt = (uint32_t)nt;
    b->h_vbase.assign((size_t)n + 1, 0);
    b->h_tbase.assign((size_t)n + 1, 0);
    for (uint32_t s = 0; s < n; ++s) {
        b->h_vbase[s + 1] = b->h_vbase[s] + scenes[s].nv;
        b->h_tbase[s + 1] = b->h_tbase[s] + scenes[s].nt;
    }
    std::vector<float> v(3 * (size_t)b->nv);
    std::vector<uint32_t> ix(3 * (size_t)b->nt), ts(b->nt);
    for (uint32_t s = 0; s < n; ++s) {
        const hr_mesh_desc& m = scenes[s];
        const uint32_t vb = b->h_vbase[s], tb = b->h_tbase[s];
        for (int k = 0; k < 3; ++k)
            std::memcpy(v.data() + (size_t)k * b->nv + vb, m.xyz + (size_t)k * m.nv, m.nv * sizeof(float));
        for (size_t i = 0; i < 3 * (size_t)m.nt; ++i) ix[3 * (size_t)tb + i] = m.idx[i] + vb;
        std::fill(ts.begin() + tb, ts.begin() + tb + m.nt, s);
    }
    HR_TRY(b->verts.reserve(v.size()));
    HR_TRY(b->idx.reserve(std::max<size_t>(ix.size(), 1)));
    HR_TRY(b->tri_scene.reserve(std::max<size_t>(ts.size(), 1)));
    HR_TRY(b->vbase.reserve((size_t)n + 1));
    HR_TRY(b->tbase.reserve((size_t)n + 1));
    HR_TRY(b->lam_top.reserve(n));
    HR_TRY(b->lam_leaf.reserve(n));
    HR_TRY(b->gps.reserve(n));
    HR_TRY(b->bounds.reserve(6 * (size_t)n));
    HR_TRY(b->cell_counts.reserve(n));
    HR_TRY(b->cell_base.reserve((size_t)n + 1));
    HR_TRY(b->top_base.reserve((size_t)n + 1));
    HR_TRY(b->leaf_base.reserve((size_t)n + 1));
    HR_TRY(b->pair_base.reserve((size_t)n + 1));
    HR_TRY(upload(b->verts.p, v.data(), v.size() * sizeof(float), b->stream));
    HR_TRY(upload(b->idx.p, ix.data(), ix.size() * sizeof(uint32_t), b->stream));
    HR_TRY(upload(b->tri_scene.p, ts.data(), ts.size() * sizeof(uint32_t), b->stream));
    HR_TRY(upload(b->vbase.p, b->h_vbase.data(), ((size_t)n + 1) * sizeof(uint32_t), b->stream));
    HR_TRY(upload(b->tbase.p, b->h_tbase.data(), ((size_t)n + 1) * sizeof(uint32_t), b->stream));
    HR_HIP(hipStreamSynchronize(b->stream));
    b->h_lt.assign(n, 1.2f);       // src/RT/RTEngine.cu:79-80
    b->h_ll.assign(n, 2.2f);
    b->pairs_cap = b->leaves_cap = b->leaf_pairs_cap = 0;
    return upload_densities(b);
}

hr_status hr_batch_set_densities(hr_batch* b, const float* lambda_top, const float* lambda_leaf)
{
    HR_TRY(check_batch(b));
    if (b->ns == 0) return fail(HR_ERR_STATE, "no scenes");
    for (uint32_t s = 0; s < b->ns; ++s) {
        const float lt = lambda_top ? lambda_top[s] : 1.2f;
        const float ll = lambda_leaf ? lambda_leaf[s] : 2.2f;
        if (!(lt > 0.f && lt < 1e6f) || !(ll > 0.f && ll < 1e6f))
            return fail(HR_ERR_ARG, "density of scene " + std::to_string(s) + " out of range");
        b->h_lt[s] = lt;
        b->h_ll[s] = ll;
    }
    return upload_densities(b);
}

hr_status hr_batch_build(hr_batch* b)
{
    HR_TRY(check_batch(b));
    return do_build(b);
}

hr_status hr_batch_get_grid_info(hr_batch* b, uint32_t s, hr_grid_info* out)
{
    HR_TRY(check_batch(b));
    if (!b->built) return fail(HR_ERR_STATE, "no batch built");
    if (s >= b->ns) return fail(HR_ERR_ARG, "scene out of range");
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    const hr::GridParams& g = b->h_gps[s];
    out->kind = HR_GRID_TWO_LEVEL;
    for (int i = 0; i < 3; ++i) {
        out->bmin[i] = g.bmin[i];
        out->bmax[i] = g.bmax[i];
        out->res[i] = g.res[i];
        out->cell_size[i] = g.cs[i];
        out->cell_size_rcp[i] = g.csr[i];
    }
    out->num_cells = b->h_cell_base[s + 1] - b->h_cell_base[s];
    out->num_leaves = b->h_leaf_base[s + 1] - b->h_leaf_base[s];
    out->num_refs = b->h_pair_base[s + 1] - b->h_pair_base[s];
    out->num_top_pairs = b->h_top_base[s + 1] - b->h_top_base[s];
    out->num_triangles = b->h_tbase[s + 1] - b->h_tbase[s];
    out->num_objects = 0;
    out->num_instances = 0;
    return HR_OK;
}

hr_status hr_batch_get_grid(hr_batch* b, uint32_t s, uint32_t* cells, uint32_t* leaves, uint32_t* prims)
{
    hr_grid_info in;
    HR_TRY(hr_batch_get_grid_info(b, s, &in));
    HR_HIP(hipStreamSynchronize(b->stream));
    const uint32_t lb = b->h_leaf_base[s], pb = b->h_pair_base[s], tb = b->h_tbase[s];
    if (cells) {
        HR_TRY(download(cells, b->cells.p + b->h_cell_base[s], 2 * (size_t)in.num_cells * sizeof(uint32_t)));
        for (uint32_t c = 0; c < in.num_cells; ++c) cells[2 * (size_t)c] -= lb;
    }
    if (leaves) {
        HR_TRY(download(leaves, b->leaves.p + lb, 2 * (size_t)in.num_leaves * sizeof(uint32_t)));
        for (uint32_t l = 0; l < in.num_leaves; ++l)
            if (leaves[2 * (size_t)l + 1] > leaves[2 * (size_t)l]) {
                leaves[2 * (size_t)l] -= pb;
                leaves[2 * (size_t)l + 1] -= pb;
            }
    }
    if (prims) {
        HR_TRY(download(prims, b->prims.p + pb, (size_t)in.num_refs * sizeof(uint32_t)));
        for (uint32_t i = 0; i < in.num_refs; ++i) prims[i] -= tb;
    }
    return HR_OK;
}

hr_status hr_batch_trace_dev(hr_batch* b, const hr_ray_batch* rb, uint32_t nb)
{
    HR_TRY(check_batch(b));
    return enqueue_trace(b, rb, nb, true);
}

hr_status hr_batch_set_tuning(hr_batch* b, const hr_batch_tuning* t)
{
    HR_TRY(check_batch(b));
    if (!t) return fail(HR_ERR_ARG, "null tuning");
    if (t->refill < 1 || t->refill > 64) return fail(HR_ERR_ARG, "refill must be in [1, 64]");
    if (t->blocks_per_cu < 0 || t->blocks_per_cu > 32) return fail(HR_ERR_ARG, "blocks_per_cu must be in [0, 32]");
    b->tuning = *t;
    return HR_OK;
}

hr_status hr_batch_get_tuning(hr_batch* b, hr_batch_tuning* t, float* lambda_top, float* lambda_leaf)
{
    HR_TRY(check_batch(b));
    if (t) *t = b->tuning;
    for (uint32_t s = 0; s < b->ns; ++s) {
        if (lambda_top) lambda_top[s] = b->h_lt[s];
        if (lambda_leaf) lambda_leaf[s] = b->h_ll[s];
    }
    return HR_OK;
}

hr_status hr_batch_get_timing(hr_batch* b, double* build_ms, double* trace_ms, uint64_t* rays_traced)
{
    HR_TRY(check_batch(b));
    HR_HIP(hipStreamSynchronize(b->stream));
    if (build_ms) *build_ms = b->build_timed ? elapsed(b->ev_b0, b->ev_b1) : 0.0;
    if (trace_ms) *trace_ms = b->trace_timed ? elapsed(b->ev_t0, b->ev_t1) : 0.0;
    if (rays_traced) {
        unsigned long long r = 0;
        HR_TRY(download(&r, b->ray_count.p, sizeof r));
        *rays_traced = r;
    }
    return HR_OK;
}

hr_status hr_batch_get_work_stats(hr_batch* b, hr_work_stats* out)
{
    HR_TRY(check_batch(b));
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    unsigned long long w[6];
    HR_HIP(hipStreamSynchronize(b->stream));
    HR_TRY(download(w, b->work.p, sizeof w));
    hr::fill_u32((uint32_t*)b->work.p, 0u, 12, nullptr, 1, b->stream);
    HR_HIP(hipStreamSynchronize(b->stream));
    out->primary.rays = w[0];
    out->primary.steps = w[1];
    out->primary.tests = w[2];
    out->secondary.rays = w[3];
    out->secondary.steps = w[4];
    out->secondary.tests = w[5];
    return HR_OK;
}

hr_status hr_batch_synchronize(hr_batch* b)
{
    HR_TRY(check_batch(b));
    HR_HIP(hipStreamSynchronize(b->stream));
    return HR_OK;
}

// Autotuner.  Scenes fall into size classes by triangle count (< 4K, < 16K,
// < 64K, >= 64K).  Starting from the reference's densities (1.2, 2.2), each
// class in turn tries every (top, leaf) candidate while the other classes
// keep their current choice; a candidate is scored by the time of a full
// batch build plus the trace of all given rays (the minimum of two runs), and
// the class keeps the fastest.  Then refill and blocks per CU are chosen the
// same way on the trace alone.  The chosen values stay set on the batch.
hr_status hr_batch_autotune(hr_batch* b, const hr_ray_batch* rb, uint32_t nb, hr_autotune_result* out)
{
    HR_TRY(check_batch(b));
    if (b->ns == 0) return fail(HR_ERR_STATE, "no scenes");
    static const float kTop[] = { 1.0f / 16, 0.125f, 0.25f, 0.5f, 1.2f };
    static const float kLeaf[] = { 1.0f, 2.2f, 3.0f, 5.0f };
    static const int kRefill[] = { 16, 32, 48, 64 };
    static const int kBlocks[] = { 0, 2, 4, 8 };
    const int nclass = 4;
    auto class_of = [&](uint32_t s) {
        const uint32_t n = b->h_tbase[s + 1] - b->h_tbase[s];
        return n < 4096 ? 0 : n < 16384 ? 1 : n < 65536 ? 2 : 3;
    };
    std::vector<float> lt(nclass, 1.2f), ll(nclass, 2.2f);
    auto apply = [&]() -> hr_status {
        for (uint32_t s = 0; s < b->ns; ++s) {
            b->h_lt[s] = lt[class_of(s)];
            b->h_ll[s] = ll[class_of(s)];
        }
        return upload_densities(b);
    };
    auto score = [&](bool with_build, float& ms) -> hr_status {
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            float t = 0.f;
            if (with_build) {
                HR_TRY(do_build(b));
                t += elapsed(b->ev_b0, b->ev_b1);
            }
            HR_TRY(enqueue_trace(b, rb, nb, true));
            HR_HIP(hipStreamSynchronize(b->stream));
            t += elapsed(b->ev_t0, b->ev_t1);
            best = std::min(best, t);
        }
        ms = best;
        return HR_OK;
    };
    int evaluations = 0;
    float best_ms = 0.f;
    HR_TRY(apply());
    HR_TRY(score(true, best_ms));
    ++evaluations;
    for (int c = 0; c < nclass; ++c) {
        bool present = false;
        for (uint32_t s = 0; s < b->ns && !present; ++s) present = class_of(s) == c;
        if (!present) continue;
        float bt = lt[c], bl = ll[c];
        for (float t : kTop)
            for (float l : kLeaf) {
                if (t == bt && l == bl) continue;
                lt[c] = t;
                ll[c] = l;
                HR_TRY(apply());
                float ms = 0.f;
                HR_TRY(score(true, ms));
                ++evaluations;
                if (ms < best_ms) {
                    best_ms = ms;
                    bt = t;
                    bl = l;
                }
            }
        lt[c] = bt;
        ll[c] = bl;
    }
    HR_TRY(apply());
    HR_TRY(do_build(b));
    float trace_best = 1e30f;
    hr_batch_tuning bestt = b->tuning;
    for (int r : kRefill)
        for (int k : kBlocks) {
            b->tuning.refill = r;
            b->tuning.blocks_per_cu = k;
            float ms = 0.f;
            HR_TRY(score(false, ms));
            ++evaluations;
            if (ms < trace_best) {
                trace_best = ms;
                bestt = b->tuning;
            }
        }
    b->tuning = bestt;
    if (out) {
        std::memset(out, 0, sizeof *out);
        for (int c = 0; c < nclass; ++c) {
            out->lambda_top[c] = lt[c];
            out->lambda_leaf[c] = ll[c];
        }
        out->tuning = bestt;
        out->evaluations = evaluations;
        out->best_build_trace_ms = best_ms;
        out->best_trace_ms = trace_best;
    }
    return HR_OK;
}

}  // extern "C"
