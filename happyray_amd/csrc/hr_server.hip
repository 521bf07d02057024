// hr_server.hip -- a request server over resident scenes (no reference
// counterpart: the reference renders one scene interactively).
//
// Clients submit trace, occlusion, path and rebuild requests from any thread.
// One worker thread runs continuous batching: while a wave of requests runs
// on the GPU, requests that arrive are admitted into the next wave as soon as
// a wave slot and device memory are free.  A wave traces the rays of all its
// requests, of any scenes, in one closest-hit and one any-hit persistent
// launch; path requests continue from their closest hits with the path
// tracer.  Scenes become resident (grid built) on first use and are evicted
// least-recently-used when the memory budget requires it; a rebuild request
// builds the scene's new frame into a second grid while earlier waves may
// still read the first, then swaps them.  A request's result depends only on
// its own rays and its scene's frame, never on what it was batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/happyray.h"
#include "hr_ctx_internal.h"
#include "hr_host.h"
#include "hr_persist.h"

using hr::DevBuf;
using hr::fail;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
    uint64_t id = 0;
    hr_request r{};
    Clock::time_point t_submit, t_done;
    hr_status status = HR_OK;
    std::string error;
    bool done = false;
    uint64_t wave = 0;
};

struct Scene {
    std::vector<float> keys;          // nframes blocks of SoA xyz
    uint32_t nframes = 0, nv = 0, nt = 0;
    std::vector<uint32_t> idx, mat_idx;
    std::vector<float> mats, weights;
    std::vector<hr_area_light> lights;
    float lt = 1.2f, ll = 2.2f;
    uint32_t k1 = 0, k2 = 0;
    float coeff = 0.f;
    hr_ctx* active = nullptr;         // resident grid
    hr_ctx* spare = nullptr;          // second grid for rebuilds
    uint64_t active_bytes = 0, spare_bytes = 0;
    uint64_t last_use = 0;            // wave stamp (LRU)
    uint64_t active_wave = 0;         // last wave that reads `active`
    uint64_t spare_wave = 0;          // last wave that reads `spare`
};

struct Slot {
    bool in_flight = false;
    uint64_t id = 0;
    std::vector<Req*> reqs;
    DevBuf<unsigned char> arena;
    DevBuf<hr::RayBatchDev> rb[2];
    DevBuf<uint32_t> prefix[2];
    DevBuf<hr::SceneDesc> desc;
    hr::RayBatchDev* h_rb = nullptr;
    uint32_t* h_prefix = nullptr;
    hr::SceneDesc* h_desc = nullptr;
    size_t h_rb_cap = 0, h_prefix_cap = 0, h_desc_cap = 0;
    hipEvent_t done = nullptr;
};

template <class T>
hr_status grow_pinned(T*& p, size_t& cap, size_t n)
{
    if (n <= cap && p) return HR_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    HR_HIP(hipHostMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
    cap = std::max<size_t>(n, 1);
    return HR_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct hr_server {
    hr_server_options opt{};
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    std::vector<std::unique_ptr<Scene>> scenes;   // appended under mu
    std::vector<Scene*> snap;                     // the worker's view of scenes
    uint32_t resident_count = 0;                  // written under mu
    uint64_t next_id = 1;
    uint64_t wave_counter = 0;
    uint64_t forming = 0;                         // id of the wave being formed
    Slot slots[2];
    std::thread worker;
    bool stop = false;
    uint64_t resident_bytes = 0, peak_bytes = 0;
    hr_server_stats stats{};
    DevBuf<uint32_t> pool;
    DevBuf<unsigned long long> ray_count;
    std::string fatal;
};

namespace {

uint64_t scene_estimate(const Scene& s)
{
    return (uint64_t)s.keys.size() * 4 + (uint64_t)s.nt * 420 + (uint64_t)s.nv * 12 + (4u << 20);
}

void destroy_ctx(hr_server* S, hr_ctx*& c, uint64_t& bytes)
{
    if (!c) return;
    hr_destroy(c);
    c = nullptr;
    S->resident_bytes -= std::min(S->resident_bytes, bytes);
    bytes = 0;
}

void count_resident(hr_server* S)
{
    uint32_t n = 0;
    for (Scene* s : S->snap) n += s->active ? 1u : 0u;
    std::lock_guard<std::mutex> lk(S->mu);
    S->resident_count = n;
}

// finish the waves whose events completed; with block, wait for the oldest
void reap(hr_server* S, bool block)
{
    for (;;) {
        Slot* oldest = nullptr;
        for (Slot& sl : S->slots)
            if (sl.in_flight && (!oldest || sl.id < oldest->id)) oldest = &sl;
        if (!oldest) return;
        if (block) {
            (void)hipEventSynchronize(oldest->done);
            block = false;
        } else if (hipEventQuery(oldest->done) != hipSuccess) {
            return;
        }
        const Clock::time_point now = Clock::now();
        std::lock_guard<std::mutex> lk(S->mu);
        for (Req* r : oldest->reqs) {
            r->done = true;
            r->t_done = now;
            S->stats.requests += 1;
            S->stats.rays += r->r.n;
        }
        oldest->reqs.clear();
        oldest->in_flight = false;
        S->cv_done.notify_all();
    }
}

// true while wave `id` is being formed or runs on the GPU
bool wave_busy(hr_server* S, uint64_t id)
{
    if (id == 0) return false;
    if (id == S->forming) return true;
    for (Slot& sl : S->slots)
        if (sl.in_flight && sl.id == id) return true;
    return false;
}

// wait until wave `id` has finished on the GPU (not for the wave being formed)
void wait_wave(hr_server* S, uint64_t id)
{
    for (;;) {
        bool running = false;
        for (Slot& sl : S->slots) running = running || (sl.in_flight && sl.id == id);
        if (!running) return;
        reap(S, true);
    }
}

void wait_all(hr_server* S)
{
    for (;;) {
        bool any = false;
        for (Slot& sl : S->slots) any = any || sl.in_flight;
        if (!any) return;
        reap(S, true);
    }
}

hr_status build_scene_ctx(hr_server* S, Scene& sc, hr_ctx** out)
{
    hr_options o{ S->opt.device, S->opt.seed_lo, S->opt.seed_hi };
    hr_ctx* c = nullptr;
    HR_TRY(hr_create(&c, &o));
    hr_status st = hr_set_keyframes(c, sc.keys.data(), sc.nframes, sc.nv, sc.idx.data(), sc.nt);
    if (st == HR_OK && !sc.mats.empty())
        st = hr_upload_materials(c, sc.mats.data(), (uint32_t)(sc.mats.size() / 12),
                                 sc.mat_idx.empty() ? nullptr : sc.mat_idx.data(), sc.nt);
    if (st == HR_OK && !sc.lights.empty())
        st = hr_set_lights(c, sc.lights.data(), sc.weights.empty() ? nullptr : sc.weights.data(),
                           (int)sc.lights.size());
    if (st == HR_OK) st = hr_set_grid(c, HR_GRID_TWO_LEVEL, sc.lt, sc.ll);
    if (st == HR_OK) st = hr_set_frame(c, sc.k1, sc.k2, sc.coeff);
    if (st == HR_OK) st = hr_build(c);
    if (st != HR_OK) {
        const std::string e = hr_last_error();
        hr_destroy(c);
        return fail(st, e);
    }
    S->stats.builds += 1;
    *out = c;
    return HR_OK;
}

// evict least-recently-used grids that no wave in flight reads until `need`
// more bytes fit the budget
bool make_room(hr_server* S, uint64_t need, const Scene* keep)
{
    for (int round = 0; round < 2; ++round) {
        while (S->resident_bytes + need > S->opt.memory_budget) {
            Scene* victim = nullptr;
            for (Scene* s : S->snap) {
                if (s == keep || (!s->active && !s->spare)) continue;
                if (wave_busy(S, s->active_wave) || wave_busy(S, s->spare_wave)) continue;
                if (!victim || s->last_use < victim->last_use) victim = s;
            }
            if (!victim) break;
            destroy_ctx(S, victim->spare, victim->spare_bytes);
            destroy_ctx(S, victim->active, victim->active_bytes);
            S->stats.evictions += 1;
            count_resident(S);
        }
        if (S->resident_bytes + need <= S->opt.memory_budget) return true;
        // grids read by waves in flight: wait for them, then try again
        wait_all(S);
    }
    return S->resident_bytes + need <= S->opt.memory_budget;
}

hr_status ensure_resident(hr_server* S, Scene& sc)
{
    if (sc.active) return HR_OK;
    const uint64_t est = scene_estimate(sc);
    if (!make_room(S, est, &sc)) return fail(HR_ERR_BUDGET, "scene does not fit the memory budget");
    HR_TRY(build_scene_ctx(S, sc, &sc.active));
    count_resident(S);
    sc.active_bytes = hr::ctx_device_bytes(sc.active);
    S->resident_bytes += sc.active_bytes;
    S->peak_bytes = std::max(S->peak_bytes, S->resident_bytes);
    return HR_OK;
}

// build frame (k1, k2, c) into the spare grid, then make it the active one
hr_status rebuild(hr_server* S, Scene& sc, const hr_request& r)
{
    if (r.k1 >= sc.nframes || r.k2 >= sc.nframes) return fail(HR_ERR_ARG, "key frame index out of range");
    sc.k1 = r.k1;
    sc.k2 = r.k2;
    sc.coeff = r.coeff;
    if (!sc.active) return HR_OK;        // built on first use with the new frame
    wait_wave(S, sc.spare_wave);
    if (!sc.spare) {
        if (!make_room(S, scene_estimate(sc), &sc)) return fail(HR_ERR_BUDGET, "no memory for a second grid");
        HR_TRY(build_scene_ctx(S, sc, &sc.spare));
        sc.spare_bytes = hr::ctx_device_bytes(sc.spare);
        S->resident_bytes += sc.spare_bytes;
    } else {
        HR_TRY(hr_set_frame(sc.spare, sc.k1, sc.k2, sc.coeff));
        HR_TRY(hr_build(sc.spare));
        S->stats.builds += 1;
        const uint64_t b = hr::ctx_device_bytes(sc.spare);
        S->resident_bytes = S->resident_bytes - std::min(S->resident_bytes, sc.spare_bytes) + b;
        sc.spare_bytes = b;
    }
    S->peak_bytes = std::max(S->peak_bytes, S->resident_bytes);
    std::swap(sc.active, sc.spare);
    std::swap(sc.active_bytes, sc.spare_bytes);
    std::swap(sc.active_wave, sc.spare_wave);
    S->stats.rebuilds += 1;
    return HR_OK;
}

void finish_now(hr_server* S, Req* r, hr_status st, const std::string& err)
{
    std::lock_guard<std::mutex> lk(S->mu);
    r->status = st;
    r->error = err;
    r->done = true;
    r->t_done = Clock::now();
    S->stats.requests += 1;
    S->cv_done.notify_all();
}

// bytes of device scratch a request needs in its wave
size_t scratch_bytes(const hr_request& r)
{
    const size_t n = r.n;
    size_t b = 0;
    if (r.layout == HR_RAYS_RECORDS)                                  // rays: planes (+ staged records)
        b += 2 * align256(12 * n) + (r.on_device ? 0 : align256(24 * n));
    else if (!r.on_device)
        b += 2 * align256(12 * n);
    if (r.kind == HR_REQ_OCCLUDED) {
        if (!r.on_device || !r.occ_out) b += align256(n);
    } else {
        if (!r.on_device || !r.t_out) b += align256(4 * n);
        if (!r.on_device || !r.prim_out || r.kind == HR_REQ_PATH) b += align256(4 * n);
        if (r.kind == HR_REQ_PATH) b += align256(28 * n) + (r.on_device && r.rgb_out ? 0 : align256(12 * n));
    }
    return b;
}

// form one wave from `list` in slot `sl` and enqueue it
void launch_wave(hr_server* S, Slot& sl, std::vector<Req*>& list)
{
    sl.id = ++S->wave_counter;
    S->forming = sl.id;
    std::vector<Req*> traced;
    std::vector<hr_ctx*> grid;           // the grid each traced request reads
    size_t scratch = 0;
    for (size_t li = 0; li < list.size(); ++li) {
        Req* q = list[li];
        hr_request& r = q->r;
        if (r.scene >= S->snap.size()) {
            finish_now(S, q, fail(HR_ERR_ARG, "scene out of range"), "scene out of range");
            continue;
        }
        Scene& sc = *S->snap[r.scene];
        if (r.kind == HR_REQ_REBUILD) {
            if (sc.spare && sc.spare_wave == sl.id) {
                // this wave still reads the grid the rebuild would overwrite:
                // the rebuild and everything after it go to the next wave
                std::lock_guard<std::mutex> lk(S->mu);
                for (size_t k = list.size(); k-- > li;) S->queue.push_front(list[k]);
                break;
            }
            const hr_status st = rebuild(S, sc, r);
            finish_now(S, q, st, st == HR_OK ? "" : hr_last_error());
            continue;
        }
        if (r.n == 0) {
            finish_now(S, q, HR_OK, "");
            continue;
        }
        const hr_status st = ensure_resident(S, sc);
        if (st != HR_OK) {
            finish_now(S, q, st, hr_last_error());
            continue;
        }
        sc.last_use = sl.id;
        sc.active_wave = sl.id;
        q->wave = sl.id;
        scratch += scratch_bytes(r);
        traced.push_back(q);
        grid.push_back(sc.active);
    }
    sl.reqs = traced;
    if (traced.empty()) {
        sl.in_flight = false;
        S->forming = 0;
        return;
    }
    // device scratch of the wave (counted against the budget while it runs)
    if (sl.arena.cap < scratch) {
        S->resident_bytes -= std::min(S->resident_bytes, (uint64_t)sl.arena.bytes());
        sl.arena.release();
        make_room(S, scratch, nullptr);
        if (sl.arena.reserve(scratch) != HR_OK) {
            for (Req* q : traced) finish_now(S, q, HR_ERR_OOM, "wave scratch allocation failed");
            sl.reqs.clear();
            S->forming = 0;
            return;
        }
        S->resident_bytes += sl.arena.bytes();
        S->peak_bytes = std::max(S->peak_bytes, S->resident_bytes);
    }
    unsigned char* base = sl.arena.p;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += align256(bytes);
        return p;
    };
    const size_t nq = traced.size();
    (void)grow_pinned(sl.h_rb, sl.h_rb_cap, nq);
    (void)grow_pinned(sl.h_prefix, sl.h_prefix_cap, 2 * (nq + 1));
    (void)grow_pinned(sl.h_desc, sl.h_desc_cap, nq);
    std::unordered_map<hr_ctx*, uint32_t> desc_of;     // grid -> descriptor index
    uint32_t ndesc = 0;
    std::vector<uint32_t> order[2];
    struct Out {
        Req* q;
        const void* src;
        void* dst;
        size_t bytes;
    };
    std::vector<Out> copies;
    std::vector<std::pair<Req*, std::pair<float*, uint32_t*>>> paths;   // hits, prims
    std::vector<hr_ctx*> path_grid;
    std::vector<float*> path_rgb;
    std::vector<hr::RayBatchDev> rbs(nq);
    for (size_t i = 0; i < nq; ++i) {
        Req* q = traced[i];
        const hr_request& r = q->r;
        Scene& sc = *S->snap[r.scene];
        auto it = desc_of.find(grid[i]);
        if (it == desc_of.end()) {
            hr::SceneDesc d;
            d.S = hr::ctx_trace_scene(grid[i]);
            d.tri_base = 0;
            d.nt = sc.nt;
            sl.h_desc[ndesc] = d;
            it = desc_of.emplace(grid[i], ndesc++).first;
        }
        const size_t n = r.n;
        const size_t stride = r.stride ? r.stride : n;
        const float* org = r.org;
        const float* dir = r.dir;
        uint32_t dev_stride = (uint32_t)stride;
        if (r.layout == HR_RAYS_RECORDS) {
            // 24-byte ray records: deinterleave on the device into SoA planes
            const float* rec = r.org;
            if (!r.on_device) {
                float* w = (float*)take(24 * n);
                (void)hipMemcpyAsync(w, r.org, 24 * n, hipMemcpyHostToDevice, S->stream);
                rec = w;
            }
            float* o = (float*)take(12 * n);
            float* d = (float*)take(12 * n);
            hr::launch_wire_to_soa(rec, (uint32_t)n, o, d, (uint32_t)n, S->stream);
            org = o;
            dir = d;
            dev_stride = (uint32_t)n;
        } else if (!r.on_device) {
            float* o = (float*)take(12 * n);
            float* d = (float*)take(12 * n);
            for (int k = 0; k < 3; ++k) {
                (void)hipMemcpyAsync(o + k * n, r.org + k * stride, 4 * n, hipMemcpyHostToDevice, S->stream);
                (void)hipMemcpyAsync(d + k * n, r.dir + k * stride, 4 * n, hipMemcpyHostToDevice, S->stream);
            }
            org = o;
            dir = d;
            dev_stride = (uint32_t)n;
        }
        hr::RayBatchDev& b = rbs[i];
        std::memset(&b, 0, sizeof b);
        b.scene = it->second;
        b.n = r.n;
        b.stride = dev_stride;
        b.org = org;
        b.dir = dir;
        if (r.kind == HR_REQ_OCCLUDED) {
            b.tmax = r.tmax;
            b.occ_out = (r.on_device && r.occ_out) ? r.occ_out : (uint8_t*)take(n);
            if (!r.on_device && r.occ_out) copies.push_back({ q, b.occ_out, r.occ_out, n });
            order[1].push_back((uint32_t)i);
        } else {
            b.t_out = (r.on_device && r.t_out) ? r.t_out : (float*)take(4 * n);
            const bool own_prims = !r.on_device || !r.prim_out || r.kind == HR_REQ_PATH;
            b.prim_out = own_prims ? (uint32_t*)take(4 * n) : r.prim_out;
            if (!r.on_device && r.t_out) copies.push_back({ q, b.t_out, r.t_out, 4 * n });
            if (r.prim_out && own_prims)
                copies.push_back({ q, b.prim_out, r.prim_out, 4 * n });
            if (r.kind == HR_REQ_PATH) {
                b.hits = (float*)take(28 * n);
                float* rgb = (r.on_device && r.rgb_out) ? r.rgb_out : (float*)take(12 * n);
                paths.push_back({ q, { b.hits, b.prim_out } });
                path_grid.push_back(grid[i]);
                path_rgb.push_back(rgb);
                if (!r.on_device && r.rgb_out) copies.push_back({ q, rgb, r.rgb_out, 12 * n });
            }
            order[0].push_back((uint32_t)i);
        }
    }
    // tables and the two persistent launches
    (void)sl.desc.reserve(std::max<uint32_t>(ndesc, 1));
    (void)hipMemcpyAsync(sl.desc.p, sl.h_desc, ndesc * sizeof(hr::SceneDesc), hipMemcpyHostToDevice, S->stream);
    uint32_t* pre[2] = { sl.h_prefix, sl.h_prefix + order[0].size() + 1 };
    hr::RayBatchDev* dst[2] = { sl.h_rb, sl.h_rb + order[0].size() };
    uint64_t total[2] = { 0, 0 };
    for (int k = 0; k < 2; ++k) {
        pre[k][0] = 0;
        for (size_t j = 0; j < order[k].size(); ++j) {
            dst[k][j] = rbs[order[k][j]];
            total[k] += rbs[order[k][j]].n;
            pre[k][j + 1] = (uint32_t)std::min<uint64_t>(total[k], 0xFFFFFFFFull);
        }
        (void)sl.rb[k].reserve(std::max<size_t>(order[k].size(), 1));
        (void)sl.prefix[k].reserve(order[k].size() + 1);
        if (!order[k].empty()) {
            (void)hipMemcpyAsync(sl.rb[k].p, dst[k], order[k].size() * sizeof(hr::RayBatchDev),
                                 hipMemcpyHostToDevice, S->stream);
            (void)hipMemcpyAsync(sl.prefix[k].p, pre[k], (order[k].size() + 1) * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, S->stream);
        }
        hr::persist_multi(sl.desc.p, sl.rb[k].p, sl.prefix[k].p, (uint32_t)order[k].size(),
                          (uint32_t)std::min<uint64_t>(total[k], 0xFFFFFFFFull), k, S->pool.p, 48, 0,
                          S->ray_count.p, nullptr, S->stream);
    }
    for (size_t i = 0; i < paths.size(); ++i) {
        Req* q = paths[i].first;
        const hr_status st = hr::ctx_path_trace(path_grid[i], paths[i].second.first, paths[i].second.second,
                                                q->r.n, q->r.depth, path_rgb[i], S->stream);
        if (st != HR_OK) {
            q->status = st;
            q->error = hr_last_error();
        }
    }
    for (const Out& c : copies) (void)hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, S->stream);
    (void)hipEventRecord(sl.done, S->stream);
    sl.in_flight = true;
    S->forming = 0;
    S->stats.waves += 1;
}

void worker_main(hr_server* S)
{
    if (hipSetDevice(S->opt.device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->fatal = "hipSetDevice failed in the server thread";
        return;
    }
    for (;;) {
        reap(S, false);
        Slot* free_slot = nullptr;
        bool any_in_flight = false;
        for (Slot& sl : S->slots) {
            if (!sl.in_flight && !free_slot) free_slot = &sl;
            any_in_flight = any_in_flight || sl.in_flight;
        }
        std::vector<Req*> list;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->queue.empty()) {
                if (S->stop && !any_in_flight) return;
                // wake for new requests, or to poll the waves in flight
                S->cv_work.wait_for(lk, std::chrono::microseconds(any_in_flight ? 50 : 2000));
                if (S->queue.empty()) continue;
            }
            // scenes added before these requests were submitted
            if (S->snap.size() != S->scenes.size()) {
                S->snap.clear();
                for (auto& up : S->scenes) S->snap.push_back(up.get());
            }
            if (!free_slot) {
                lk.unlock();
                reap(S, true);
                continue;
            }
            uint64_t rays = 0;
            while (!S->queue.empty()) {
                Req* q = S->queue.front();
                if (!list.empty() && rays + q->r.n > S->opt.max_wave_rays) break;
                rays += q->r.n;
                list.push_back(q);
                S->queue.pop_front();
            }
        }
        launch_wave(S, *free_slot, list);
    }
}

}  // namespace

extern "C" {

hr_status hr_server_create(hr_server** out, const hr_server_options* opt)
{
    if (!out) return fail(HR_ERR_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(HR_ERR_DEVICE, "no HIP device available");
    hr_server* S = new hr_server();
    if (opt) S->opt = *opt;
    if (S->opt.memory_budget == 0) S->opt.memory_budget = 200ull << 30;
    if (S->opt.max_wave_rays == 0) S->opt.max_wave_rays = 16u << 20;
    if (S->opt.device < 0 || S->opt.device >= n) {
        delete S;
        return fail(HR_ERR_ARG, "device ordinal out of range");
    }
    hr_status st = HR_OK;
    do {
        if (hipSetDevice(S->opt.device) != hipSuccess) { st = fail(HR_ERR_DEVICE, "hipSetDevice failed"); break; }
        if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
            st = fail(HR_ERR_DEVICE, "hipStreamCreate failed");
            break;
        }
        if ((st = S->pool.reserve(1)) != HR_OK) break;
        if ((st = S->ray_count.reserve(1)) != HR_OK) break;
        for (Slot& sl : S->slots)
            if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
                st = fail(HR_ERR_DEVICE, "hipEventCreate failed");
    } while (0);
    if (st != HR_OK) {
        hr_server_destroy(S);
        return st;
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return HR_OK;
}

hr_status hr_server_destroy(hr_server* S)
{
    if (!S) return HR_OK;
    if (S->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv_work.notify_all();
        S->worker.join();
    }
    (void)hipSetDevice(S->opt.device);
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    for (auto& up : S->scenes) {
        destroy_ctx(S, up->spare, up->spare_bytes);
        destroy_ctx(S, up->active, up->active_bytes);
    }
    for (Slot& sl : S->slots) {
        sl.arena.release();
        sl.rb[0].release();
        sl.rb[1].release();
        sl.prefix[0].release();
        sl.prefix[1].release();
        sl.desc.release();
        if (sl.h_rb) (void)hipHostFree(sl.h_rb);
        if (sl.h_prefix) (void)hipHostFree(sl.h_prefix);
        if (sl.h_desc) (void)hipHostFree(sl.h_desc);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    S->pool.release();
    S->ray_count.release();
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return HR_OK;
}

hr_status hr_server_add_scene(hr_server* S, const hr_server_scene* d, uint32_t* id)
{
    if (!S || !d || !id) return fail(HR_ERR_ARG, "null argument");
    if (!d->keys || d->nframes == 0 || d->nv == 0 || (d->nt > 0 && !d->idx))
        return fail(HR_ERR_ARG, "a scene needs key frames and indices");
    for (size_t i = 0; i < 3 * (size_t)d->nt; ++i)
        if (d->idx[i] >= d->nv) return fail(HR_ERR_ARG, "vertex index out of range");
    if (d->nmat > 0 && d->mat_idx)
        for (uint32_t i = 0; i < d->nt; ++i)
            if (d->mat_idx[i] >= d->nmat) return fail(HR_ERR_ARG, "material index out of range");
    auto sc = std::make_unique<Scene>();
    sc->nframes = d->nframes;
    sc->nv = d->nv;
    sc->nt = d->nt;
    sc->keys.assign(d->keys, d->keys + (size_t)d->nframes * 3 * d->nv);
    sc->idx.assign(d->idx, d->idx + 3 * (size_t)d->nt);
    if (d->mats && d->nmat) sc->mats.assign(d->mats, d->mats + 12 * (size_t)d->nmat);
    if (d->mat_idx && d->nmat) sc->mat_idx.assign(d->mat_idx, d->mat_idx + d->nt);
    if (d->lights && d->nlights > 0) {
        sc->lights.assign(d->lights, d->lights + d->nlights);
        if (d->light_weights) sc->weights.assign(d->light_weights, d->light_weights + d->nlights);
    }
    sc->lt = d->lambda_top > 0.f ? d->lambda_top : 1.2f;
    sc->ll = d->lambda_leaf > 0.f ? d->lambda_leaf : 2.2f;
    std::lock_guard<std::mutex> lk(S->mu);
    *id = (uint32_t)S->scenes.size();
    S->scenes.push_back(std::move(sc));
    return HR_OK;
}

hr_status hr_server_submit(hr_server* S, const hr_request* r, uint64_t* id)
{
    if (!S || !r || !id) return fail(HR_ERR_ARG, "null argument");
    if (r->kind < HR_REQ_CLOSEST || r->kind > HR_REQ_REBUILD) return fail(HR_ERR_ARG, "unknown request kind");
    if (r->layout != HR_RAYS_SOA && r->layout != HR_RAYS_RECORDS) return fail(HR_ERR_ARG, "unknown ray layout");
    if (r->kind != HR_REQ_REBUILD && r->n > 0 && (!r->org || (!r->dir && r->layout == HR_RAYS_SOA)))
        return fail(HR_ERR_ARG, "null rays");
    if (r->kind == HR_REQ_OCCLUDED && !(r->tmax > 0.f)) return fail(HR_ERR_ARG, "tmax must be positive");
    if (r->kind == HR_REQ_PATH && (r->depth < 1 || r->depth > 64)) return fail(HR_ERR_ARG, "depth must be in [1, 64]");
    if (r->n > (1u << 30)) return fail(HR_ERR_ARG, "at most 2^30 rays per request");
    if (r->stride != 0 && r->stride < r->n) return fail(HR_ERR_ARG, "stride smaller than the ray count");
    auto q = std::make_unique<Req>();
    q->r = *r;
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->fatal.empty()) return fail(HR_ERR_DEVICE, S->fatal);
        if (r->scene >= S->scenes.size()) return fail(HR_ERR_ARG, "scene out of range");
        q->id = S->next_id++;
        *id = q->id;
        S->queue.push_back(q.get());
        S->reqs.emplace(q->id, std::move(q));
    }
    S->cv_work.notify_one();
    return HR_OK;
}

hr_status hr_server_wait(hr_server* S, uint64_t id, hr_request_result* out)
{
    if (!S) return fail(HR_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HR_ERR_ARG, "unknown request id");
    Req* q = it->second.get();
    S->cv_done.wait(lk, [&] { return q->done || !S->fatal.empty(); });
    if (!q->done) return fail(HR_ERR_DEVICE, S->fatal);
    if (out) {
        out->status = q->status;
        out->latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
        out->wave = q->wave;
    }
    const hr_status st = q->status;
    const std::string err = q->error;
    S->reqs.erase(it);
    if (st != HR_OK) return fail(st, err);
    return HR_OK;
}

hr_status hr_server_poll(hr_server* S, uint64_t id, int* done)
{
    if (!S || !done) return fail(HR_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(HR_ERR_ARG, "unknown request id");
    *done = it->second->done ? 1 : 0;
    return HR_OK;
}

hr_status hr_server_get_stats(hr_server* S, hr_server_stats* out)
{
    if (!S || !out) return fail(HR_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    out->resident_bytes = S->resident_bytes;
    out->peak_resident_bytes = S->peak_bytes;
    out->resident_scenes = S->resident_count;
    return HR_OK;
}

}  // extern "C"
