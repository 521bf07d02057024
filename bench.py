"""Benchmark of the per-frame hot path on one MI355X.

Workload (default): 1,000,000 triangles -- 900k small triangles in a unit
cube inside a 100k-triangle room box of extent 100 -- two-level grid with
top density 1/16 and leaf density 3, 1024x1024 primary rays plus 8 ambient-
occlusion rays per hit (occlusion distance 0.05 * |bounds diagonal|, the
reference's alpha).  One step = one frame of the hot path on resident data:
full two-level rebuild (interpolation, bounds, build) + AO render.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints one JSON line.  `value` is Mrays/s over the timed steps (all rays:
primary + AO), with the rebuild inside the step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s (AO, path); grid rebuild ms/frame; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic work per DDA step / triangle test (flops), and bytes gathered
FLOPS_PER_STEP = 20
FLOPS_PER_TEST = 50
BYTES_PER_STEP = 8          # one (begin, end) cell or leaf
BYTES_PER_TEST = 4 + 36     # primitive id + triangle record (9 floats)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-pipelined", action="store_true",
                    help="soup_ao: skip the extra run of the same steps on two frame lanes")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration)")
    ap.add_argument("--gather-calibration", action="store_true",
                    help="also run a 1 GiB gather of 128-byte rows in random order (counter calibration)")
    ap.add_argument("--scenes", type=int, default=256, help="batched workload: number of scenes")
    ap.add_argument("--requests", type=int, default=2000, help="served workload: number of requests")
    ap.add_argument("--workload", default="soup_ao",
                    choices=["soup_ao", "keyframes_pt", "batched", "served", "instanced_ao"],
                    help="soup_ao: 1M-triangle two-level rebuild + AO frame (default); "
                         "keyframes_pt: 708x708-quad heightfield over 30 key frames, "
                         "interpolation + rebuild + path tracing per frame, captured as HIP graphs; "
                         "batched: 256 scenes of 1k-200k triangles with 4k-1M rays each, one segmented "
                         "build + mixed persistent traces after autotuning, against per-scene launches; "
                         "served: Poisson requests over 64 resident scenes through the request server, "
                         "against sequential per-request launches; "
                         "instanced_ao: 10,000 instances of a 20k-triangle sphere and a 100k-triangle soup "
                         "in a room, instanced hierarchy rebuild + AO frame (4 AO rays per hit)")
    return ap.parse_args()


def copy_calibration(nbytes=256 << 20):
    """One plain copy kernel of known size through the HIP runtime, so that a
    profiler's FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0      # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def gather_calibration(rows=1 << 23, row_floats=32):
    """One gather of known bytes: `rows` rows of 128 bytes (one cache line
    each, 1 GiB in all, beyond the 256 MB last-level cache) read in a random
    permutation by torch's index_select, so that a profiler's read counters
    can be scaled for scattered line-sized reads as the copy scales them for
    streaming reads.  Reads rows * 128 bytes of rows plus 8 bytes of index
    per row; writes rows * 128 bytes."""
    import torch
    src = torch.ones((rows, row_floats), dtype=torch.float32, device="cuda")
    idx = torch.randperm(rows, device="cuda")
    for _ in range(2):
        out = torch.index_select(src, 0, idx)
    torch.cuda.synchronize()
    del src, idx, out


def host_cpu():
    """Core count and model of the host the baseline runs on."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return n, model


def roofline_ao(ws, npix_hit, ao_rays, kernel_ms):
    """Algorithmic bytes and flops of one AO-ray kernel launch."""
    steps, tests = ws.secondary.steps, ws.secondary.tests
    bytes_alg = BYTES_PER_STEP * steps + BYTES_PER_TEST * tests + 32 * npix_hit + ao_rays
    flops = FLOPS_PER_STEP * steps + FLOPS_PER_TEST * tests
    sec = kernel_ms * 1e-3
    gbs = bytes_alg / sec / 1e9
    tflops = flops / sec / 1e12
    return bytes_alg, flops, gbs, tflops


COUNTER_FILE = os.path.join(ROOT, "profiles", "counters_soup_in_room_ao_1024.md")


def committed_counter(key):
    """A figure of the committed rocprofv3 counter summary of this workload
    (`key = value` lines), or None."""
    import re
    try:
        text = open(COUNTER_FILE).read()
    except OSError:
        return None
    m = re.search(r"^" + re.escape(key) + r" = ([0-9.eE+-]+)$", text, re.M)
    return float(m.group(1)) if m else None


def build_bytes(nt, nv, info):
    """Algorithmic bytes of one two-level rebuild: each phase reads its inputs
    and writes its outputs once, sort = one read and one write of keys and
    values (interpolation 36 V; top level 24 N + 24 V + 40 P_t + 8 C_t; leaf
    resolution and begin scan 20 C_t; leaf pair count and write
    2 (8 P_t + 8 C_t + 12 N + 12 V); leaf pairs, sort, ranges 36 P_l + 8 L)."""
    V, N = nv, nt
    Ct, Pt, L, Pl = info.num_cells, info.num_top_pairs, info.num_leaves, info.num_refs
    return 36 * V + (24 * N + 24 * V + 40 * Pt + 8 * Ct) + 20 * Ct + 2 * (8 * Pt + 8 * Ct + 12 * N + 12 * V) \
        + 36 * Pl + 8 * L


def cpu_baseline(scene, cam, ao_dist, threads):
    """The CPU oracle (a scalar port of the reference algorithm) on the same
    frame -- scene, view, 1024x1024 primary rays, 8 AO rays per hit -- on
    `threads` host threads, and its two-level build of the full scene on one
    thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    v = scene.verts
    m = O.OMesh(v[0], v[1], v[2], scene.idx)
    bmin, bmax = O.bounds(v)
    t0 = time.perf_counter()
    g = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    build_s = time.perf_counter() - t0
    c = O.camera(cam["pos"], cam["dir"], cam["up"], cam["fov"], *cam["res"])
    p = O.params(seed=(0x1234, 0x5678), jitter=1, ao_samples=8, ao_dist=ao_dist, nthreads=threads)
    t0 = time.perf_counter()
    _, _, st = O.render_ao(g, c, p, scene.materials, scene.mat_idx)
    render_s = time.perf_counter() - t0
    # the same frame on one thread
    p1 = O.params(seed=(0x1234, 0x5678), jitter=1, ao_samples=8, ao_dist=ao_dist, nthreads=1)
    t0 = time.perf_counter()
    _, _, st1 = O.render_ao(g, c, p1, scene.materials, scene.mat_idx)
    one_s = time.perf_counter() - t0
    nproc, model = host_cpu()
    return {"value": round(st.rays / render_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"one full AO frame of the same workload ({st.rays} rays, "
                      f"{render_s:.2f} s on {threads} threads, {one_s:.2f} s on 1 thread); two-level build of all "
                      f"{scene.nt} triangles on 1 thread: {build_s * 1e3:.1f} ms",
            "value_1thread": round(st1.rays / one_s / 1e6, 3),
            "nproc": nproc, "cpu_model": model,
            "build_ms_1thread": round(build_s * 1e3, 2)}


def cpu_baseline_pt(scene, keys, cam, threads, t_frame):
    """The CPU oracle on one frame of the key-frame workload: interpolation,
    two-level build (1 thread) and the path-traced frame on `threads` threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    k1, k2, c = O.frame_ids(t_frame, 0.4, keys.shape[0])
    t0 = time.perf_counter()
    v, bmin, bmax = O.lerp_bounds(keys[k1], keys[k2], c)
    m = O.OMesh(v[0], v[1], v[2], scene.idx)
    g = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    build_s = time.perf_counter() - t0
    oc = O.camera(cam["pos"], cam["dir"], cam["up"], cam["fov"], *cam["res"])
    p = O.params(seed=(0x1234, 0x5678), jitter=1, pt_depth=4, nthreads=threads)
    t0 = time.perf_counter()
    _, st = O.render_pt(g, oc, p, scene.materials, scene.mat_idx, scene.lights)   # primary rays included
    render_s = time.perf_counter() - t0
    rays = st.rays
    return {"value": round(rays / render_s / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"one path-traced frame of the same workload ({rays} rays, {render_s:.2f} s on "
                      f"{threads} threads); interpolation + two-level build of all {scene.nt} "
                      f"triangles on 1 thread: {build_s * 1e3:.1f} ms",
            "build_ms_1thread": round(build_s * 1e3, 2)}


def main_keyframes(args):
    """Config: 708x708-quad heightfield (1,002,528 triangles, 502,681 vertices)
    over 30 key frames; per frame: interpolation (step 0.4), full two-level
    rebuild (top 1/16, leaf 3), 1024x768 primary rays and 1-spp path tracing
    with 4 path vertices and Russian roulette, all replayed from HIP graphs."""
    from happyray_amd import Engine, GRID_TWO_LEVEL, MODE_PATH, scenes
    from happyray_amd.engine import CAMERA_KERNEL_TRIALS
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    scene, keys = scenes.heightfield_keyframes()
    cam = scene.camera
    e = Engine(seed=(0x1234, 0x5678))
    e.set_keyframes(keys, scene.idx)
    e.upload_materials(scene.materials, scene.mat_idx)
    e.set_lights(scene.lights)
    e.set_camera_dict(cam)
    e.set_grid(GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    nf = keys.shape[0]
    step_size = 0.4

    def frame_ids(t):
        k1 = int(t)
        return k1, (k1 + 1 + int(step_size)) % nf, np.float32(t) - np.float32(k1)

    state = {"t": np.float32(0.0)}

    def step(i, wait=False):
        k1, k2, c = frame_ids(float(state["t"]))
        if wait:
            e.render_frame_graph(k1, k2, float(c), image_id=i, mode=MODE_PATH, jitter=1, pt_depth=4,
                                 download=False)
        else:   # queued: the next frame's rebuild follows this render with no host round trip
            e.submit_frame_graph(k1, k2, float(c), image_id=i, mode=MODE_PATH, jitter=1, pt_depth=4)
        t = np.float32(state["t"]) + np.float32(step_size)
        if int(t) >= nf:
            t -= np.float32(nf)
        state["t"] = t

    for i in range(CAMERA_KERNEL_TRIALS):   # the camera-ray kernel choice, before warmup
        step(i)
    for i in range(args.warmup):
        step(i)
    e.synchronize()
    r0 = e.timing().rays_traced
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    e.synchronize()
    wall = time.perf_counter() - t0
    rays = e.timing().rays_traced - r0
    # per-graph times of the same frames, each waited for (outside the timed region)
    build_ms, render_ms, prim_ms, path_ms = [], [], [], []
    for i in range(args.steps):
        step(args.warmup + args.steps + i, wait=True)
        tm = e.timing()
        build_ms.append(tm.build_ms)
        render_ms.append(tm.render_ms)
        prim_ms.append(tm.primary_ms)
        path_ms.append(tm.secondary_ms)
    info = e.grid_info()
    out = {
        "metric": METRIC, "value": round(rays / wall / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "keyframes_pt: 708x708-quad heightfield (1,002,528 triangles, 502,681 "
                               "vertices), 30 key frames of a seeded sum of 4 sinusoids, step 0.4; per "
                               "frame interpolation + two-level rebuild (top 1/16, leaf 3) + 1024x768 "
                               "1-spp path tracing (4 vertices, Russian roulette), replayed HIP graphs, "
                               "each frame queued while the one before it renders",
                   "triangles": scene.nt, "rays_per_step": int(rays // args.steps),
                   "top_res": list(info.res), "leaves": info.num_leaves, "leaf_refs": info.num_refs},
        "rebuild_ms": round(float(np.mean(build_ms)), 4),
        "render_ms": round(float(np.mean(render_ms)), 4),
        "render_mrays_s": round(rays / args.steps / (float(np.mean(render_ms)) * 1e-3) / 1e6, 3),
        "kernel_ms": {"primary": round(float(np.mean(prim_ms)), 4), "paths": round(float(np.mean(path_ms)), 4)},
        "camera_kernel": e.timing().primary_kernel,
    }
    e.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_pt(scene, keys, cam, args.cpu_threads, 0.0)
    print(json.dumps(out), flush=True)


def main_batched(args):
    """Config: 256 scenes (random soups and icosphere clusters) of 1k-200k
    triangles, log-uniform, each with 4k-1M rays (half primary closest-hit,
    half ambient-occlusion any-hit from random surface points).  One step
    rebuilds every scene and traces every ray.  Batched: one segmented
    two-level build and two persistent launches (closest, any) over all
    scenes, densities and persistence chosen by the autotuner beforehand.
    Plain: per scene, its own context, hr_build then hr_trace_dev and
    hr_occluded_dev with the default densities, waiting for each request."""
    import ctypes as C
    from happyray_amd import Batch, Engine, GRID_TWO_LEVEL, scenes
    from happyray_amd.devmem import hip
    t0 = time.perf_counter()
    scs, items = scenes.batched_workload(n_scenes=args.scenes)
    gen_s = time.perf_counter() - t0
    nt_total = sum(int(f.shape[0]) for _, f in scs)
    b = Batch()
    b.set_scenes(scs)
    rays = b.make_rays(items)
    b.build()
    t0 = time.perf_counter()
    tune = b.autotune(rays)
    tune_s = time.perf_counter() - t0

    def step():
        b.build()
        b.trace(rays)
        b.synchronize()

    for _ in range(args.warmup):
        step()
    builds, traces = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        bm, tm, _ = b.timing()
        builds.append(bm)
        traces.append(tm)
    wall = time.perf_counter() - t0
    batched = rays.rays * args.steps / wall / 1e6
    res_b = rays.results()

    # plain per-request path on the same resident rays
    engines = []
    for v, f in scs:
        e = Engine(seed=(0x1234, 0x5678))
        e.upload(v, None, 1.0, f)
        e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
        engines.append(e)
    by_scene = {}
    for (s, org, d, anyhit, tmax), bufs in zip(items, rays.bufs):
        by_scene.setdefault(s, []).append((org.shape[1], anyhit, tmax, bufs))

    def plain_step():
        for s, e in enumerate(engines):
            try:
                e.build_acc_struct()
            except Exception as ex:
                free, total = C.c_size_t(), C.c_size_t()
                hip().hipMemGetInfo(C.byref(free), C.byref(total))
                print(f"plain build of scene {s} ({scs[s][1].shape[0]} triangles) failed: {ex}; device memory "
                      f"free {free.value / 2**30:.1f} of {total.value / 2**30:.1f} GiB", file=sys.stderr, flush=True)
                raise
            for n, anyhit, tmax, (o, d, t, p, occ) in by_scene.get(s, []):
                if anyhit:
                    e.occluded_dev(o.value, d.value, n, tmax, occ.value)
                else:
                    e.trace_dev(o.value, d.value, n, t.value, p.value)
            e.synchronize()

    plain_steps = max(1, min(args.steps, 3))
    plain_step()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        plain_step()
    plain_wall = time.perf_counter() - t0
    plain = rays.rays * plain_steps / plain_wall / 1e6
    res_p = rays.results()
    same = total = 0
    for (s, org, d, anyhit, tmax), rb_, rp_ in zip(items, res_b, res_p):
        same += int(np.sum(rb_[0] == rp_[0]) if anyhit else np.sum((rb_[0] == rp_[0]) & (rb_[1] == rp_[1])))
        total += org.shape[1]
    lt, ll, _ = b.densities()
    out = {
        "metric": METRIC, "value": round(batched, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(scs)} scenes (even: random soups, odd: icosphere clusters), "
                               "1k-200k triangles log-uniform, 4k-1M rays each log-uniform (half primary "
                               "closest hit, half AO any hit from random surface points); step = rebuild of "
                               "every scene + trace of every ray",
                   "scenes": len(scs), "triangles": nt_total, "rays_per_step": rays.rays,
                   "seed": 4},
        "batched_build_ms": round(float(np.mean(builds)), 3),
        "batched_trace_ms": round(float(np.mean(traces)), 3),
        "plain_mrays_s": round(plain, 3),
        "plain_ms_per_step": round(plain_wall * 1e3 / plain_steps, 3),
        "speedup_vs_plain": round(batched / plain, 3),
        "results_identical_to_plain": round(same / max(total, 1), 6),
        "autotune": {"lambda_top": [round(x, 4) for x in tune.lambda_top],
                     "lambda_leaf": [round(x, 4) for x in tune.lambda_leaf],
                     "size_classes": ["<4K", "<16K", "<64K", ">=64K"],
                     "refill": tune.tuning.refill, "blocks_per_cu": tune.tuning.blocks_per_cu,
                     "evaluations": tune.evaluations, "seconds": round(tune_s, 2)},
        "generation_s": round(gen_s, 1),
    }
    for e in engines:
        e.close()
    out["two_batches"] = two_batches(b, rays, scs, items, args.steps, args.warmup)
    rays.free()
    b.close()
    print(json.dumps(out), flush=True)


def two_batches(a, rays_a, scs, items, steps, warmup):
    """The same steps on two Batch objects with the same scenes, rays and
    tuning, taking steps in turn without waiting, so that one batch's rebuild
    overlaps the other's traces (results per batch unchanged)."""
    from happyray_amd import Batch
    lt, ll, (refill, bpc) = a.densities()
    b = Batch()
    b.set_scenes(scs)
    rays_b = b.make_rays(items)
    b.set_densities(lt, ll)
    b.set_tuning(refill, bpc)
    pair = [(a, rays_a), (b, rays_b)]
    for i in range(2 * warmup):
        x, r = pair[i % 2]
        x.build()
        x.trace(r)
    a.synchronize()
    b.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        x, r = pair[i % 2]
        x.build()
        x.trace(r)
    a.synchronize()
    b.synchronize()
    wall = time.perf_counter() - t0
    same = all(all(np.array_equal(u, v) for u, v in zip(p, q)) for p, q in zip(rays_a.results(), rays_b.results()))
    rays_b.free()
    b.close()
    return {"value": round(rays_a.rays * steps / wall / 1e6, 3), "unit": "Mrays/s",
            "ms_per_step": round(wall * 1e3 / steps, 3), "results_identical": bool(same),
            "note": "two batches stepping in turn; each build waits only for its own batch's last traces"}


def main_served(args):
    """Config: 64 resident scenes of 10k-2M triangles; Poisson arrivals at
    2,000 requests/s; each request a rebuild (5 %) or a trace of 1k-4M rays
    (closest hit, occlusion or depth-4 path, uniformly).  Rays come from
    per-scene pools resident on the device (requests reference sub-ranges).
    Served: the request server with a 200 GB budget, requests submitted at
    their arrival times by a client thread.  Plain: one context per scene,
    each request run alone with the single-scene device entry points in
    arrival order (its completion time is its start -- the later of its
    arrival and the previous completion -- plus its measured run time)."""
    import ctypes as C
    import threading
    from happyray_amd import Engine, GRID_TWO_LEVEL, Server, scenes
    from happyray_amd.devmem import DeviceArray, copy_planes, hip
    from happyray_amd.server import REQ_CLOSEST, REQ_OCCLUDED, REQ_PATH
    t0 = time.perf_counter()
    scs, pools, reqs = scenes.served_workload(n_requests=args.requests)
    gen_s = time.perf_counter() - t0
    P = 1 << 22                      # device pool size per kind and scene (4M rays)
    rng = np.random.default_rng(55)
    offsets = [int(rng.integers(0, P - r[3] + 1)) if r[3] else 0 for r in reqs]
    dev = []
    for po, pd, ao, ad, tmax in pools:
        reps = P // po.shape[1]
        dev.append(tuple(DeviceArray.from_host(np.tile(a, (1, reps))) for a in (po, pd, ao, ad)) + (tmax,))
    kinds = {"closest": REQ_CLOSEST, "occluded": REQ_OCCLUDED, "path": REQ_PATH}
    nt_total = sum(int(f.shape[0]) for _, f, _ in scs)
    mat = np.stack([scenes.material((0.5, 0.5, 0.5))])

    srv = Server(memory_budget=200 << 30)
    for keys, f, light in scs:
        srv.add_scene(keys, f, mat, None, light[None])
    # make every scene resident before the timed stream
    warm = [srv.submit_device(REQ_CLOSEST, s, 64, dev[s][0].value, dev[s][1].value, stride=P)
            for s in range(len(scs))]
    for w in warm:
        w.wait()
    st0 = srv.stats()

    def ray_ptrs(s, kind, off):
        o, d = (dev[s][0], dev[s][1]) if kind != "occluded" else (dev[s][2], dev[s][3])
        return o.value + 4 * off, d.value + 4 * off

    pending = [None] * len(reqs)
    start = time.perf_counter() + 0.05

    def client():
        for i, (t, kind, s, n, _, k1, k2, c) in enumerate(reqs):
            delay = start + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if kind == "rebuild":
                pending[i] = srv.submit_rebuild(s, k1, k2, c)
            else:
                o, d = ray_ptrs(s, kind, offsets[i])
                pending[i] = srv.submit_device(kinds[kind], s, n, o, d, stride=P, tmax=dev[s][4], depth=4)

    th = threading.Thread(target=client)
    th.start()
    th.join()
    lat = np.zeros(len(reqs))
    for i, p in enumerate(pending):
        p.wait()
        lat[i] = p.latency_ms
    end = time.perf_counter()
    st = srv.stats()
    rays = sum(r[3] for r in reqs)
    served_wall = end - start
    served = rays / served_wall / 1e6

    # capacity: the same requests submitted at once (arrival times ignored), so
    # the server is never idle waiting for an arrival
    cap_pending = []
    cap_start = time.perf_counter()
    for i, (t, kind, s, n, _, k1, k2, c) in enumerate(reqs):
        if kind == "rebuild":
            cap_pending.append(srv.submit_rebuild(s, k1, k2, c))
        else:
            o, d = ray_ptrs(s, kind, offsets[i])
            cap_pending.append(srv.submit_device(kinds[kind], s, n, o, d, stride=P, tmax=dev[s][4], depth=4))
    for p in cap_pending:
        p.wait()
    cap_wall = time.perf_counter() - cap_start
    capacity = rays / cap_wall / 1e6
    srv.close()

    # plain: sequential single-scene launches on the same resident pools
    engines = []
    for keys, f, light in scs:
        e = Engine(seed=(0x1234, 0x5678))
        e.set_keyframes(keys, f)
        e.upload_materials(mat)
        e.set_lights(light[None])
        e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
        e.set_frame(0, 0, 0.0)
        e.build_acc_struct()
        engines.append(e)
    nmax = max(r[3] for r in reqs)
    so, sd = DeviceArray((3, nmax), np.float32), DeviceArray((3, nmax), np.float32)
    t_s, p_s = DeviceArray((nmax,), np.float32), DeviceArray((nmax,), np.uint32)
    occ_s, rgb_s = DeviceArray((nmax,), np.uint8), DeviceArray((nmax, 3), np.float32)
    service = np.zeros(len(reqs))
    by_kind = {"rebuild": 0.0, "closest": 0.0, "occluded": 0.0, "path": 0.0}
    for i, (t, kind, s, n, _, k1, k2, c) in enumerate(reqs):
        e = engines[s]
        t1 = time.perf_counter()
        if kind == "rebuild":
            e.set_frame(k1, k2, c)
            e.build_acc_struct()
        else:
            o, d = ray_ptrs(s, kind, offsets[i])
            strm = e.stream()
            copy_planes(so.value, o, n, P, stream=strm)
            copy_planes(sd.value, d, n, P, stream=strm)
            if kind == "closest":
                e.trace_dev(so.value, sd.value, n, t_s.value, p_s.value)
            elif kind == "occluded":
                e.occluded_dev(so.value, sd.value, n, dev[s][4], occ_s.value)
            else:
                e.path_trace_dev(so.value, sd.value, n, 4, rgb_s.value)
            e.synchronize()
        service[i] = time.perf_counter() - t1
        by_kind[kind] += service[i]
    arr = np.array([r[0] for r in reqs])
    done = np.zeros(len(reqs))
    clock = 0.0
    for i in range(len(reqs)):
        clock = max(clock, arr[i]) + service[i]
        done[i] = clock
    plain_lat = (done - arr) * 1e3
    plain = rays / done[-1] / 1e6
    for e in engines:
        e.close()
    out = {
        "metric": METRIC, "value": round(served, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(served_wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"served: {len(reqs)} Poisson requests at 2000/s over {len(scs)} resident "
                               "scenes of 10k-2M triangles (soups and icosphere clusters, two key frames "
                               "each); 5% rebuilds, else 1k-4M rays of closest hit, occlusion or depth-4 "
                               "paths; 200 GB budget; step = the whole request stream",
                   "scenes": len(scs), "triangles": nt_total, "requests": len(reqs), "request_rays": rays,
                   "seed": 5},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "mean": round(float(lat.mean()), 3)},
        "plain_mrays_s": round(plain, 3),
        "plain_latency_ms": {"p50": round(float(np.percentile(plain_lat, 50)), 3),
                             "p99": round(float(np.percentile(plain_lat, 99)), 3)},
        "speedup_vs_plain": round(served / plain, 3),
        "capacity_mrays_s": round(capacity, 3),
        "capacity_speedup_vs_plain": round(capacity / (rays / service.sum() / 1e6), 3),
        "plain_service_s": {k: round(v, 4) for k, v in by_kind.items()},
        "server": {"waves": st.waves - st0.waves, "builds": st.builds, "rebuilds": st.rebuilds,
                   "evictions": st.evictions, "peak_resident_gib": round(st.peak_resident_bytes / 2**30, 2),
                   "resident_scenes": st.resident_scenes},
        "offered_mrays_s": round(rays / (arr[-1] - arr[0]) / 1e6, 1),
        "generation_s": round(gen_s, 1),
    }
    print(json.dumps(out), flush=True)


def main_instanced(args):
    """Config: the instanced hierarchy (the reference's TLGRIDHIERARCHY build).
    Objects: a room of extent 100 (768 triangles), a unit icosphere (20,480
    triangles) and a 100,000-triangle soup in [-1, 1]^3 (corners within 0.02 of
    their centre); 10,000 instances (the
    room once, spheres and soups with random rotation, scale 0.5-1.5 and
    position).  One step = hierarchy rebuild (object grids + top grid) + the
    hierarchy's AO frame: 1024x1024 primary rays + 4 AO rays per hit
    (NUMAMBIENTOCCLUSIONSAMPLES_TLGH, TLGridHierarchyAOIntegrator.h:47)."""
    from happyray_amd import Engine, GRID_HIERARCHY, MODE_AO, scenes
    sc, ins = scenes.instanced_scene(seed=12, n_inst=10000, subdiv=5, n_soup=100000, extent=100.0,
                                     res=(1024, 1024), soup_size=0.02)
    e = Engine(seed=(0x1234, 0x5678))
    e.upload_scene(sc)
    e.set_instances(ins.obj_first, ins.objects, ins.matrices, ins.aabb_min, ins.aabb_max)
    e.set_grid(GRID_HIERARCHY, 1.2, 2.2)
    e.set_camera_dict(sc.camera)
    e.build()
    ao_dist = 0.05 * e.get_bounding_box_diagonal_length()      # RTEngine.cu:261
    S = 4

    def step(i):
        e.build()
        e.render_frame(i, MODE_AO, jitter=1, ao_samples=S, ao_dist=ao_dist, download=False)

    for i in range(args.warmup):   # (the hierarchy's camera rays take no timed kernel choice)
        step(i)
    e.synchronize()
    r0 = e.timing().rays_traced
    build_ms, render_ms, prim_ms, ao_ms = [], [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
        tm = e.timing()
        build_ms.append(tm.build_ms)
        render_ms.append(tm.render_ms)
        prim_ms.append(tm.primary_ms)
        ao_ms.append(tm.secondary_ms)
    e.synchronize()
    wall = time.perf_counter() - t0
    rays = e.timing().rays_traced - r0
    info = e.grid_info()
    out = {
        "metric": METRIC, "value": round(rays / wall / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "instanced: 10,000 instances of a 20,480-triangle icosphere and a "
                               "100,000-triangle soup in a 768-triangle room of extent 100; instanced "
                               "hierarchy top 1.2 leaf 2.2; 1024x1024 primary + 4 AO rays per hit; "
                               "step = hierarchy rebuild + AO frame",
                   "mesh_triangles": sc.nt, "instances": int(len(ins.objects)),
                   "rays_per_step": int(rays // args.steps), "top_res": list(info.res),
                   "object_cells": info.num_leaves, "object_refs": info.num_refs,
                   "instance_refs": info.num_top_pairs, "ao_dist": round(ao_dist, 4)},
        "rebuild_ms": round(float(np.mean(build_ms)), 4),
        "render_ms": round(float(np.mean(render_ms)), 4),
        "render_mrays_s": round(rays / args.steps / (float(np.mean(render_ms)) * 1e-3) / 1e6, 3),
        "kernel_ms": {"primary": round(float(np.mean(prim_ms)), 4), "ao_rays": round(float(np.mean(ao_ms)), 4)},
    }
    e.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        v = sc.verts
        _, bmin, bmax = O.lerp_bounds(v, v, 1.0)
        t0 = time.perf_counter()
        h = O.OHier(O.OMesh(v[0], v[1], v[2], sc.idx), v, bmin, bmax, ins.obj_first, ins.objects, ins.matrices,
                    ins.aabb_min, ins.aabb_max)
        build_s = time.perf_counter() - t0
        c = sc.camera
        rx, ry = 128, 128
        oc = O.camera(c["pos"], c["dir"], c["up"], c["fov"], rx, ry)
        p = O.params(seed=(0x1234, 0x5678), jitter=1, ao_samples=S, ao_dist=ao_dist)
        t0 = time.perf_counter()
        _, occ, hits, _, _ = O.render_ao_hier(h, oc, p, sc.materials, sc.mat_idx)
        render_s = time.perf_counter() - t0
        nr = rx * ry + int((hits[6] < 3.4e38).sum()) * S
        out["cpu_baseline"] = {
            "value": round(nr / render_s / 1e6, 4), "unit": "Mrays/s", "cores": 1, "kind": "port",
            "sample": f"the same view at {rx}x{ry} ({nr} rays, {render_s:.2f} s on 1 thread); "
                      f"hierarchy build on 1 thread: {build_s * 1e3:.1f} ms"}
    print(json.dumps(out), flush=True)


def soup_two_lanes(scene, ao_dist, S, steps, warmup):
    """The same steps with frame_lanes 2, where a frame's rebuild and camera
    rays also overlap the previous frame's AO rays: whole-job rate and the
    per-frame times from the events of the timed frames.  Reported beside the
    headline line, whose AO kernel has the GPU to itself."""
    from happyray_amd import Engine, GRID_TWO_LEVEL, MODE_AO
    from happyray_amd.engine import CAMERA_KERNEL_TRIALS
    e = Engine(seed=(0x1234, 0x5678))
    e.set_keyframes(np.asarray(scene.verts, np.float32)[None], scene.idx)
    e.upload_materials(scene.materials, scene.mat_idx)
    e.set_grid(GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    e.set_camera_dict(scene.camera)
    e.set_tuning(frame_lanes=2)

    def step(i):
        e.submit_frame_graph(0, 0, 0.0, image_id=i, mode=MODE_AO, jitter=1, ao_samples=S, ao_dist=ao_dist)

    for i in range(CAMERA_KERNEL_TRIALS + warmup):
        step(i)
    e.synchronize()
    t0_dev = e.timing()
    t0 = time.perf_counter()
    for i in range(steps):
        step(i)
    e.synchronize()
    wall = time.perf_counter() - t0
    t1_dev = e.timing()
    n = max(t1_dev.frames_replayed - t0_dev.frames_replayed, 1)

    def per_frame(name):
        return round((getattr(t1_dev, name) - getattr(t0_dev, name)) / n, 4)

    out = {"value": round((t1_dev.rays_traced - t0_dev.rays_traced) / wall / 1e6, 3), "unit": "Mrays/s",
           "ms_per_step": round(wall * 1e3 / steps, 3), "frame_lanes": 2, "frames_timed": int(n),
           "rebuild_ms": per_frame("replay_build_ms"),
           "kernel_ms": {"primary": per_frame("replay_primary_ms"), "ao_rays": per_frame("replay_secondary_ms"),
                         "ao_shade": per_frame("replay_shade_ms")},
           "note": "frame_lanes 2: kernels of neighbouring frames share the GPU, so each takes longer "
                   "than alone while the frame rate rises"}
    e.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; sharding over several GPUs is "
                                   "described, not benchmarked", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "keyframes_pt":
        return main_keyframes(args)
    if args.workload == "batched":
        return main_batched(args)
    if args.workload == "served":
        return main_served(args)
    if args.workload == "instanced_ao":
        return main_instanced(args)
    from happyray_amd import Engine, GRID_TWO_LEVEL, MODE_AO, scenes
    from happyray_amd.engine import CAMERA_KERNEL_TRIALS

    scene = scenes.soup_in_room()
    cam = scene.camera
    e = Engine(seed=(0x1234, 0x5678))
    # the mesh as one resident key frame: every step interpolates it (as hr_build does), rebuilds the
    # grid and renders the AO frame, as captured frames on two lanes where a frame's rebuild runs beside
    # the previous frame's camera rays and the AO rays run alone
    e.set_keyframes(np.asarray(scene.verts, np.float32)[None], scene.idx)
    e.upload_materials(scene.materials, scene.mat_idx)
    e.set_grid(GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    e.set_camera_dict(cam)
    e.set_tuning(frame_lanes=3)
    e.set_frame(0, 0, 0.0)
    e.build_acc_struct()
    ao_dist = 0.05 * e.get_bounding_box_diagonal_length()
    S = 8

    def step(i):
        e.submit_frame_graph(0, 0, 0.0, image_id=i, mode=MODE_AO, jitter=1, ao_samples=S, ao_dist=ao_dist)

    for i in range(CAMERA_KERNEL_TRIALS):   # the camera-ray kernel choice, before warmup
        step(i)
    for i in range(args.warmup):
        step(CAMERA_KERNEL_TRIALS + i)
    e.synchronize()
    t0_dev = e.timing()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(CAMERA_KERNEL_TRIALS + args.warmup + i)
    e.synchronize()
    wall = time.perf_counter() - t0
    t1_dev = e.timing()
    rays = t1_dev.rays_traced - t0_dev.rays_traced
    ms_step = wall * 1e3 / args.steps
    value = rays / wall / 1e6
    # per-frame kernel times from the HIP events recorded on the lanes' streams over the timed frames
    n_rep = max(t1_dev.frames_replayed - t0_dev.frames_replayed, 1)

    def per_frame(name):
        return (getattr(t1_dev, name) - getattr(t0_dev, name)) / n_rep

    prim_ms, ao_avg, shade_ms = (per_frame("replay_primary_ms"), per_frame("replay_secondary_ms"),
                                 per_frame("replay_shade_ms"))
    rebuild_overlapped = per_frame("replay_build_ms")
    # the rebuild alone (outside the timed region): the same interpolation and build, uncaptured
    iso = []
    for i in range(5):
        e.build_acc_struct()
        iso.append(e.timing().build_ms)
    rebuild_ms = float(np.mean(iso))

    # work counters of one more frame of the same grid (uncaptured, outside the timed region)
    e.work_stats()
    e.render_frame(0, MODE_AO, jitter=1, ao_samples=S, ao_dist=ao_dist, download=False, collect_stats=True)
    ws = e.work_stats()
    planes, prims = e.primary_hits()
    npix_hit = int((prims < scene.nt).sum())
    ao_rays = ws.secondary.rays
    bytes_alg, flops, gbs, tflops = roofline_ao(ws, npix_hit, ao_rays, ao_avg)
    render_ms = prim_ms + ao_avg + shade_ms
    info = e.grid_info()
    tun = e.tuning()
    if not tun.persistent_secondary:
        ao_kernel = "k_ao_rays<true>"
    elif tun.test_batch:
        ao_kernel = "k_persist_bt<true, true, GenAO, 12>"
    else:
        ao_kernel = f"k_persist<true, true, GenAO, {'true' if tun.fuse_units else 'false'}>"

    bb = build_bytes(scene.nt, scene.nv, info)
    bcnt = committed_counter("build_counter_bytes")
    bbusy = committed_counter("build_counter_busy_us")
    tr = committed_counter(f"traffic_bytes_per_launch {ao_kernel}")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "soup_in_room: 900k small triangles in a unit cube inside a "
                               "100k-triangle room of extent 100 (1,000,000 triangles); two-level "
                               "grid top 1/16 leaf 3; 1024x1024 primary + 8 AO rays per hit; "
                               "step = rebuild + AO frame, as captured frames on two lanes (a frame's "
                               "rebuild runs beside the previous frame's camera rays, the AO rays run alone)",
                   "triangles": scene.nt, "rays_per_step": int(rays // args.steps),
                   "top_res": list(info.res), "leaves": info.num_leaves,
                   "leaf_refs": info.num_refs, "ao_dist": round(ao_dist, 4)},
        "rebuild_ms": round(rebuild_ms, 4),
        "rebuild_ms_beside_camera_rays": round(rebuild_overlapped, 4),
        "render_ms": round(render_ms, 4),
        "render_mrays_s": round(rays / args.steps / (render_ms * 1e-3) / 1e6, 3),
        "kernel_ms": {"primary": round(prim_ms, 4), "ao_rays": round(ao_avg, 4), "ao_shade": round(shade_ms, 4)},
        "frames_timed": int(n_rep),
        "camera_kernel": e.timing().primary_kernel,
        "roofline": {"kernel": ao_kernel, "bound": "hbm", "achieved": round(gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                     "traffic": None if tr is None else int(tr),
                     "traffic_source": os.path.relpath(COUNTER_FILE, ROOT),
                     "bytes_per_launch": int(bytes_alg),
                     "bytes": "algorithmic gathers (8 B per DDA step, 40 B per triangle test, 32 B per hit "
                              "pixel, 1 B per AO ray), served mostly by the L2 and the last-level cache",
                     "limiter": "VALU issue under divergence: the VALU is busy in about four fifths of the "
                                "cycles while an instruction computes for about two fifths of a wave's lanes "
                                "(the batched triangle tests run at full width, the grid steps diverge)",
                     "counters": {"valu_busy": committed_counter(f"counter {ao_kernel} valu_busy"),
                                  "lanes_active": committed_counter(f"counter {ao_kernel} lanes_active"),
                                  "l2_hit": committed_counter(f"counter {ao_kernel} l2_hit")},
                     "valu": {"achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 5),
                              "flops_per_launch": int(flops)},
                     "work": {"ao_rays": int(ao_rays), "steps": int(ws.secondary.steps),
                              "tests": int(ws.secondary.tests)}},
        "build_roofline": {"bound": "hbm", "algorithmic_bytes": int(bb),
                           "achieved": round(bb / (rebuild_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(bb / (rebuild_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "counter_bytes": None if bcnt is None else int(bcnt),
                           "counter_frac": None if bcnt is None or not bbusy else
                           round(bcnt / (bbusy * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    e.close()
    if not args.no_pipelined:
        out["two_lanes"] = soup_two_lanes(scene, ao_dist, S, args.steps, args.warmup)
    if args.copy_calibration:
        copy_calibration()
    if args.gather_calibration:
        gather_calibration()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, cam, ao_dist, args.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
