"""The five benchmark workloads checked on the GPU at their stated sizes (the
served and batched ones with fewer scenes and shorter ray lists):

1. 16 icospheres + ground quad (81,922 triangles), exact uniform grid at
   density 5, a 256x256 AO frame with 4 rays per hit within 2.0: grid,
   occlusion bits and image against the oracle;
2. 900k-triangle soup in a 100k-triangle room, two-level grid (1/16, 3): the
   grid bit-exact against the oracle (3.45M leaf pairs: the radix sort runs
   several 2048-key tiles per workgroup), then a 1024x1024 AO frame with 8 rays
   per hit within 0.05 of the diagonal, occlusion bit-exact;
3. the 708x708-quad heightfield (1,002,528 triangles, 30 key frames): frames
   replayed from captured HIP graphs equal the oracle's interpolation,
   two-level build and depth-4 path tracing at 1024x768;
4. 32 scenes of the batched generator (1k-200k triangles) with autotuned
   densities: every scene's grid and every ray equal single-scene runs, three
   scenes also against the oracle;
5. 8 scenes of the served generator (10k-2M triangles, two key frames each)
   and 200 requests (closest hit, occlusion, depth-4 paths, 5 % rebuilds):
   alone = in the mix under a budget that evicts = single-scene entry points.
"""
import threading

import numpy as np
import pytest

import pyoracle as O
from gpu_util import SEED, assert_grid_equal, gpu_engine, oracle_grid
from happyray_amd import (Batch, Engine, GRID_TWO_LEVEL, GRID_UNIFORM_EXACT, MODE_AO, MODE_PATH, Server,
                          scenes)

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
IMG_FRAC = 0.999


def assert_image_match(a, b):
    err = np.abs(a - b).max(axis=1)
    frac = (err <= IMG_TOL).mean()
    assert frac >= IMG_FRAC, f"only {frac:.6f} of pixels within {IMG_TOL} (max err {err.max()})"


def ocam(cam):
    return O.camera(cam["pos"], cam["dir"], cam["up"], cam["fov"], *cam["res"])


def test_config1_cluster_256():
    sc = scenes.icosphere_cluster(res=256)
    assert sc.nt == 81922
    print("config 1: cluster 256x256 AO", flush=True)
    e = gpu_engine(sc, GRID_UNIFORM_EXACT, 5.0, 1.0)
    og = oracle_grid(sc, GRID_UNIFORM_EXACT, 5.0, 1.0)
    assert_grid_equal(e, og)
    e.set_camera_dict(sc.camera)
    fb_g = e.render_frame(0, MODE_AO, jitter=1, ao_samples=4, ao_dist=2.0)
    occ_g = e.ao_occlusion(4)
    p = O.params(seed=SEED, jitter=1, ao_samples=4, ao_dist=2.0, nthreads=16)
    fb_o, occ_o, _ = O.render_ao(og, ocam(sc.camera), p, sc.materials, sc.mat_idx)
    np.testing.assert_array_equal(occ_g, occ_o)
    assert_image_match(fb_g, fb_o)
    e.close()


def test_config2_soup_in_room_full_size():
    sc = scenes.soup_in_room()
    assert sc.nt == 1_000_000
    print("config 2: soup in room, two-level (1/16, 3) build", flush=True)
    e = gpu_engine(sc, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    og = oracle_grid(sc, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    info = assert_grid_equal(e, og)
    assert info.num_refs > 2048 * 1024            # more keys than one 2048-key tile per workgroup
    e.set_camera_dict(sc.camera)
    dist = 0.05 * e.get_bounding_box_diagonal_length()
    print(f"config 2: 1024x1024 AO, 8 rays per hit within {dist}", flush=True)
    fb_g = e.render_frame(0, MODE_AO, jitter=1, ao_samples=8, ao_dist=dist)
    occ_g = e.ao_occlusion(8)
    planes_g, prims_g = e.primary_hits()
    p = O.params(seed=SEED, jitter=1, ao_samples=8, ao_dist=float(np.float32(dist)), nthreads=16)
    cam = ocam(sc.camera)
    _, prims_o, _ = O.primary(og, cam, p)
    np.testing.assert_array_equal(prims_g, prims_o)
    fb_o, occ_o, _ = O.render_ao(og, cam, p, sc.materials, sc.mat_idx)
    np.testing.assert_array_equal(occ_g, occ_o)
    assert 0.01 < occ_o.mean() < 0.99
    assert_image_match(fb_g, fb_o)
    e.close()


def test_config3_heightfield_graph_frames():
    sc, keys = scenes.heightfield_keyframes()
    assert sc.nt == 1_002_528 and keys.shape == (30, 3, 502_681)
    e = Engine(seed=SEED)
    e.set_keyframes(keys, sc.idx)
    e.upload_materials(sc.materials, sc.mat_idx)
    e.set_grid(GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    e.set_camera_dict(sc.camera)
    e.set_lights(sc.lights)
    cam = ocam(sc.camera)
    for i, cur in enumerate((3.3, 17.8)):
        k1, k2, c = O.frame_ids(cur, 0.25, 30)
        # the first call records the graphs, the second replays them
        for rep in range(2):
            print(f"config 3: frame t={cur} keys=({k1},{k2}) c={c} call {rep}", flush=True)
            fb_g = e.render_frame_graph(k1, k2, c, image_id=0, mode=MODE_PATH, jitter=1, pt_depth=4)
        verts, bmin, bmax = O.lerp_bounds(keys[k1], keys[k2], c)
        m = O.OMesh(verts[0], verts[1], verts[2], sc.idx)
        og = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, 1.0 / 16, 3.0)
        assert_grid_equal(e, og)
        p = O.params(seed=SEED, jitter=1, pt_depth=4, nthreads=16)
        fb_o, _ = O.render_pt(og, cam, p, sc.materials, sc.mat_idx, sc.lights)
        assert fb_o.max() > 0
        assert_image_match(fb_g, fb_o)
    e.close()


def test_config4_batched_workload():
    scs, items = scenes.batched_workload(n_scenes=32, scale=0.05)
    print(f"config 4: {len(scs)} scenes, {sum(f.shape[0] for _, f in scs)} triangles, "
          f"{sum(it[1].shape[1] for it in items)} rays", flush=True)
    b = Batch()
    b.set_scenes(scs)
    rays = b.make_rays(items)
    r = b.autotune(rays)
    lt, ll, tun = b.densities()
    print(f"config 4: autotuned {r.evaluations} evaluations, refill {tun[0]}", flush=True)
    b.trace(rays)
    b.synchronize()
    res = rays.results()
    engines = []
    for s, (v, f) in enumerate(scs):
        e = Engine(seed=SEED)
        e.upload(v, v, 0.0, f)
        e.set_grid(GRID_TWO_LEVEL, float(lt[s]), float(ll[s]))
        e.build_acc_struct()
        gb, cb, lb, pb = b.grid(s)
        ge, ce, le, pe = e.grid()
        assert tuple(gb.res) == tuple(ge.res)
        np.testing.assert_array_equal(cb, ce)
        np.testing.assert_array_equal(lb, le)
        np.testing.assert_array_equal(pb, pe)
        engines.append(e)
    for (s, org, d, anyhit, tmax), out in zip(items, res):
        e = engines[s]
        if anyhit:
            np.testing.assert_array_equal(out[0], e.occluded(org, d, tmax))
        else:
            t, p = e.trace(org, d)
            np.testing.assert_array_equal(out[0], t)
            np.testing.assert_array_equal(out[1], p)
    # three scenes (smallest, a middle one, largest) against the oracle
    sizes = [f.shape[0] for _, f in scs]
    for s in (int(np.argmin(sizes)), len(scs) // 2, int(np.argmax(sizes))):
        v, f = scs[s]
        m = O.OMesh(v[0], v[1], v[2], f)
        bmin, bmax = O.bounds(v)
        og = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, float(lt[s]), float(ll[s]))
        _, cells, leaves, prims = b.grid(s)
        np.testing.assert_array_equal(cells, og.cells())
        np.testing.assert_array_equal(leaves, og.leaves())
        np.testing.assert_array_equal(prims, og.prims())
        for (s2, org, d, anyhit, tmax), out in zip(items, res):
            if s2 != s:
                continue
            if anyhit:
                _, occ, _ = O.trace(og, org, d, anyhit=True, tmax=tmax, nthreads=16)
                np.testing.assert_array_equal(out[0], occ.astype(np.uint8))
            else:
                t, p, _ = O.trace(og, org, d, nthreads=16)
                np.testing.assert_array_equal(out[0], t)
                np.testing.assert_array_equal(out[1], p)
    rays.free()
    for e in engines:
        e.close()
    b.close()


def served_requests(n_scenes=8, n_requests=200):
    scs, pools, reqs = scenes.served_workload(n_scenes=n_scenes, n_requests=n_requests,
                                              ray_range=(1000, 200000), pool=65536)
    rng = np.random.default_rng(77)
    out = []
    for (t, kind, s, n, _, k1, k2, c) in reqs:
        if kind == "rebuild":
            out.append(("rebuild", s, k1, k2, c))
            continue
        po, pd, ao, ad, tmax = pools[s]
        sel = (int(rng.integers(0, po.shape[1])) + np.arange(n)) % po.shape[1]
        if kind == "occluded":
            out.append((kind, s, ao[:, sel].copy(), ad[:, sel].copy(), tmax))
        else:
            out.append((kind, s, po[:, sel].copy(), pd[:, sel].copy(), 0.0))
    return scs, out


def _submit(srv, r):
    if r[0] == "rebuild":
        return srv.submit_rebuild(r[1], r[2], r[3], r[4])
    if r[0] == "closest":
        return srv.submit_closest(r[1], r[2], r[3])
    if r[0] == "occluded":
        return srv.submit_occluded(r[1], r[2], r[3], r[4])
    return srv.submit_path(r[1], r[2], r[3], depth=4)


def _out(p):
    o = p.wait()
    if o is None:
        return None
    return tuple(x.copy() for x in o) if isinstance(o, tuple) else o.copy()


def _same(a, b):
    if a is None:
        assert b is None
        return
    if isinstance(a, tuple):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    else:
        np.testing.assert_array_equal(a, b)


def test_config5_served_workload_with_eviction():
    scs, reqs = served_requests()
    mat = np.stack([scenes.material((0.5, 0.5, 0.5))])
    ntri = sum(f.shape[0] for _, f, _ in scs)
    print(f"config 5: {len(scs)} scenes, {ntri} triangles, {len(reqs)} requests", flush=True)

    def add(srv):
        for keys, f, light in scs:
            srv.add_scene(keys, f, mat, None, light[None])

    # alone: every request waits for the previous one, no budget pressure
    srv = Server(seed=SEED)
    add(srv)
    alone = [_out(_submit(srv, r)) for r in reqs]
    st = srv.stats()
    peak = st.peak_resident_bytes
    srv.close()
    # in the mix: two client threads, a budget of 3/4 of the alone run's peak
    budget = int(0.75 * peak)
    print(f"config 5: alone peak {peak / 2**20:.0f} MiB, mix budget {budget / 2**20:.0f} MiB", flush=True)
    srv = Server(seed=SEED, memory_budget=budget, max_wave_rays=300000)
    add(srv)
    pend = [None] * len(reqs)
    lock = threading.Lock()
    order = iter(range(len(reqs)))

    def client():
        while True:
            with lock:
                i = next(order, None)
                if i is None:
                    return
                pend[i] = _submit(srv, reqs[i])

    th = [threading.Thread(target=client) for _ in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mixed = [_out(p) for p in pend]
    st2 = srv.stats()
    srv.close()
    print(f"config 5: waves alone {st.waves}, mix {st2.waves}; evictions {st2.evictions}", flush=True)
    assert st2.evictions > 0
    assert st2.waves < st.waves
    # single-scene entry points replaying the stream
    engines = []
    for keys, f, light in scs:
        e = Engine(seed=SEED)
        e.set_keyframes(keys, f)
        e.upload_materials(mat)
        e.set_lights(light[None])
        e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
        e.set_frame(0, 0, 0.0)
        e.build_acc_struct()
        engines.append(e)
    for i, r in enumerate(reqs):
        if r[0] == "rebuild":
            engines[r[1]].set_frame(r[2], r[3], r[4])
            engines[r[1]].build_acc_struct()
            ref = None
        elif r[0] == "closest":
            ref = engines[r[1]].trace(r[2], r[3])
        elif r[0] == "occluded":
            ref = engines[r[1]].occluded(r[2], r[3], r[4])
        else:
            ref = engines[r[1]].path_trace(r[2], r[3], depth=4)
        _same(alone[i], mixed[i])
        _same(alone[i], ref)
    for e in engines:
        e.close()
