"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer outputs (grids, primitive ids) must be
bit-exact; floating-point outputs use the tolerances of the hot-path
contract: hit t to 1e-5 relative, images to 1e-4 per pixel on >= 99.9 % of
pixels.  The kernels and the oracle evaluate the same IEEE operations, so
in practice these comparisons are exact."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import (SEED, assert_grid_equal, gpu_engine, oracle_grid, random_rays)
from happyray_amd import (Engine, GRID_TWO_LEVEL, GRID_UNIFORM, GRID_UNIFORM_EXACT, MODE_AO,
                          MODE_PATH, scenes)

pytestmark = pytest.mark.gpu

T_RTOL = 1e-5          # hit distance, relative
IMG_TOL = 1e-4         # per pixel, absolute
IMG_FRAC = 0.999       # fraction of pixels within IMG_TOL


@pytest.fixture(scope="module")
def cluster():
    return scenes.icosphere_cluster(res=128)


@pytest.fixture(scope="module")
def box():
    return scenes.small_box()


def assert_hits_match(t_g, p_g, t_o, p_o, nt):
    hit_g = p_g < nt
    hit_o = p_o < nt
    np.testing.assert_array_equal(hit_g, hit_o)
    np.testing.assert_allclose(t_g[hit_o], t_o[hit_o], rtol=T_RTOL, atol=0)
    # ids equal except at equal-t ties (none expected with identical arithmetic)
    diff = p_g != p_o
    assert np.all(t_g[diff] == t_o[diff])
    assert diff.mean() < 1e-4


def assert_image_match(a, b):
    err = np.abs(a - b).max(axis=1)
    frac = (err <= IMG_TOL).mean()
    assert frac >= IMG_FRAC, f"only {frac:.6f} of pixels within {IMG_TOL} (max err {err.max()})"


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_UNIFORM, 5.0, 1.0),
                                        (GRID_TWO_LEVEL, 1.0 / 16, 3.0),
                                        (GRID_TWO_LEVEL, 1.2, 2.2)])
def test_build_parity_cluster(cluster, kind, lt, ll):
    print(f"build parity icosphere_cluster kind={kind} lt={lt} ll={ll}", flush=True)
    e = gpu_engine(cluster, kind, lt, ll)
    og = oracle_grid(cluster, kind, lt, ll)
    assert_grid_equal(e, og)
    e.close()


@pytest.mark.parametrize("kind", [GRID_UNIFORM_EXACT, GRID_TWO_LEVEL])
def test_build_parity_box_and_soup(box, kind):
    soup = scenes.random_soup(7, 3000, extent=10.0, size=1.5)
    for sc in (box, soup):
        print(f"build parity {sc.name} kind={kind}", flush=True)
        e = gpu_engine(sc, kind, 5.0 if kind != GRID_TWO_LEVEL else 0.5, 3.0)
        assert_grid_equal(e, oracle_grid(sc, kind, 5.0 if kind != GRID_TWO_LEVEL else 0.5, 3.0))
        e.close()


def test_build_capacity_regrowth():
    """Few large triangles make far more pairs than the first capacity guess,
    so the build overflows, grows its buffers and runs again."""
    print("build capacity regrowth", flush=True)
    sc = scenes.random_soup(11, 40, extent=10.0, size=8.0)
    for kind, lt, ll in ((GRID_UNIFORM_EXACT, 200.0, 1.0), (GRID_TWO_LEVEL, 50.0, 40.0)):
        e = gpu_engine(sc, kind, lt, ll)
        assert_grid_equal(e, oracle_grid(sc, kind, lt, ll))
        e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 1.0 / 16, 3.0)])
def test_trace_parity(cluster, kind, lt, ll):
    print(f"trace parity kind={kind}", flush=True)
    e = gpu_engine(cluster, kind, lt, ll)
    og = oracle_grid(cluster, kind, lt, ll)
    org, d = random_rays(cluster, 20000, seed=5)
    t_g, p_g = e.trace(org, d)
    t_o, p_o, _ = O.trace(og, org, d, nthreads=8)
    assert_hits_match(t_g, p_g, t_o, p_o, cluster.nt)
    assert (p_g < cluster.nt).mean() > 0.3
    for tmax in (0.5, 3.0, 1e30):
        occ_g = e.occluded(org, d, tmax)
        _, occ_o, _ = O.trace(og, org, d, anyhit=True, tmax=tmax, nthreads=8)
        np.testing.assert_array_equal(occ_g, occ_o.astype(np.uint8))
    e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 1.0 / 16, 3.0)])
@pytest.mark.parametrize("jitter", [0, 1])
def test_primary_and_ao_parity(cluster, kind, lt, ll, jitter):
    print(f"AO parity kind={kind} jitter={jitter}", flush=True)
    e = gpu_engine(cluster, kind, lt, ll)
    e.set_camera_dict(cluster.camera)
    og = oracle_grid(cluster, kind, lt, ll)
    cam = O.camera(**{k: cluster.camera[k] for k in ("pos", "dir", "up", "fov")},
                   res_x=cluster.camera["res"][0], res_y=cluster.camera["res"][1])
    p = O.params(seed=SEED, jitter=jitter, ao_samples=4, ao_dist=2.0, nthreads=8)
    fb_g = e.render_frame(0, MODE_AO, jitter=jitter, ao_samples=4, ao_dist=2.0)
    planes_g, prims_g = e.primary_hits()
    planes_o, prims_o, _ = O.primary(og, cam, p)
    np.testing.assert_array_equal(prims_g, prims_o)
    hit = prims_o < cluster.nt
    np.testing.assert_array_equal(planes_g[:, hit], planes_o[:, hit])
    fb_o, occ_o, _ = O.render_ao(og, cam, p, cluster.materials, cluster.mat_idx)
    occ_g = e.ao_occlusion(4)
    np.testing.assert_array_equal(occ_g, occ_o)
    assert_image_match(fb_g, fb_o)
    # progressive accumulation over a second image
    fb_g2 = e.render_frame(1, MODE_AO, jitter=jitter, ao_samples=4, ao_dist=2.0)
    fb_o2, _, _ = O.render_ao(og, cam, O.params(seed=SEED, jitter=jitter, ao_samples=4,
                                                 ao_dist=2.0, nthreads=8),
                              cluster.materials, cluster.mat_idx, image_id=1, fb=fb_o)
    assert_image_match(fb_g2, fb_o2)
    e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 1.0 / 16, 3.0)])
def test_odd_resolution_parity(cluster, kind, lt, ll):
    """An image that is not a whole number of 8x8 tiles (camera and AO rays
    are handed to lanes in tiled order): every camera-ray kernel and both AO
    kernels equal the oracle."""
    cam_d = dict(cluster.camera, res=(70, 45))
    og = oracle_grid(cluster, kind, lt, ll)
    cam = O.camera(**{k: cam_d[k] for k in ("pos", "dir", "up", "fov")}, res_x=70, res_y=45)
    p = O.params(seed=SEED, jitter=1, ao_samples=3, ao_dist=2.0, nthreads=8)
    planes_o, prims_o, _ = O.primary(og, cam, p)
    fb_o, occ_o, _ = O.render_ao(og, cam, p, cluster.materials, cluster.mat_idx)
    hit = prims_o < cluster.nt
    e = gpu_engine(cluster, kind, lt, ll)
    e.set_camera_dict(cam_d)
    for pp, pers in ((0, 1), (1, 1), (3, 1), (0, 0)):
        print(f"odd resolution kind={kind} camera kernel={pp} persistent AO={pers}", flush=True)
        e.set_tuning(persistent_primary=pp, persistent_secondary=pers)
        fb_g = e.render_frame(0, MODE_AO, jitter=1, ao_samples=3, ao_dist=2.0)
        planes_g, prims_g = e.primary_hits()
        np.testing.assert_array_equal(prims_g, prims_o)
        np.testing.assert_array_equal(planes_g[:, hit], planes_o[:, hit])
        np.testing.assert_array_equal(e.ao_occlusion(3), occ_o)
        assert_image_match(fb_g, fb_o)
    e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 1.0 / 16, 3.0)])
def test_camera_tiles_from_counter_parity(cluster, kind, lt, ll):
    """Camera rays whose 8x8 tiles are taken from a counter by one workgroup
    per CU (camera_blocks 1: 256 workgroups for the 1,344 tiles of a 333x251
    image) equal the oracle's, as do those of one workgroup per 256 pixels,
    and both count the same rays."""
    cam_d = dict(cluster.camera, res=(333, 251))
    og = oracle_grid(cluster, kind, lt, ll)
    cam = O.camera(**{k: cam_d[k] for k in ("pos", "dir", "up", "fov")}, res_x=333, res_y=251)
    p = O.params(seed=SEED, jitter=1, ao_samples=1, ao_dist=2.0, nthreads=8)
    planes_o, prims_o, _ = O.primary(og, cam, p)
    hit = prims_o < cluster.nt
    e = gpu_engine(cluster, kind, lt, ll)
    e.set_camera_dict(cam_d)
    counts = []
    for blocks in (1, 0):
        print(f"camera tiles kind={kind} camera_blocks={blocks}", flush=True)
        e.set_tuning(persistent_primary=0, camera_blocks=blocks)
        r0 = e.timing().rays_traced
        e.render_frame(0, MODE_AO, jitter=1, ao_samples=1, ao_dist=2.0, download=False)
        counts.append(e.timing().rays_traced - r0)
        planes_g, prims_g = e.primary_hits()
        np.testing.assert_array_equal(prims_g, prims_o)
        np.testing.assert_array_equal(planes_g[:, hit], planes_o[:, hit])
    assert counts[0] == counts[1] >= 333 * 251
    e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0)])
def test_path_tracing_parity(cluster, kind, lt, ll):
    print(f"PT parity kind={kind}", flush=True)
    e = gpu_engine(cluster, kind, lt, ll)
    e.set_camera_dict(cluster.camera)
    e.set_lights(cluster.lights)
    og = oracle_grid(cluster, kind, lt, ll)
    cam = O.camera(**{k: cluster.camera[k] for k in ("pos", "dir", "up", "fov")},
                   res_x=cluster.camera["res"][0], res_y=cluster.camera["res"][1])
    p = O.params(seed=SEED, jitter=1, pt_depth=4, nthreads=8)
    fb_g = e.render_frame(0, MODE_PATH, jitter=1, pt_depth=4)
    fb_o, _ = O.render_pt(og, cam, p, cluster.materials, cluster.mat_idx, cluster.lights)
    assert fb_o.max() > 0
    assert_image_match(fb_g, fb_o)
    e.close()


def mixed_material_cluster(res=128):
    """The cluster with its spheres in turn diffuse, glossy (Phong lobe),
    clear glass and absorbing glass (IOR 1.5), so that paths take every
    branch of the path tracers (PathTracer.h:133-313)."""
    sc = scenes.icosphere_cluster(res=res)
    mats = np.concatenate([sc.materials, np.stack([
        scenes.material((0.05, 0.05, 0.05), ks=(0.25, 0.25, 0.2), spec_exp=30.0),
        scenes.material((0.0, 0.0, 0.0), ior=1.5),
        scenes.material((0.08, 0.02, 0.01), ior=1.5)])]).astype(np.float32)
    nm = len(sc.materials)
    mi = sc.mat_idx.copy()
    per = 5120
    for s in range(16):
        mi[s * per:(s + 1) * per] = (mi[s * per], nm, nm + 1, nm + 2)[s % 4]
    sc.materials, sc.mat_idx = mats, mi
    return sc


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0)])
def test_path_tracing_parity_glass_and_glossy(kind, lt, ll):
    sc = mixed_material_cluster()
    print(f"PT parity glass/glossy kind={kind}", flush=True)
    e = gpu_engine(sc, kind, lt, ll)
    e.set_camera_dict(sc.camera)
    e.set_lights(sc.lights)
    og = oracle_grid(sc, kind, lt, ll)
    cam = O.camera(**{k: sc.camera[k] for k in ("pos", "dir", "up", "fov")},
                   res_x=sc.camera["res"][0], res_y=sc.camera["res"][1])
    for image in (0, 1):
        p = O.params(seed=SEED, jitter=1, pt_depth=8, nthreads=8)
        fb_g = e.render_frame(image, MODE_PATH, jitter=1, pt_depth=8)
        fb_o, _ = O.render_pt(og, cam, p, sc.materials, sc.mat_idx, sc.lights, image_id=image,
                              fb=None if image == 0 else fb_o)
        assert fb_o.max() > 0
        assert_image_match(fb_g, fb_o)
    e.close()


def test_keyframe_interpolation_parity():
    print("keyframe parity", flush=True)
    sc, keys = scenes.heightfield_keyframes(quads=64, frames=5, res=(64, 48))
    e = Engine(seed=SEED)
    e.set_keyframes(keys, sc.idx)
    e.upload_materials(sc.materials, sc.mat_idx)
    e.set_grid(GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    for cur in (0.0, 0.4, 1.7, 3.9):
        k1, k2, c = O.frame_ids(cur, 0.4, 5)
        e.set_frame(k1, k2, c)
        e.build_acc_struct()
        verts, bmin, bmax = O.lerp_bounds(keys[k1], keys[k2], c)
        m = O.OMesh(verts[0], verts[1], verts[2], sc.idx)
        og = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, 1.0 / 16, 3.0)
        assert_grid_equal(e, og)
    e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 1.0 / 16, 3.0)])
def test_persistent_matches_one_ray_per_lane(cluster, kind, lt, ll):
    """The persistent kernels (dynamic ray fetch) and the one-ray-per-lane
    kernels apply the same operations to every ray: bit-identical outputs and
    identical work counters, for several refill thresholds, with per-lane tests
    (one unit of work per loop iteration, or a step and a test fused) and with
    wave-batched tests at 1 to 4 grid steps per iteration."""
    print(f"persistent vs one-ray-per-lane kind={kind}", flush=True)
    e = gpu_engine(cluster, kind, lt, ll)
    e.set_camera_dict(cluster.camera)
    org, d = random_rays(cluster, 30000, seed=17)
    out = {}
    for pers, refill, fuse, batch, steps in ((0, 16, 1, 0, 1), (1, 1, 0, 0, 1), (1, 16, 1, 0, 1), (1, 64, 0, 0, 1),
                                            (1, 48, 1, 0, 1), (1, 56, 1, 1, 1), (1, 56, 1, 1, 2), (1, 1, 1, 1, 4),
                                            (1, 64, 1, 1, 3), (1, 16, 0, 1, 2)):
        e.set_tuning(pers, pers, refill, fuse_units=fuse, test_batch=batch, advance_steps=steps)
        e.work_stats()
        img = e.render_frame(0, MODE_AO, jitter=1, ao_samples=4, ao_dist=2.0, collect_stats=True)
        ws = e.work_stats()
        occ = e.ao_occlusion(4)
        planes, prims = e.primary_hits()
        t, p = e.trace(org, d)
        o = e.occluded(org, d, 1.5)
        out[(pers, refill, fuse, batch, steps)] = (img, occ, planes, prims, t, p, o,
                                                   (ws.primary.rays, ws.primary.steps, ws.primary.tests,
                                                    ws.secondary.rays, ws.secondary.steps, ws.secondary.tests))
    ref = out[(0, 16, 1, 0, 1)]
    for key, val in out.items():
        for a, b in zip(val[:7], ref[:7]):
            np.testing.assert_array_equal(a, b, err_msg=str(key))
        assert val[7] == ref[7], (key, val[7], ref[7])
    e.close()


def test_camera_kernel_choice(cluster):
    """persistent_primary 2 (the default) runs the camera-ray kernels 0, 1, 3,
    0, 1, 3 on its first six timed frames and then keeps the fastest; a camera
    change starts the trials again.  Every frame equals every fixed choice bit
    for bit."""
    print("automatic camera-ray kernel", flush=True)
    e = gpu_engine(cluster, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    e.set_camera_dict(cluster.camera)
    assert e.tuning().persistent_primary == 2
    ref = None
    for pp in (0, 1, 3):
        e.set_tuning(persistent_primary=pp)
        img = e.render_frame(0, MODE_AO, jitter=1, ao_samples=4, ao_dist=2.0)
        assert e.timing().primary_kernel == pp
        got = (img, e.ao_occlusion(4)) + e.primary_hits()
        if ref is None:
            ref = got
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"kernel {pp}")
    e.set_tuning(persistent_primary=2)
    used = []
    for i in range(14):
        if i == 8:
            e.set_camera_dict(cluster.camera)
        img = e.render_frame(0, MODE_AO, jitter=1, ao_samples=4, ao_dist=2.0)
        used.append(e.timing().primary_kernel)
        got = (img, e.ao_occlusion(4)) + e.primary_hits()
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"frame {i}")
    assert used[:6] == [0, 1, 3] * 2 and used[8:14] == [0, 1, 3] * 2, used
    assert used[6] in (0, 1, 3) and used[7] == used[6], used
    assert e.tuning().persistent_primary == 2
    e.close()


def test_work_counters_match_oracle(cluster):
    """DDA steps and triangle tests counted by the kernels equal the oracle's."""
    print("work counters vs oracle", flush=True)
    kind, lt, ll = GRID_TWO_LEVEL, 1.0 / 16, 3.0
    e = gpu_engine(cluster, kind, lt, ll)
    e.set_camera_dict(cluster.camera)
    e.work_stats()
    e.render_frame(0, MODE_AO, jitter=1, ao_samples=4, ao_dist=2.0, collect_stats=True)
    ws = e.work_stats()
    og = oracle_grid(cluster, kind, lt, ll)
    cam = O.camera(**{k: cluster.camera[k] for k in ("pos", "dir", "up", "fov")},
                   res_x=cluster.camera["res"][0], res_y=cluster.camera["res"][1])
    p = O.params(seed=SEED, jitter=1, ao_samples=4, ao_dist=2.0, nthreads=8)
    _, _, sp = O.primary(og, cam, p)
    _, _, sa = O.render_ao(og, cam, p, cluster.materials, cluster.mat_idx)
    assert (ws.primary.rays, ws.primary.steps, ws.primary.tests) == (sp.rays, sp.steps, sp.tests)
    assert (ws.secondary.rays, ws.secondary.steps, ws.secondary.tests) == \
        (sa.rays - sp.rays, sa.steps - sp.steps, sa.tests - sp.tests)
    e.close()


def test_device_pointer_entry_points(cluster):
    """hr_trace_dev / hr_occluded_dev on resident rays equal hr_trace / hr_occluded."""
    from happyray_amd.devmem import DeviceArray
    print("device-pointer trace entry points", flush=True)
    e = gpu_engine(cluster, GRID_TWO_LEVEL, 1.0 / 16, 3.0)
    org, d = random_rays(cluster, 5000, seed=23)
    o_d, d_d = DeviceArray.from_host(org), DeviceArray.from_host(d)
    t_d, p_d = DeviceArray((5000,), np.float32), DeviceArray((5000,), np.uint32)
    occ_d = DeviceArray((5000,), np.uint8)
    e.trace_dev(o_d.value, d_d.value, 5000, t_d.value, p_d.value)
    e.occluded_dev(o_d.value, d_d.value, 5000, 2.5, occ_d.value)
    e.synchronize()
    t, p = e.trace(org, d)
    np.testing.assert_array_equal(t_d.to_host(), t)
    np.testing.assert_array_equal(p_d.to_host(), p)
    np.testing.assert_array_equal(occ_d.to_host(), e.occluded(org, d, 2.5))
    for a in (o_d, d_d, t_d, p_d, occ_d):
        a.free()
    e.close()


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0)])
def test_persistent_paths_match_one_path_per_lane(kind, lt, ll):
    """The persistent path tracer (dynamic pixel fetch, batched shading) and
    the one-path-per-lane megakernel apply the same operations to every path:
    bit-identical images over diffuse, glossy and glass materials, identical
    work counters, for several refill / shading thresholds, and the oracle's
    image; paths from given rays (hr_path_trace) likewise."""
    sc = mixed_material_cluster(res=96)
    print(f"persistent paths vs one path per lane kind={kind}", flush=True)
    e = gpu_engine(sc, kind, lt, ll)
    e.set_camera_dict(sc.camera)
    e.set_lights(sc.lights)
    org, d = random_rays(sc, 20000, seed=41)
    out = {}
    for pers, refill, shade in ((0, 32, 16), (1, 1, 1), (1, 32, 16), (1, 64, 64), (1, 48, 8)):
        e.set_tuning(persistent_path=pers, path_refill=refill, path_shade=shade)
        e.work_stats()
        imgs = [e.render_frame(i, MODE_PATH, jitter=1, pt_depth=6, collect_stats=True) for i in (0, 1)]
        ws = e.work_stats()
        rgb = e.path_trace(org, d, depth=5)
        out[(pers, refill, shade)] = (imgs[0], imgs[1], rgb, (ws.secondary.rays, ws.secondary.steps,
                                                              ws.secondary.tests))
    ref = out[(0, 32, 16)]
    for key, val in out.items():
        for a, b in zip(val[:3], ref[:3]):
            np.testing.assert_array_equal(a, b, err_msg=str(key))
        assert val[3] == ref[3], (key, val[3], ref[3])
    og = oracle_grid(sc, kind, lt, ll)
    cam = O.camera(**{k: sc.camera[k] for k in ("pos", "dir", "up", "fov")},
                   res_x=sc.camera["res"][0], res_y=sc.camera["res"][1])
    fb_o, _ = O.render_pt(og, cam, O.params(seed=SEED, jitter=1, pt_depth=6, nthreads=8), sc.materials,
                          sc.mat_idx, sc.lights)
    assert fb_o.max() > 0
    assert_image_match(ref[0], fb_o)
    e.close()
