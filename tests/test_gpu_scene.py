"""GPU parity of scenes read from files (hr_scene_load + hr_scene_upload):
an animated key-frame scene with per-corner normals, a camera.cfg and a
light.cfg is stepped through AnimationManager time; every frame's grid, AO
image and path-traced image (shaded with interpolated vertex normals) match
the oracle on the same inputs."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import SEED, assert_grid_equal
from happyray_amd import GRID_TWO_LEVEL, MODE_AO, MODE_PATH, Engine, SceneFile
from scene_files import animated_heightfield_scene

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
IMG_FRAC = 0.999


def assert_image_match(a, b):
    err = np.abs(a - b).max(axis=1)
    frac = (err <= IMG_TOL).mean()
    assert frac >= IMG_FRAC, f"only {frac:.6f} of pixels within {IMG_TOL} (max err {err.max()})"


def test_scene_file_frames_match_oracle(tmp_path):
    cfg = animated_heightfield_scene(str(tmp_path), q=24, frames=3, step=0.4, res=(96, 64))
    sc = SceneFile.load(cfg)
    keys = sc.keyframes()
    idx, nidx, mi = sc.triangles()
    mats = sc.materials()
    lights, weights = sc.lights()
    cam = sc.camera
    ocam = O.camera(cam["pos"], cam["dir"], cam["up"], cam["fov"], *cam["res"])
    lt, ll = 1.2, 2.2
    e = Engine(seed=SEED)
    e.load_scene(sc)
    e.set_grid(GRID_TWO_LEVEL, lt, ll)
    t = 0.0
    for step in range(4):
        k1, k2, c = sc.frame(t)
        print(f"scene frame t={t:.2f} keys=({k1},{k2}) c={c:.3f}", flush=True)
        e.set_frame(k1, k2, c)
        e.build_acc_struct()
        verts, bmin, bmax = O.lerp_bounds(keys[k1], keys[k2], c)
        # normals are uploaded with coefficient 0: n_k1 * 1 + n_k2 * 0
        nrm = sc.normals(k1) * np.float32(1) + sc.normals(k2) * np.float32(0)
        m = O.OMesh(verts[0], verts[1], verts[2], idx, normals=nrm, nidx=nidx)
        og = O.OGrid(m, bmin, bmax, O.GRID_TWO_LEVEL, lt, ll)
        assert_grid_equal(e, og)
        fb_g = e.render_frame(0, MODE_PATH, jitter=1, pt_depth=3)
        p = O.params(seed=SEED, jitter=1, pt_depth=3, nthreads=8)
        fb_o, _ = O.render_pt(og, ocam, p, mats, mi, lights, weights)
        assert fb_o.max() > 0
        assert_image_match(fb_g, fb_o)
        # the interpolated normals change the image (the flat-normal oracle differs)
        flat = O.OMesh(verts[0], verts[1], verts[2], idx)
        fb_flat, _ = O.render_pt(O.OGrid(flat, bmin, bmax, O.GRID_TWO_LEVEL, lt, ll), ocam, p, mats,
                                 mi, lights, weights)
        assert np.abs(fb_flat - fb_o).max() > 1e-3
        ao_g = e.render_frame(0, MODE_AO, jitter=0, ao_samples=4, ao_dist=0.3)
        ao_o, _, _ = O.render_ao(og, ocam, O.params(seed=SEED, jitter=0, ao_samples=4, ao_dist=0.3,
                                                    nthreads=8), mats, mi)
        assert_image_match(ao_g, ao_o)
        t = sc.next_frame(t)
    e.close()


def test_normals_for_uploaded_mesh(tmp_path):
    """hr_set_normals after hr_upload_mesh (two frames): the normals of the
    first frame shade the path tracer; new geometry drops them."""
    cfg = animated_heightfield_scene(str(tmp_path), q=12, frames=2, res=(64, 48))
    sc = SceneFile.load(cfg)
    keys = sc.keyframes()
    idx, nidx, mi = sc.triangles()
    mats = sc.materials()
    lights, weights = sc.lights()
    cam = sc.camera
    ocam = O.camera(cam["pos"], cam["dir"], cam["up"], cam["fov"], *cam["res"])
    e = Engine(seed=SEED)
    e.upload(keys[0], keys[1], 0.3, idx)
    e.set_normals(np.stack([sc.normals(0), sc.normals(1)]), nidx)
    e.upload_materials(mats, mi)
    e.set_lights(lights, weights)
    e.set_camera_dict(cam)
    e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
    e.build_acc_struct()
    fb_g = e.render_frame(0, MODE_PATH, jitter=1, pt_depth=2)
    verts, bmin, bmax = O.lerp_bounds(keys[0], keys[1], 0.3)
    nrm = sc.normals(0) * np.float32(1) + sc.normals(1) * np.float32(0)
    og = O.OGrid(O.OMesh(verts[0], verts[1], verts[2], idx, normals=nrm, nidx=nidx), bmin, bmax,
                 O.GRID_TWO_LEVEL, 1.2, 2.2)
    p = O.params(seed=SEED, jitter=1, pt_depth=2, nthreads=8)
    fb_o, _ = O.render_pt(og, ocam, p, mats, mi, lights, weights)
    assert_image_match(fb_g, fb_o)
    # re-uploading geometry drops the normals: geometric-normal shading again
    e.upload(keys[0], keys[1], 0.3, idx)
    e.build_acc_struct()
    fb_g2 = e.render_frame(0, MODE_PATH, jitter=1, pt_depth=2)
    flat = O.OMesh(verts[0], verts[1], verts[2], idx)
    fb_f, _ = O.render_pt(O.OGrid(flat, bmin, bmax, O.GRID_TWO_LEVEL, 1.2, 2.2), ocam, p, mats, mi,
                          lights, weights)
    assert_image_match(fb_g2, fb_f)
    e.close()
