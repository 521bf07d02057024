"""The headless renderer `happyray scene.cfg -nw` (src/Main.cpp:54-71): with
the render mode and the samples per pixel on standard input it builds frame 0
of the scene, renders the samples into the running mean and writes
screenshot0.png; the file holds the gamma-corrected bytes of the same
renders made through the library (hr_scene_upload, frame 0 with coefficient
1, two-level grid 1.2 / 2.2, image ids 0..spp-1)."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import SEED
from happyray_amd import (GRID_TWO_LEVEL, MODE_AO, MODE_PATH, MODE_PATH_MULTI, MODE_SIMPLE, Engine,
                          SceneFile, image)
from scene_files import animated_heightfield_scene
from test_image import decode_png

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode_in,spp", [(0, 5), (1, 3), (2, 2), (3, 2)])
def test_cli_screenshot_equals_library_render(tmp_path, mode_in, spp):
    cfg = animated_heightfield_scene(str(tmp_path), q=20, frames=2, res=(64, 48))
    print(f"cli mode={mode_in} spp={spp}", flush=True)
    r = subprocess.run([image.CLI_PATH, cfg, "-nw"], input=f"{mode_in}\n{spp}\n", capture_output=True,
                       text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Wrote screenshot0.png" in r.stdout
    shot = decode_png(os.path.join(tmp_path, "screenshot0.png"))
    sc = SceneFile.load(cfg)
    e = Engine(seed=SEED)
    e.load_scene(sc)
    e.set_frame(0, 0, 1.0)
    e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
    e.build_acc_struct()
    gi = e.grid_info()
    diag = np.float32(np.sqrt(np.float32(sum((np.float32(gi.bmax[i]) - np.float32(gi.bmin[i])) ** 2
                                               for i in range(3)))))
    mode = [MODE_SIMPLE, MODE_PATH_MULTI, MODE_PATH, MODE_AO][mode_in]
    fb = None
    for i in range(spp):
        fb = e.render_frame(i, mode, jitter=1, ao_samples=4, ao_dist=float(diag * np.float32(0.05)), pt_depth=64)
    w, h = sc.camera["res"]
    assert fb.max() > 0
    np.testing.assert_array_equal(shot, image.to_rgb8(fb, w, h))
    e.close()


def test_cli_instanced_scene_renders_the_hierarchy(tmp_path):
    """A scene with an instances file renders through the instanced hierarchy
    (ambient occlusion whatever the mode asked), as the reference's
    TLGRIDHIERARCHY build does."""
    from happyray_amd import GRID_HIERARCHY, scenes
    from scene_files import instanced_scene_files
    sc, ins = scenes.instanced_scene(seed=21, n_inst=24, n_soup=100, res=(48, 32))
    cfg = instanced_scene_files(str(tmp_path), sc, ins)
    print("cli instanced scene", flush=True)
    r = subprocess.run([image.CLI_PATH, cfg, "-nw"], input="2\n3\n", capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "ambient occlusion" in r.stdout and "Wrote screenshot0.png" in r.stdout
    shot = decode_png(os.path.join(tmp_path, "screenshot0.png"))
    sf = SceneFile.load(cfg)
    e = Engine(seed=SEED)
    e.load_scene(sf)
    e.set_frame(0, 0, 1.0)
    e.set_grid(GRID_HIERARCHY, 1.2, 2.2)
    e.build_acc_struct()
    gi = e.grid_info()
    diag = np.float32(np.sqrt(np.float32(sum((np.float32(gi.bmax[i]) - np.float32(gi.bmin[i])) ** 2
                                               for i in range(3)))))
    fb = None
    for i in range(3):
        fb = e.render_frame(i, MODE_AO, jitter=1, ao_samples=4, ao_dist=float(diag * np.float32(0.05)))
    w, h = sf.camera["res"]
    assert fb.max() > 0
    np.testing.assert_array_equal(shot, image.to_rgb8(fb, w, h))
    e.close()
