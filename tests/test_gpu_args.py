"""Argument and state checks of the C ABI that return before any kernel runs:
frame selection after a new mesh replaced the resident key frames, and AO
frames whose ray ids would not fit 32 bits."""
import numpy as np
import pytest

from happyray_amd import Engine, GRID_TWO_LEVEL, HappyRayError, MODE_AO, scenes
from happyray_amd._lib import STATUS_NAMES

pytestmark = pytest.mark.gpu

ERR = {v: k for k, v in STATUS_NAMES.items()}


def test_upload_mesh_drops_resident_key_frames():
    sc, keys = scenes.heightfield_keyframes(quads=8, frames=4, res=(8, 8))
    big = scenes.random_soup(3, 500, extent=10.0, size=0.5)
    assert big.nv > keys.shape[2]
    e = Engine()
    e.set_keyframes(keys, sc.idx)
    e.set_frame(3, 2, 0.5)
    e.upload(big.verts, big.verts, 0.0, big.idx)
    for k in ((0, 0), (3, 3), (1, 2)):
        with pytest.raises(HappyRayError) as ex:
            e.set_frame(k[0], k[1], 0.0)
        assert ex.value.status == ERR["HR_ERR_STATE"]
    e.set_camera_dict(sc.camera)
    with pytest.raises(HappyRayError) as ex:
        e.render_frame_graph(1, 2, 0.5, image_id=0, mode=MODE_AO)
    assert ex.value.status == ERR["HR_ERR_STATE"]
    # the new mesh itself builds and renders
    e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
    e.build_acc_struct()
    img = e.render_frame(0, MODE_AO, ao_samples=2, ao_dist=1.0)
    assert img.shape == (64, 3)
    e.close()


def test_ao_ray_count_must_fit_32_bits():
    sc = scenes.small_box()
    e = Engine()
    e.upload_scene(sc)
    e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
    e.build_acc_struct()
    e.set_camera(sc.camera["pos"], sc.camera["dir"], sc.camera["up"], sc.camera["fov"], 16384, 16384)
    for s in (16, 17, 256):
        with pytest.raises(HappyRayError) as ex:
            e.render_frame(0, MODE_AO, ao_samples=s, ao_dist=0.5, download=False)
        assert ex.value.status == ERR["HR_ERR_ARG"]
        assert "2^32" in str(ex.value)
    e.close()
