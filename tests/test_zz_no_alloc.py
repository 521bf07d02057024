"""hr_path_trace_dev_noalloc (the allocation-free path-trace entry point)
against hr_path_trace_dev on the same device buffers: bit-identical radiance
on two calls in a row, the second with fewer rays, with uploaded materials and
with the default material, and the size limit as an argument error."""
import numpy as np
import pytest

from gpu_util import SEED, random_rays
from happyray_amd import Engine, GRID_TWO_LEVEL, GRID_UNIFORM_EXACT, HappyRayError, scenes
from happyray_amd._lib import STATUS_NAMES
from happyray_amd.devmem import DeviceArray

pytestmark = pytest.mark.gpu

ERR = {v: k for k, v in STATUS_NAMES.items()}
MAX_RAYS = 1 << 21


def _run(e, org, d, depth, noalloc):
    n = org.shape[1]
    o_d, d_d = DeviceArray.from_host(org), DeviceArray.from_host(d)
    rgb_d = DeviceArray((n, 3), np.float32)
    if noalloc:
        e.path_trace_dev_noalloc(o_d.value, d_d.value, n, depth, rgb_d.value)
    else:
        e.path_trace_dev(o_d.value, d_d.value, n, depth, rgb_d.value)
    e.synchronize()
    out = rgb_d.to_host()
    for a in (o_d, d_d, rgb_d):
        a.free()
    return out


@pytest.mark.parametrize("kind,lt,ll,materials", [(GRID_TWO_LEVEL, 1.0 / 16, 3.0, True),
                                                  (GRID_UNIFORM_EXACT, 5.0, 1.0, False)])
def test_noalloc_path_trace_equals_path_trace_dev(kind, lt, ll, materials):
    sc = scenes.icosphere_cluster(res=64)
    print(f"no-alloc path trace kind={kind} materials={materials}", flush=True)
    e = Engine(seed=SEED)
    if materials:
        e.upload_scene(sc)
    else:
        e.upload(sc.verts, sc.verts, 0.0, sc.idx)
    e.set_lights(sc.lights)
    e.set_grid(kind, lt, ll)
    e.build_acc_struct()
    for n, seed in ((40000, 31), (7001, 32)):
        org, d = random_rays(sc, n, seed=seed)
        print(f"  call with n={n}", flush=True)
        ref = _run(e, org, d, 4, noalloc=False)
        got = _run(e, org, d, 4, noalloc=True)
        assert ref.max() > 0
        np.testing.assert_array_equal(got, ref)
    e.close()


def test_noalloc_rejects_more_rays_than_its_scratch():
    sc = scenes.small_box()
    e = Engine(seed=SEED)
    e.upload_scene(sc)
    e.set_grid(GRID_TWO_LEVEL, 1.2, 2.2)
    e.build_acc_struct()
    buf = DeviceArray((16,), np.float32)
    with pytest.raises(HappyRayError) as ex:
        e.path_trace_dev_noalloc(buf.value, buf.value, MAX_RAYS + 1, 4, buf.value)
    assert ex.value.status == ERR["HR_ERR_ARG"]
    buf.free()
    e.close()
