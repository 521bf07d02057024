"""Ray records on the GPU (24 bytes per ray, origin xyz then direction xyz;
SimpleRayTraverser.h:200-262, AORayExporter.h:50-68): record entry points
equal the SoA ones, input rays through the simple integrator and the AO ray
dump equal the oracle, and server requests carrying records equal the same
requests carrying SoA planes."""
import numpy as np
import pytest

import pyoracle as O
from gpu_util import SEED, gpu_engine, oracle_grid, random_rays
from happyray_amd import GRID_TWO_LEVEL, GRID_UNIFORM_EXACT, MODE_AO, Server, rayfile, scenes
from happyray_amd.server import REQ_CLOSEST, REQ_OCCLUDED, REQ_PATH

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,lt,ll", [(GRID_UNIFORM_EXACT, 5.0, 1.0), (GRID_TWO_LEVEL, 0.5, 3.0)])
def test_records_equal_planes_and_oracle(kind, lt, ll):
    sc = scenes.icosphere_cluster(res=64)
    print(f"ray records kind={kind}", flush=True)
    e = gpu_engine(sc, kind, lt, ll)
    og = oracle_grid(sc, kind, lt, ll)
    org, d = random_rays(sc, 20000, seed=17)
    rec = rayfile.to_records(org, d)
    t0, p0 = e.trace(org, d)
    t1, p1 = e.trace_records(rec)
    np.testing.assert_array_equal(t0, t1)
    np.testing.assert_array_equal(p0, p1)
    ext = float((sc.verts.max(1) - sc.verts.min(1)).max())
    np.testing.assert_array_equal(e.occluded(org, d, 0.1 * ext), e.occluded_records(rec, 0.1 * ext))
    rgb = e.render_rays(rec)
    rgb_o, _ = O.render_rays_simple(og, rec, nthreads=8)
    np.testing.assert_array_equal(rgb, rgb_o)
    # the AO dump of a jittered image 2
    e.set_camera_dict(sc.camera)
    e.render_frame(2, MODE_AO, jitter=1, ao_samples=3, ao_dist=1.5)
    dump = e.dump_ao_rays(3)
    cam = O.camera(sc.camera["pos"], sc.camera["dir"], sc.camera["up"], sc.camera["fov"], *sc.camera["res"])
    dump_o = O.dump_ao_rays(og, cam, O.params(seed=SEED, jitter=1, ao_samples=3, ao_dist=1.5, nthreads=8), 2)
    np.testing.assert_array_equal(dump, dump_o)
    e.close()


def test_server_record_payloads():
    sc = scenes.icosphere_cluster(res=48)
    srv = Server(seed=SEED)
    sid = srv.add_scene(sc.verts[None].copy(), sc.idx, sc.materials, sc.mat_idx, sc.lights)
    org, d = random_rays(sc, 5000, seed=23)
    rec = rayfile.to_records(org, d)
    ext = float((sc.verts.max(1) - sc.verts.min(1)).max())
    a = [srv.submit_closest(sid, org, d), srv.submit_occluded(sid, org, d, 0.2 * ext),
         srv.submit_path(sid, org, d, depth=3)]
    b = [srv.submit_records(REQ_CLOSEST, sid, rec), srv.submit_records(REQ_OCCLUDED, sid, rec, tmax=0.2 * ext),
         srv.submit_records(REQ_PATH, sid, rec, depth=3)]
    for x, y in zip(a, b):
        ox, oy = x.wait(), y.wait()
        if isinstance(ox, tuple):
            for u, v in zip(ox, oy):
                np.testing.assert_array_equal(u, v)
        else:
            np.testing.assert_array_equal(ox, oy)
    srv.close()
