"""Scene files through the C ABI (hr_scene_*), on the CPU: scene.cfg keys and
frame naming, OBJ fans / normals / materials, MTL Phong set-up, camera.cfg
defaults and rotation, light.cfg, light derivation from emitters, the default
light and AnimationManager frame stepping.  Expected values are written from
the reference's rules and closed forms, not from the loader's output."""
import math
import os

import numpy as np
import pytest

import pyoracle as O
from scene_files import (animated_heightfield_scene, write_camera, write_cfg, write_instances,
                         write_lights, write_mtl, write_obj)

PI = np.float32(math.pi)
SQ = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]


@pytest.fixture(scope="module")
def hrlib(hr):
    from happyray_amd import SceneFile, _lib
    return SceneFile, _lib


def load(hrlib, path):
    SceneFile, _ = hrlib
    return SceneFile.load(str(path))


def status_of(hrlib, fn):
    from happyray_amd import HappyRayError
    with pytest.raises(HappyRayError) as ei:
        fn()
    return ei.value.status


def phong(kd, ks, ns):
    """Material::setupPhongCoefficients in float32 (WFObjectReader.cpp:49-57)."""
    kd = np.float32(kd)
    ks = np.float32(ks)
    c099 = np.float32(0.99)
    ks2 = np.minimum(ks, c099 - np.minimum(c099, kd)) * (np.float32(ns) + np.float32(2)) \
        * np.float32(0.5) / PI
    return np.minimum(kd, c099) / PI, ks2


def test_obj_fan_normals_and_materials(hrlib, tmp_path):
    write_mtl(tmp_path / "m.mtl", {"red": {"Kd": (0.8, 0.1, 0.1), "Ks": (0.5, 0.5, 0.5), "Ns": 20}})
    write_obj(tmp_path / "a.obj", SQ, [[(1, 1), (2, 1), (3, 1), (4, 1)], [1, 2, 4]],
              normals=[(0, 0, 1)], mtllib="m.mtl", groups=[(0, "red"), (1, "Default")],
              extra=["v 1 2 3 4", "# comment", "vt 0.5 0.5", "g group", "s 1"])
    write_cfg(tmp_path / "s.cfg", [("obj", "a.obj")])
    sc = load(hrlib, tmp_path / "s.cfg")
    # the 4-coordinate vertex line is skipped (exactly three coordinates); no
    # emitter and no light file: the default light's quad follows (4 + 4 vertices)
    assert sc.info.num_vertices == 8 and sc.info.num_triangles == 5
    np.testing.assert_array_equal(sc.vertices(0)[:, :4], np.array(SQ, np.float32).T)
    idx, nidx, mi = sc.triangles()
    # quad -> fan (c0, c1, c2), (c0, c2, c3); then the triangle without normals
    np.testing.assert_array_equal(idx[:3], [[0, 1, 2], [0, 2, 3], [0, 1, 3]])
    np.testing.assert_array_equal(mi[:3], [1, 1, 0])
    # explicit normals keep their index; the face without one gets its face normal
    np.testing.assert_array_equal(nidx[:2], 0)
    np.testing.assert_array_equal(nidx[2], [1, 1, 1])
    np.testing.assert_allclose(sc.normals(0)[:, 1], [0, 0, 1])
    mats = sc.materials()
    assert mats.shape == (3, 12)
    kd, ks = phong([0.6774] * 3, [0, 0, 0], 1)
    np.testing.assert_array_equal(mats[0], np.r_[kd, 1.0, ks, 1.0, 0, 0, 0, 0].astype(np.float32))
    kd, ks = phong([0.8, 0.1, 0.1], [0.5] * 3, 20)
    np.testing.assert_array_equal(mats[1], np.r_[kd, 1.0, ks, 20.0, 0, 0, 0, 0].astype(np.float32))
    # closed forms: kd/pi, and the specular clamp 0.99 - kd = 0.19 on the red channel
    np.testing.assert_allclose(mats[1, 0], 0.8 / math.pi, rtol=1e-6)
    np.testing.assert_allclose(mats[1, 4], 0.19 * 22 * 0.5 / math.pi, rtol=1e-5)
    np.testing.assert_allclose(mats[1, 5], 0.5 * 22 * 0.5 / math.pi, rtol=1e-6)


def test_mtl_rules(hrlib, tmp_path):
    """Materials the OBJ never selects are errors that leave the previous
    material current; KD (upper-case second letter) assigns nothing; illum > 3
    marks the material refractive (index of refraction + 10 in the table)."""
    write_mtl(tmp_path / "m.mtl", {"glass": {"Kd": (0.1, 0.1, 0.1), "Ni": 1.5, "illum": 7},
                                   "unused": {"Kd": (0.0, 0.9, 0.0), "KD": (0.3, 0.3, 0.3)}})
    write_obj(tmp_path / "a.obj", SQ[:3], [[1, 2, 3]], mtllib="m.mtl", groups=[(0, "glass")])
    SceneFile, _ = hrlib
    sc = SceneFile.load_obj(str(tmp_path / "a.obj"))
    m = sc.materials()
    kd, _ = phong([0.0, 0.9, 0.0], [0, 0, 0], 1)
    np.testing.assert_array_equal(m[1, :3], kd)
    assert m[1, 3] == np.float32(11.5)


def test_frame_file_names(hrlib, tmp_path):
    """prefix + index zero-padded to the digits of num_frames - 1 + suffix."""
    for n in (1, 3, 12):
        d = tmp_path / f"n{n}"
        d.mkdir()
        width = len(str(n - 1)) if n > 1 else 0
        for i in range(n):
            write_obj(d / ("k" + str(i).zfill(width) + ".obj"), SQ[:3], [[1, 2, 3]])
        write_cfg(d / "s.cfg", [("framename", "k"), ("framesuffix", ".obj"), ("framenumber", n)])
        sc = load(hrlib, d / "s.cfg")
        assert sc.num_frames == n
        assert sc.info.frame_step == np.float32(0.8)     # AnimationManager default step


def test_config_keys_and_comments(hrlib, tmp_path):
    write_obj(tmp_path / "k0.obj", SQ[:3], [[1, 2, 3]])
    write_obj(tmp_path / "k1.obj", [(0, 0, 1), (1, 0, 1), (1, 1, 1)], [[1, 2, 3]])
    write_cfg(tmp_path / "s.cfg", ["# comment line", "frame prefix=k", "frame suffix = .obj  # trailing",
                                   "number of frames = 2", "step size = 0.25", "",
                                   "unknown token = 3"])
    sc = load(hrlib, tmp_path / "s.cfg")
    assert sc.num_frames == 2 and sc.info.frame_step == np.float32(0.25)
    np.testing.assert_array_equal(sc.vertices(1)[2, :3], [1, 1, 1])
    assert sc.camera is None


def test_camera_defaults_and_rotation(hrlib, tmp_path):
    write_obj(tmp_path / "a.obj", SQ[:3], [[1, 2, 3]])
    # no keys at all: CameraManager defaults
    (tmp_path / "empty.cfg").write_text("# nothing\n")
    write_cfg(tmp_path / "s0.cfg", [("obj", "a.obj"), ("camera", "empty.cfg")])
    cam = load(hrlib, tmp_path / "s0.cfg").camera
    assert cam["res"] == (512, 512) and cam["fov"] == 66.0
    assert cam["dir"] == [0, 0, 1] and cam["up"] == [0, -1, 0] and cam["pos"] == [0, 0, 0]
    # 90 degrees about y turns +z into -x and keeps +y
    write_camera(tmp_path / "c1.cfg", (1, 2, 3), (0, 0, 1), (0, 1, 0), fov=45, res=(64, 32),
                 rotation=(0, 90, 0))
    write_cfg(tmp_path / "s1.cfg", [("obj", "a.obj"), ("cam", "c1.cfg")])
    cam = load(hrlib, tmp_path / "s1.cfg").camera
    np.testing.assert_allclose(cam["dir"], [-1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(cam["up"], [0, 1, 0], atol=1e-6)
    assert cam["res"] == (64, 32) and cam["fov"] == 45.0 and cam["pos"] == [1, 2, 3]
    # the reference's second column has z = sx*sy*sz - sx*cz (a rotation matrix
    # has cx*sy*sz - sx*cz): for (0, 90, 90) the up vector's z stays 0, not 1
    write_camera(tmp_path / "c2.cfg", (0, 0, 0), (0, 0, 1), (0, 1, 0), rotation=(0, 90, 90))
    write_cfg(tmp_path / "s2.cfg", [("obj", "a.obj"), ("view", "c2.cfg")])
    cam = load(hrlib, tmp_path / "s2.cfg").camera
    assert abs(cam["up"][2]) < 1e-6


TWO_LIGHTS = [
    {"position": (0, 5, 0), "intensity": (1, 2, 2), "edge1": (2, 0, 0), "edge2": (0, 0, 3)},
    {"position": (1, 5, 1), "normal": (0, 1, 0), "intensity": (3, 0, 4), "edge1": (1, 0, 0),
     "edge2": (0, 0, 1)}]


def test_light_file_blank_line_repeats_previous_key(hrlib, tmp_path):
    """A line without a key keeps the previous line's key (LightSourceLoader.cpp:58-62
    reads into a string that survives the loop): the blank line after the first light
    marks edge2 as read again, so the second light is complete at its edge1 line and
    keeps the first light's edge2."""
    write_obj(tmp_path / "a.obj", SQ, [[1, 2, 3, 4]])
    write_lights(tmp_path / "l.cfg", TWO_LIGHTS, blank_lines=True)
    write_cfg(tmp_path / "s.cfg", [("obj", "a.obj"), ("lights", "l.cfg")])
    L, w = load(hrlib, tmp_path / "s.cfg").lights()
    assert len(L) == 2
    np.testing.assert_array_equal(L[1, 12:15], [0, 0, 3])
    np.testing.assert_allclose(w, [6 * 3, 3 * 5], rtol=1e-6)


def test_light_file(hrlib, tmp_path):
    write_obj(tmp_path / "a.obj", SQ, [[1, 2, 3, 4]])
    write_lights(tmp_path / "l.cfg", TWO_LIGHTS)
    write_cfg(tmp_path / "s.cfg", [("obj", "a.obj"), ("lights", "l.cfg")])
    sc = load(hrlib, tmp_path / "s.cfg")
    L, w = sc.lights()
    assert len(L) == 2
    # missing normal -> normalize(edge1 x edge2) = (0, -1, 0); weight = area * |I|
    np.testing.assert_array_equal(L[0, 3:6], [0, -1, 0])
    np.testing.assert_allclose(w, [6 * 3, 1 * 5], rtol=1e-6)
    np.testing.assert_array_equal(L[1, 3:6], [0, 1, 0])
    # the first light's quad joins the geometry; the second sees an emitter and is skipped
    assert sc.info.num_triangles == 2 + 2
    assert sc.info.num_vertices == 4 + 4
    v = sc.vertices(0)
    np.testing.assert_array_equal(v[:, 4:].T, [(0, 5, 0), (2, 5, 0), (2, 5, 3), (0, 5, 3)])
    idx, nidx, mi = sc.triangles()
    np.testing.assert_array_equal(idx[2:], [[4, 5, 6], [4, 6, 7]])
    m = sc.materials()
    np.testing.assert_array_equal(m[mi[2], 8:11], [1, 2, 2])


def test_lights_from_emissive_quads(hrlib, tmp_path):
    write_mtl(tmp_path / "m.mtl", {"lamp": {"Ke": (4, 4, 4)}})
    verts = SQ + [(0, 0, 2), (3, 0, 2), (3, 2, 2)]
    write_obj(tmp_path / "a.obj", verts, [[1, 2, 3], [1, 3, 4], [5, 6, 7]], normals=[(0, 0, -1)],
              mtllib="m.mtl", groups=[(0, "lamp"), (2, "Default")])
    sc = hrlib[0].load_obj(str(tmp_path / "a.obj"))
    L, w = sc.lights()
    assert len(L) == 1
    # pairing (0,1,2)+(0,2,3): position V1, edge1 = V0 - V1, edge2 = V2 - V1 (the unit square)
    np.testing.assert_array_equal(L[0, 0:3], [1, 0, 0])
    np.testing.assert_array_equal(L[0, 9:12], [-1, 0, 0])
    np.testing.assert_array_equal(L[0, 12:15], [0, 1, 0])
    np.testing.assert_array_equal(L[0, 6:9], [4, 4, 4])
    np.testing.assert_allclose(w, [math.sqrt(48)], rtol=1e-6)
    assert sc.info.num_triangles == 3       # emitters present: nothing inserted


def test_default_light(hrlib, tmp_path):
    cube = [(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)]
    write_obj(tmp_path / "a.obj", cube, [[1, 2, 4], [5, 6, 8]])
    sc = hrlib[0].load_obj(str(tmp_path / "a.obj"))
    L, w = sc.lights()
    assert len(L) == 1
    lo, hi = np.float32(0.025), np.float32(1) - np.float32(0.025)
    np.testing.assert_array_equal(L[0, 0:3], [lo, lo, hi])
    np.testing.assert_array_equal(L[0, 3:6], [0, 0, -1])
    np.testing.assert_array_equal(L[0, 6:9], [1, 1, 1])
    np.testing.assert_array_equal(L[0, 9:12], [0, hi - lo, 0])
    np.testing.assert_array_equal(L[0, 12:15], [hi - lo, hi - lo, 0])
    assert sc.info.num_triangles == 4 and sc.info.num_materials == 2


def test_animation_frames_follow_animation_manager(hrlib, tmp_path):
    cfg = animated_heightfield_scene(str(tmp_path), q=4, frames=3, step=0.4)
    sc = load(hrlib, cfg)
    assert sc.num_frames == 3 and sc.info.frame_step == np.float32(0.4)
    t = 0.0
    for _ in range(12):
        k1, k2, c = sc.frame(t)
        assert (k1, k2, c) == O.frame_ids(t, 0.4, 3)
        t2 = sc.next_frame(t)
        expect = np.float32(t) + np.float32(0.4)
        if int(expect) >= 3:
            expect -= np.float32(3)
        assert np.float32(t2) == expect
        t = t2
    # a single OBJ is not interpolatable: k2 = k1 and time advances by whole frames
    write_obj(tmp_path / "one.obj", SQ[:3], [[1, 2, 3]])
    one = hrlib[0].load_obj(str(tmp_path / "one.obj"))
    assert one.frame(0.0) == (0, 0, 0.0)
    assert one.next_frame(0.0) == 0.0


def test_scene_errors(hrlib, tmp_path):
    from happyray_amd._lib import STATUS_NAMES
    SceneFile, _ = hrlib
    names = {v: k for k, v in STATUS_NAMES.items()}
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "missing.cfg"))) == names["HR_ERR_IO"]
    write_cfg(tmp_path / "none.cfg", [("camera", "c.cfg")])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "none.cfg"))) == names["HR_ERR_PARSE"]
    write_obj(tmp_path / "bad.obj", SQ[:3], [[1, 2, 9]])
    write_cfg(tmp_path / "bad.cfg", [("obj", "bad.obj")])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "bad.cfg"))) == names["HR_ERR_PARSE"]
    write_obj(tmp_path / "k0.obj", SQ[:3], [[1, 2, 3]])
    write_obj(tmp_path / "k1.obj", SQ, [[1, 2, 3]])
    write_cfg(tmp_path / "mix.cfg", [("frame prefix", "k"), ("frame suffix", ".obj"), ("number", 2)])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "mix.cfg"))) == names["HR_ERR_PARSE"]
    write_obj(tmp_path / "nomtl.obj", SQ[:3], [[1, 2, 3]], mtllib="absent.mtl")
    assert status_of(hrlib, lambda: SceneFile.load_obj(str(tmp_path / "nomtl.obj"))) == names["HR_ERR_IO"]
    assert status_of(hrlib, lambda: load(hrlib, cfg_for_frame_error(tmp_path)).frame(5.0)) == \
        names["HR_ERR_ARG"]


def cfg_for_frame_error(tmp_path):
    write_obj(tmp_path / "one.obj", SQ[:3], [[1, 2, 3]])
    write_cfg(tmp_path / "one.cfg", [("scene", "one.obj")])
    return tmp_path / "one.cfg"


def test_scene_vertices_and_normals_roundtrip(hrlib, tmp_path):
    cfg = animated_heightfield_scene(str(tmp_path), q=3, frames=2)
    sc = load(hrlib, cfg)
    from scene_files import heightfield
    for f in range(2):
        v, n, faces = heightfield(3, 0.7 * f)
        got = sc.vertices(f)
        np.testing.assert_array_equal(got[:, :len(v)], v.T.astype(np.float32))
        np.testing.assert_array_equal(sc.normals(f)[:, :len(n)], n.T.astype(np.float32))
    idx, nidx, mi = sc.triangles()
    assert idx.shape == (2 * 9 + 2, 3)                  # + the light quad
    np.testing.assert_array_equal(idx[:18], nidx[:18])  # corners written as v//v
    assert set(mi[:9]) == {1} and set(mi[9:18]) == {2}


def test_objects_and_instances_file(hrlib, tmp_path):
    """'o' lines split the faces into objects (an empty object is not recorded,
    the last one always closes); an instances file sets each instance's rows and
    box, a missing box stays unbounded (WFObjectReader.cpp:516-583, :762-824)."""
    faces = [[1, 2, 3], [1, 3, 4], [1, 2, 4], [2, 3, 4], [1, 2, 3]]
    # objects start before faces 0 (twice: the first is empty), 2 and 4
    write_obj(tmp_path / "o.obj", SQ, faces, objects=[0, 0, 2, 4])
    m0 = [[1, 0, 0, 0.5], [0, 2, 0, -1], [0, 0, 0.5, 2]]
    m1 = [[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 1, 0.25]]
    write_instances(tmp_path / "inst.txt", [2, 0], [m0, m1])
    with open(tmp_path / "inst.txt", "a") as fh:
        fh.write("AABB_min -1 -2 -3   # trailing comment\nAABB_max 1 2 3\nunknown_key 7\n")
    write_cfg(tmp_path / "s.cfg", [("obj", "o.obj"), ("instances", "inst.txt")])
    sf = load(hrlib, tmp_path / "s.cfg")
    assert sf.info.num_objects == 3 and sf.info.num_instances == 2
    first, objects, m, lo, hi = sf.instances()
    # 5 faces of the file plus the default light quad's 2, outside every object
    assert sf.info.num_triangles == 7
    np.testing.assert_array_equal(first, [0, 2, 4, 5])
    np.testing.assert_array_equal(objects, [2, 0])
    np.testing.assert_array_equal(m[0], np.float32(m0))
    np.testing.assert_array_equal(m[1], np.float32(m1))
    big = np.float32(3.402823466e38)
    np.testing.assert_array_equal(lo[0], [-big] * 3)
    np.testing.assert_array_equal(hi[0], [big] * 3)
    np.testing.assert_array_equal(lo[1], [-1, -2, -3])
    np.testing.assert_array_equal(hi[1], [1, 2, 3])
    # without 'o' lines the whole file is one object
    write_obj(tmp_path / "p.obj", SQ, faces)
    write_cfg(tmp_path / "p.cfg", [("obj", "p.obj")])
    sp = load(hrlib, tmp_path / "p.cfg")
    assert sp.info.num_objects == 1 and sp.info.num_instances == 0
    np.testing.assert_array_equal(sp.instances()[0], [0, 5])


def test_instances_file_errors(hrlib, tmp_path):
    from happyray_amd._lib import STATUS_NAMES
    SceneFile, _ = hrlib
    names = {v: k for k, v in STATUS_NAMES.items()}
    write_obj(tmp_path / "o.obj", SQ, [[1, 2, 3]])
    eye = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]]
    write_instances(tmp_path / "bad_id.txt", [1], [eye])          # one object only
    write_cfg(tmp_path / "a.cfg", [("obj", "o.obj"), ("instances", "bad_id.txt")])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "a.cfg"))) == names["HR_ERR_PARSE"]
    with open(tmp_path / "early.txt", "w") as fh:
        fh.write("m_row_0 1 0 0 0\nobj_id 0\n")
    write_cfg(tmp_path / "b.cfg", [("obj", "o.obj"), ("instances", "early.txt")])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "b.cfg"))) == names["HR_ERR_PARSE"]
    with open(tmp_path / "short.txt", "w") as fh:
        fh.write("obj_id 0\nm_row_1 0 1 0\n")
    write_cfg(tmp_path / "c.cfg", [("obj", "o.obj"), ("instances", "short.txt")])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "c.cfg"))) == names["HR_ERR_PARSE"]
    write_cfg(tmp_path / "d.cfg", [("obj", "o.obj"), ("instances", "absent.txt")])
    assert status_of(hrlib, lambda: SceneFile.load(str(tmp_path / "d.cfg"))) == names["HR_ERR_IO"]
