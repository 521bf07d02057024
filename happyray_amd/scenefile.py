"""Scene files over the C ABI: SceneLoader::loadScene and AnimationManager.

    sc = SceneFile.load("scene.cfg")        # SceneLoader::loadScene
    sc = SceneFile.load_obj("model.obj")    # AnimationManager::read(file)
    engine.load_scene(sc)                   # key frames, normals, materials, lights, camera
    k1, k2, c = sc.frame(t)                 # getFrameId / getNextFrameId / coefficient
    t = sc.next_frame(t)                    # nextFrame

Parsing runs in the native library (happyray_amd/csrc/hr_scene.cpp) and
needs no GPU.  References: src/Application/SceneLoader.cpp:212-332,
src/Application/AnimationManager.hpp:46-138.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


class SceneFile:
    def __init__(self, handle):
        self._L = _lib.load()
        self._h = handle
        info = _lib.SceneInfo()
        check(self._L.hr_scene_get_info(self._h, C.byref(info)), "hr_scene_get_info")
        self.info = info

    @classmethod
    def load(cls, cfg_path):
        L = _lib.load()
        h = C.c_void_p()
        check(L.hr_scene_load(str(cfg_path).encode(), C.byref(h)), "hr_scene_load")
        return cls(h)

    @classmethod
    def load_obj(cls, obj_path):
        L = _lib.load()
        h = C.c_void_p()
        check(L.hr_scene_load_obj(str(obj_path).encode(), C.byref(h)), "hr_scene_load_obj")
        return cls(h)

    def close(self):
        if getattr(self, "_h", None):
            self._L.hr_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def num_frames(self):
        return self.info.num_frames

    @property
    def camera(self):
        """dict for Engine.set_camera_dict, or None without a camera file."""
        i = self.info
        if not i.has_camera:
            return None
        return {"pos": list(i.cam_pos), "dir": list(i.cam_dir), "up": list(i.cam_up),
                "fov": i.cam_fov, "res": (i.cam_res_x, i.cam_res_y)}

    def vertices(self, frame=0):
        out = np.empty((3, self.info.num_vertices), np.float32)
        check(self._L.hr_scene_get_vertices(self._h, frame, out.ctypes.data_as(_f32p)),
              "hr_scene_get_vertices")
        return out

    def keyframes(self):
        return np.stack([self.vertices(f) for f in range(self.num_frames)])

    def normals(self, frame=0):
        out = np.empty((3, self.info.num_normals), np.float32)
        check(self._L.hr_scene_get_normals(self._h, frame, out.ctypes.data_as(_f32p)),
              "hr_scene_get_normals")
        return out

    def triangles(self):
        """(idx, nidx, mat_idx): (nt, 3), (nt, 3), (nt,) uint32."""
        nt = self.info.num_triangles
        idx = np.empty((nt, 3), np.uint32)
        nidx = np.empty((nt, 3), np.uint32)
        mi = np.empty(nt, np.uint32)
        check(self._L.hr_scene_get_triangles(self._h, idx.ctypes.data_as(_u32p),
                                             nidx.ctypes.data_as(_u32p), mi.ctypes.data_as(_u32p)),
              "hr_scene_get_triangles")
        return idx, nidx, mi

    def materials(self):
        out = np.empty((self.info.num_materials, 12), np.float32)
        check(self._L.hr_scene_get_materials(self._h, out.ctypes.data_as(_f32p)),
              "hr_scene_get_materials")
        return out

    def lights(self):
        """(lights (n, 15) float32, weights (n,) float32)."""
        n = self.info.num_lights
        L = np.empty((max(n, 1), 15), np.float32)
        w = np.empty(max(n, 1), np.float32)
        check(self._L.hr_scene_get_lights(self._h, L.ctypes.data_as(_f32p), w.ctypes.data_as(_f32p)),
              "hr_scene_get_lights")
        return L[:n], w[:n]

    def instances(self):
        """(obj_first (num_objects + 1,), objects, matrices (n, 3, 4), aabb_min, aabb_max)
        as hr_set_instances / Engine.set_instances take them."""
        from . import _lib
        nobj = self.info.num_objects
        n = self.info.num_instances
        first = np.zeros(nobj + 1, np.uint32)
        arr = (_lib.Instance * max(n, 1))()
        check(self._L.hr_scene_get_instances(self._h, first.ctypes.data_as(_u32p), arr),
              "hr_scene_get_instances")
        objects = np.array([arr[i].object for i in range(n)], np.uint32)
        m = np.array([[list(arr[i].m[r]) for r in range(3)] for i in range(n)], np.float32).reshape(n, 3, 4)
        lo = np.array([list(arr[i].aabb_min) for i in range(n)], np.float32).reshape(n, 3)
        hi = np.array([list(arr[i].aabb_max) for i in range(n)], np.float32).reshape(n, 3)
        return first, objects, m, lo, hi

    def frame(self, current):
        k1, k2, c = C.c_uint32(), C.c_uint32(), C.c_float()
        check(self._L.hr_scene_frame(self._h, float(current), C.byref(k1), C.byref(k2), C.byref(c)),
              "hr_scene_frame")
        return k1.value, k2.value, c.value

    def next_frame(self, current):
        t = C.c_float(current)
        check(self._L.hr_scene_next_frame(self._h, C.byref(t)), "hr_scene_next_frame")
        return t.value
