"""Python mirror of the reference's RTEngine facade over the C ABI.

RTEngine (src/RT/RTEngine.h:36-68) is a static, single-instance class; here
each Engine is one hr_ctx.  Method names follow the reference:

    init / cleanup              -> Engine() / close()
    upload(frame1, frame2, c)   -> upload(verts1, verts2, coeff, idx)
    setGridDensities(top, leaf) -> set_grid_densities(top, leaf)
    buildAccStruct()            -> build_acc_struct()
    setCamera(...)              -> set_camera(...)
    setLights(...)              -> set_lights(...)
    renderFrame(fb, id, mode)   -> render_frame(image_id, mode, ...)
    getBoundingBoxDiagonalLength() -> get_bounding_box_diagonal_length()
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (GRID_HIERARCHY, GRID_TWO_LEVEL, GRID_UNIFORM, GRID_UNIFORM_EXACT, MODE_AO, MODE_PATH, MODE_PATH_MULTI,
                   MODE_SIMPLE, check)

# timed frames that persistent_primary 2 spends choosing the camera-ray kernel
# after the mesh, key frames, grid, camera or tuning change (kernels 0, 1, 3,
# twice each); a caller that times frames renders these first
CAMERA_KERNEL_TRIALS = 6

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)


def _f(a):
    return a.ctypes.data_as(_f32p)


def _u(a):
    return a.ctypes.data_as(_u32p)


def _vec3(v):
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))


class Engine:
    def __init__(self, device=0, seed=(0x1234, 0x5678)):
        self._L = _lib.load()
        h = C.c_void_p()
        opts = _lib.Options(device, seed[0] & 0xFFFFFFFF, seed[1] & 0xFFFFFFFF)
        check(self._L.hr_create(C.byref(h), C.byref(opts)), "hr_create")
        self._h = h
        self._res = None

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.hr_destroy(self._h)
            self._h = None

    cleanup = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def stream(self):
        s = C.c_void_p()
        check(self._L.hr_get_stream(self._h, C.byref(s)), "hr_get_stream")
        return s.value

    # -- geometry ---------------------------------------------------------
    def upload(self, verts1, verts2, coeff, idx):
        """verts: (3, nv) float32 SoA key frames; idx: (nt, 3) uint32."""
        v1 = np.ascontiguousarray(verts1, np.float32)
        v2 = v1 if verts2 is None else np.ascontiguousarray(verts2, np.float32)
        ix = np.ascontiguousarray(idx, np.uint32).reshape(-1)
        self._nt = len(ix) // 3
        check(self._L.hr_upload_mesh(self._h, _f(v1), _f(v2), float(coeff), v1.shape[1], _u(ix),
                                     self._nt), "hr_upload_mesh")

    def upload_scene(self, scene):
        self.upload(scene.verts, None, 1.0, scene.idx)
        self.upload_materials(scene.materials, scene.mat_idx)

    def upload_materials(self, mats, mat_idx=None):
        m = np.ascontiguousarray(mats, np.float32).reshape(-1, 12)
        mi = None if mat_idx is None else np.ascontiguousarray(mat_idx, np.uint32)
        check(self._L.hr_upload_materials(self._h, _f(m), len(m), None if mi is None else _u(mi),
                                          0 if mi is None else len(mi)), "hr_upload_materials")

    def set_keyframes(self, keys, idx):
        """keys: (frames, 3, nv) float32."""
        k = np.ascontiguousarray(keys, np.float32)
        ix = np.ascontiguousarray(idx, np.uint32).reshape(-1)
        self._nt = len(ix) // 3
        check(self._L.hr_set_keyframes(self._h, _f(k), k.shape[0], k.shape[2], _u(ix), self._nt),
              "hr_set_keyframes")

    def set_frame(self, k1, k2, coeff):
        check(self._L.hr_set_frame(self._h, k1, k2, float(coeff)), "hr_set_frame")

    def set_normals(self, normals, nidx):
        """normals: (3, nn) or (frames, 3, nn) float32 SoA; nidx: (nt, 3) normal indices."""
        n = np.ascontiguousarray(normals, np.float32)
        if n.ndim == 2:
            n = n[None]
        ni = np.ascontiguousarray(nidx, np.uint32).reshape(-1)
        check(self._L.hr_set_normals(self._h, _f(n), n.shape[0], n.shape[2], _u(ni), len(ni) // 3),
              "hr_set_normals")

    def load_scene(self, scene_file):
        """Make a parsed SceneFile this engine's geometry (hr_scene_upload)."""
        check(self._L.hr_scene_upload(self._h, scene_file.handle), "hr_scene_upload")
        self._nt = scene_file.info.num_triangles
        if scene_file.info.has_camera:
            i = scene_file.info
            self._res = (i.cam_res_x, i.cam_res_y)

    # -- build ------------------------------------------------------------
    def set_grid(self, kind, lambda_top, lambda_leaf=2.2):
        check(self._L.hr_set_grid(self._h, kind, float(lambda_top), float(lambda_leaf)),
              "hr_set_grid")
        self._kind = kind

    def set_grid_densities(self, top, leaf):
        self.set_grid(getattr(self, "_kind", GRID_TWO_LEVEL), top, leaf)

    def build_acc_struct(self):
        check(self._L.hr_build(self._h), "hr_build")

    build = build_acc_struct

    def grid_info(self):
        gi = _lib.GridInfo()
        check(self._L.hr_get_grid_info(self._h, C.byref(gi)), "hr_get_grid_info")
        return gi

    def grid(self):
        """Returns (info, cells (C, 2), leaves (L, 2), prims (P,)) copied to the host."""
        gi = self.grid_info()
        cells = np.zeros((gi.num_cells, 2), np.uint32)
        leaves = np.zeros((gi.num_leaves, 2), np.uint32)
        prims = np.zeros(gi.num_refs, np.uint32)
        check(self._L.hr_get_grid(self._h, _u(cells), _u(leaves) if gi.num_leaves else None,
                                  _u(prims) if gi.num_refs else None), "hr_get_grid")
        return gi, cells, leaves, prims

    # -- instanced hierarchy ----------------------------------------------
    def set_instances(self, obj_first, objects, matrices, aabb_min, aabb_max):
        """Objects are triangle ranges [obj_first[k], obj_first[k+1]); instance i
        shows object objects[i] through the world-to-object rows matrices[i]
        (3x4, the instances file's m_row_0..2) inside the world box
        [aabb_min[i], aabb_max[i]] (WFObject::loadInstances,
        src/Application/WFObjectReader.cpp:762-824)."""
        first = np.ascontiguousarray(obj_first, np.uint32)
        objects = np.asarray(objects, np.uint32).reshape(-1)
        m = np.asarray(matrices, np.float32).reshape(-1, 3, 4)
        lo = np.asarray(aabb_min, np.float32).reshape(-1, 3)
        hi = np.asarray(aabb_max, np.float32).reshape(-1, 3)
        n = objects.shape[0]
        arr = (_lib.Instance * max(n, 1))()
        for i in range(n):
            arr[i].object = int(objects[i])
            for r in range(3):
                for c in range(4):
                    arr[i].m[r][c] = float(m[i, r, c])
                arr[i].aabb_min[r] = float(lo[i, r])
                arr[i].aabb_max[r] = float(hi[i, r])
        check(self._L.hr_set_instances(self._h, _u(first), first.shape[0] - 1, arr, n), "hr_set_instances")

    def object_grids(self):
        """(list of dicts per object grid, instance references of the top cells)."""
        gi = self.grid_info()
        grids = (_lib.ObjectGrid * max(gi.num_objects, 1))()
        refs = np.zeros(gi.num_top_pairs, np.uint32)
        check(self._L.hr_get_object_grids(self._h, grids, _u(refs) if gi.num_top_pairs else None),
              "hr_get_object_grids")
        out = []
        for g in grids[:gi.num_objects]:
            out.append(dict(bmin=np.array(g.bmin[:], np.float32), bmax=np.array(g.bmax[:], np.float32),
                            res=list(g.res), cell_size=np.array(g.cell_size[:], np.float32),
                            cell_size_rcp=np.array(g.cell_size_rcp[:], np.float32),
                            first_cell=g.first_cell, num_cells=g.num_cells))
        return out, refs

    def trace_instanced(self, org, dirs):
        """Closest hits on the instanced hierarchy: (t, prim, instance)."""
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[1]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.uint32)
        inst = np.zeros(n, np.uint32)
        check(self._L.hr_trace_instanced(self._h, _f(o), _f(d), n, _f(t), _u(prim), _u(inst)),
              "hr_trace_instanced")
        return t, prim, inst

    def primary_instances(self):
        n = self._res[0] * self._res[1]
        inst = np.zeros(n, np.uint32)
        check(self._L.hr_get_primary_instances(self._h, _u(inst)), "hr_get_primary_instances")
        return inst

    def get_bounding_box_diagonal_length(self):
        gi = self.grid_info()
        d = np.array(gi.bmax[:], np.float32) - np.array(gi.bmin[:], np.float32)
        return float(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))

    # -- camera / lights --------------------------------------------------
    def set_camera(self, pos, dir, up, fov, res_x, res_y):
        check(self._L.hr_set_camera(self._h, _f(_vec3(pos)), _f(_vec3(dir)), _f(_vec3(up)),
                                    float(fov), int(res_x), int(res_y)), "hr_set_camera")
        self._res = (int(res_x), int(res_y))

    def set_camera_dict(self, cam):
        self.set_camera(cam["pos"], cam["dir"], cam["up"], cam["fov"], *cam["res"])

    def set_lights(self, lights, weights=None):
        L = np.ascontiguousarray(lights, np.float32).reshape(-1, 15)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        check(self._L.hr_set_lights(self._h, _f(L), None if w is None else _f(w), len(L)),
              "hr_set_lights")

    # -- render -----------------------------------------------------------
    def render_frame(self, image_id=0, mode=MODE_AO, jitter=1, ao_samples=4, ao_dist=2.0,
                     pt_depth=4, download=True, collect_stats=False):
        p = _lib.RenderParams(int(jitter), int(ao_samples), float(ao_dist), int(pt_depth),
                              int(bool(collect_stats)))
        out = None
        if download:
            out = np.zeros((self._res[0] * self._res[1], 3), np.float32)
        check(self._L.hr_render(self._h, mode, int(image_id), C.byref(p),
                                None if out is None else _f(out)), "hr_render")
        return out

    def render_frame_graph(self, k1, k2, coeff, image_id=0, mode=MODE_PATH, jitter=1, ao_samples=4,
                           ao_dist=2.0, pt_depth=4, download=True, collect_stats=False):
        """One captured frame of the resident key frames: interpolation, rebuild,
        primary rays and the integrator replayed from recorded HIP graphs."""
        p = _lib.RenderParams(int(jitter), int(ao_samples), float(ao_dist), int(pt_depth),
                              int(bool(collect_stats)))
        out = None
        if download:
            out = np.zeros((self._res[0] * self._res[1], 3), np.float32)
        check(self._L.hr_render_frame_graph(self._h, int(k1), int(k2), float(coeff), mode, int(image_id),
                                            C.byref(p), None if out is None else _f(out)),
              "hr_render_frame_graph")
        return out

    def submit_frame_graph(self, k1, k2, coeff, image_id=0, mode=MODE_PATH, jitter=1, ao_samples=4,
                           ao_dist=2.0, pt_depth=4, collect_stats=False):
        """The same frame as render_frame_graph, queued without waiting for it:
        returns once the frame submitted two calls earlier has finished.  Any
        other call on the engine (framebuffer(), synchronize(), ...) first
        completes the frames in flight."""
        p = _lib.RenderParams(int(jitter), int(ao_samples), float(ao_dist), int(pt_depth),
                              int(bool(collect_stats)))
        check(self._L.hr_submit_frame_graph(self._h, int(k1), int(k2), float(coeff), mode, int(image_id),
                                            C.byref(p)), "hr_submit_frame_graph")

    def framebuffer(self):
        """The framebuffer after the last render, (res_x*res_y, 3) float32."""
        out = np.zeros((self._res[0] * self._res[1], 3), np.float32)
        check(self._L.hr_get_framebuffer(self._h, _f(out)), "hr_get_framebuffer")
        return out

    def primary_hits(self):
        n = self._res[0] * self._res[1]
        planes = np.zeros((7, n), np.float32)
        prims = np.zeros(n, np.uint32)
        check(self._L.hr_get_primary_hits(self._h, _f(planes), _u(prims)), "hr_get_primary_hits")
        return planes, prims

    def ao_occlusion(self, ao_samples):
        n = self._res[0] * self._res[1]
        occ = np.zeros(n * ao_samples, np.uint8)
        check(self._L.hr_get_ao_occlusion(self._h, occ.ctypes.data_as(_u8p)), "hr_get_ao_occlusion")
        return occ

    # -- ray batches ------------------------------------------------------
    def trace(self, org, dirs):
        """org, dirs: (3, n) float32. Closest hit: (t, prim)."""
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[1]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.uint32)
        check(self._L.hr_trace(self._h, _f(o), _f(d), n, _f(t), _u(prim)), "hr_trace")
        return t, prim

    def occluded(self, org, dirs, tmax):
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[1]
        occ = np.zeros(n, np.uint8)
        check(self._L.hr_occluded(self._h, _f(o), _f(d), n, float(tmax), occ.ctypes.data_as(_u8p)),
              "hr_occluded")
        return occ

    # -- ray records (24 bytes per ray: origin xyz, direction xyz) --------
    def trace_records(self, records):
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 6)
        n = len(r)
        t = np.zeros(n, np.float32)
        p = np.zeros(n, np.uint32)
        check(self._L.hr_trace_wire(self._h, _f(r), n, _f(t), p.ctypes.data_as(_u32p)), "hr_trace_wire")
        return t, p

    def occluded_records(self, records, tmax):
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 6)
        occ = np.zeros(len(r), np.uint8)
        check(self._L.hr_occluded_wire(self._h, _f(r), len(r), float(tmax), occ.ctypes.data_as(_lib._u8p)),
              "hr_occluded_wire")
        return occ

    def render_rays(self, records):
        """renderFrame mode 0 with input rays: (n, 3) simple shading of each ray's hit."""
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 6)
        rgb = np.zeros((len(r), 3), np.float32)
        check(self._L.hr_render_rays(self._h, _f(r), len(r), _f(rgb)), "hr_render_rays")
        return rgb

    def dump_ao_rays(self, ao_samples=None):
        """The last AO render's rays as AORayExporter dumps them: (npix * (S + 1), 6).
        ao_samples, when given, is checked against the dump's size."""
        cnt = C.c_uint64()
        check(self._L.hr_dump_ao_rays(self._h, None, 0, C.byref(cnt)), "hr_dump_ao_rays")
        n = cnt.value
        if ao_samples is not None and n != self._res[0] * self._res[1] * (int(ao_samples) + 1):
            raise ValueError("the last AO render used another sample count or resolution")
        out = np.zeros((n, 6), np.float32)
        check(self._L.hr_dump_ao_rays(self._h, _f(out), n, None), "hr_dump_ao_rays")
        return out

    def path_trace(self, org, dirs, depth=4):
        """org, dirs: (3, n) float32. Radiance (n, 3) of a path from each ray."""
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[1]
        rgb = np.zeros((n, 3), np.float32)
        check(self._L.hr_path_trace(self._h, _f(o), _f(d), n, int(depth), _f(rgb)), "hr_path_trace")
        return rgb

    def path_trace_dev(self, org, dirs, n, depth, rgb_out):
        check(self._L.hr_path_trace_dev(self._h, org, dirs, n, int(depth), rgb_out), "hr_path_trace_dev")

    def path_trace_dev_noalloc(self, org, dirs, n, depth, rgb_out):
        """path_trace_dev without any device allocation (at most 2,097,152 rays;
        the scratch is shared by every context on the device)."""
        check(self._L.hr_path_trace_dev_noalloc(self._h, org, dirs, n, int(depth), rgb_out),
              "hr_path_trace_dev_noalloc")

    def trace_dev(self, org, dirs, n, t_out, prim_out):
        """Closest hit of n rays whose SoA buffers are device pointers (ints)."""
        check(self._L.hr_trace_dev(self._h, org, dirs, n, t_out, prim_out), "hr_trace_dev")

    def occluded_dev(self, org, dirs, n, tmax, occ_out):
        check(self._L.hr_occluded_dev(self._h, org, dirs, n, float(tmax), occ_out), "hr_occluded_dev")

    def timing(self):
        t = _lib.Timing()
        check(self._L.hr_get_timing(self._h, C.byref(t)), "hr_get_timing")
        return t

    def set_tuning(self, persistent_primary=None, persistent_secondary=None, refill=None, hier_refill=None,
                   hier_bounds_blocks=None, persistent_path=None, path_refill=None, path_shade=None,
                   fuse_units=None, test_batch=None, advance_steps=None, frame_lanes=None,
                   camera_blocks=None):
        """Kernel choices (hr_tuning).  Every field keeps its current value
        when not given."""
        cur = self.tuning()

        def pick(v, name):
            return int(getattr(cur, name) if v is None else v)

        t = _lib.Tuning(pick(persistent_primary, "persistent_primary"),
                        pick(persistent_secondary, "persistent_secondary"), pick(refill, "refill"),
                        pick(hier_refill, "hier_refill"), pick(hier_bounds_blocks, "hier_bounds_blocks"),
                        pick(persistent_path, "persistent_path"), pick(path_refill, "path_refill"),
                        pick(path_shade, "path_shade"), pick(fuse_units, "fuse_units"),
                        pick(test_batch, "test_batch"), pick(advance_steps, "advance_steps"),
                        pick(frame_lanes, "frame_lanes"), pick(camera_blocks, "camera_blocks"))
        check(self._L.hr_set_tuning(self._h, C.byref(t)), "hr_set_tuning")

    def tuning(self):
        t = _lib.Tuning()
        check(self._L.hr_get_tuning(self._h, C.byref(t)), "hr_get_tuning")
        return t

    def work_stats(self):
        w = _lib.WorkStats()
        check(self._L.hr_get_work_stats(self._h, C.byref(w)), "hr_get_work_stats")
        return w

    def synchronize(self):
        check(self._L.hr_synchronize(self._h), "hr_synchronize")


__all__ = ["Engine", "GRID_HIERARCHY", "GRID_UNIFORM", "GRID_UNIFORM_EXACT", "GRID_TWO_LEVEL", "MODE_AO",
           "MODE_PATH", "MODE_PATH_MULTI", "MODE_SIMPLE"]
