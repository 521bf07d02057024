"""Batches of scenes over the C ABI (hr_batch_*): many resident scenes, one
segmented two-level build, persistent traces that mix rays of all scenes, and
the autotuner.

    b = Batch()
    b.set_scenes([(verts_soa, idx), ...])      # verts (3, nv) float32, idx (nt, 3)
    b.build()
    rays = b.make_rays([(scene, org, dir, anyhit, tmax), ...])   # resident on the device
    b.trace(rays); b.synchronize(); results = rays.results()
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .devmem import DeviceArray

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


class RaySet:
    """Ray batches resident in device memory, with their output buffers and
    the hr_ray_batch table that names them."""

    def __init__(self, items):
        self.items = []
        self.bufs = []
        table = (_lib.RayBatch * max(len(items), 1))()
        for i, (scene, org, d, anyhit, tmax) in enumerate(items):
            org = np.ascontiguousarray(org, np.float32).reshape(3, -1)
            d = np.ascontiguousarray(d, np.float32).reshape(3, -1)
            n = org.shape[1]
            o = DeviceArray.from_host(org)
            dd = DeviceArray.from_host(d)
            t = DeviceArray((max(n, 1),), np.float32)
            p = None if anyhit else DeviceArray((max(n, 1),), np.uint32)
            occ = DeviceArray((max(n, 1),), np.uint8) if anyhit else None
            self.bufs.append((o, dd, t, p, occ))
            self.items.append((int(scene), n, bool(anyhit), float(tmax)))
            table[i] = _lib.RayBatch(int(scene), n, int(bool(anyhit)), float(tmax), o.value, dd.value, t.value,
                                     None if p is None else p.value, None if occ is None else occ.value)
        self.table = table
        self.count = len(items)
        self.rays = sum(it[1] for it in self.items)

    def results(self):
        """Per batch: (t, prim) for closest hit, (occ, t) for any hit."""
        out = []
        for (scene, n, anyhit, tmax), (o, d, t, p, occ) in zip(self.items, self.bufs):
            if anyhit:
                out.append((occ.to_host()[:n], t.to_host()[:n]))
            else:
                out.append((t.to_host()[:n], p.to_host()[:n]))
        return out

    def free(self):
        for bufs in self.bufs:
            for b in bufs:
                if b is not None:
                    b.free()
        self.bufs = []


class Batch:
    def __init__(self, device=0):
        self._L = _lib.load()
        h = C.c_void_p()
        check(self._L.hr_batch_create(C.byref(h), C.byref(_lib.Options(device, 0, 0))), "hr_batch_create")
        self._h = h
        self.ns = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.hr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scenes(self, scenes):
        """scenes: list of (verts (3, nv) float32 SoA, idx (nt, 3) uint32)."""
        keep = []
        descs = (_lib.MeshDesc * len(scenes))()
        for i, (v, idx) in enumerate(scenes):
            v = np.ascontiguousarray(v, np.float32)
            ix = np.ascontiguousarray(idx, np.uint32).reshape(-1)
            keep.append((v, ix))
            descs[i] = _lib.MeshDesc(v.ctypes.data_as(_f32p), v.shape[1], ix.ctypes.data_as(_u32p), len(ix) // 3)
        check(self._L.hr_batch_set_scenes(self._h, descs, len(scenes)), "hr_batch_set_scenes")
        self.ns = len(scenes)
        self.nts = [len(k[1]) // 3 for k in keep]

    def set_densities(self, lambda_top, lambda_leaf):
        lt = np.ascontiguousarray(np.broadcast_to(np.float32(lambda_top), (self.ns,)), np.float32)
        ll = np.ascontiguousarray(np.broadcast_to(np.float32(lambda_leaf), (self.ns,)), np.float32)
        check(self._L.hr_batch_set_densities(self._h, lt.ctypes.data_as(_f32p), ll.ctypes.data_as(_f32p)),
              "hr_batch_set_densities")

    def densities(self):
        lt = np.zeros(self.ns, np.float32)
        ll = np.zeros(self.ns, np.float32)
        t = _lib.BatchTuning()
        check(self._L.hr_batch_get_tuning(self._h, C.byref(t), lt.ctypes.data_as(_f32p), ll.ctypes.data_as(_f32p)),
              "hr_batch_get_tuning")
        return lt, ll, (t.refill, t.blocks_per_cu)

    def set_tuning(self, refill=48, blocks_per_cu=0):
        check(self._L.hr_batch_set_tuning(self._h, C.byref(_lib.BatchTuning(refill, blocks_per_cu))),
              "hr_batch_set_tuning")

    def build(self):
        check(self._L.hr_batch_build(self._h), "hr_batch_build")

    def grid_info(self, s):
        gi = _lib.GridInfo()
        check(self._L.hr_batch_get_grid_info(self._h, s, C.byref(gi)), "hr_batch_get_grid_info")
        return gi

    def grid(self, s):
        """Scene-local (info, cells (C, 2), leaves (L, 2), prims) as hr_get_grid lays them out."""
        gi = self.grid_info(s)
        cells = np.zeros((gi.num_cells, 2), np.uint32)
        leaves = np.zeros((gi.num_leaves, 2), np.uint32)
        prims = np.zeros(gi.num_refs, np.uint32)
        check(self._L.hr_batch_get_grid(self._h, s, cells.ctypes.data_as(_u32p),
                                        leaves.ctypes.data_as(_u32p) if gi.num_leaves else None,
                                        prims.ctypes.data_as(_u32p) if gi.num_refs else None),
              "hr_batch_get_grid")
        return gi, cells, leaves, prims

    @staticmethod
    def make_rays(items):
        return RaySet(items)

    def trace(self, rays):
        check(self._L.hr_batch_trace_dev(self._h, rays.table, rays.count), "hr_batch_trace_dev")

    def autotune(self, rays):
        r = _lib.AutotuneResult()
        check(self._L.hr_batch_autotune(self._h, rays.table, rays.count, C.byref(r)), "hr_batch_autotune")
        return r

    def timing(self):
        b, t, n = C.c_double(), C.c_double(), C.c_uint64()
        check(self._L.hr_batch_get_timing(self._h, C.byref(b), C.byref(t), C.byref(n)), "hr_batch_get_timing")
        return b.value, t.value, n.value

    def work_stats(self):
        ws = _lib.WorkStats()
        check(self._L.hr_batch_get_work_stats(self._h, C.byref(ws)), "hr_batch_get_work_stats")
        return ws

    def synchronize(self):
        check(self._L.hr_batch_synchronize(self._h), "hr_batch_synchronize")
