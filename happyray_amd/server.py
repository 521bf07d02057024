"""The request server over the C ABI (hr_server_*).

    srv = Server(memory_budget=8 << 30)
    sid = srv.add_scene(keys, idx, materials, mat_idx, lights, weights)
    p = srv.submit_closest(sid, org, dirs)      # returns at once
    t, prim = p.wait()                          # blocks until the request is done
    srv.close()                                 # finishes what was submitted
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

REQ_CLOSEST, REQ_OCCLUDED, REQ_PATH, REQ_REBUILD = 0, 1, 2, 3

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


class Pending:
    """A submitted request; keeps its host buffers alive until wait()."""

    def __init__(self, server, rid, kind, outputs, keep):
        self.server = server
        self.id = rid
        self.kind = kind
        self.outputs = outputs
        self._keep = keep
        self.result = None

    def wait(self):
        res = _lib.RequestResult()
        check(self.server._L.hr_server_wait(self.server._h, self.id, C.byref(res)), "hr_server_wait")
        self.result = res
        self._keep = None
        return self.outputs

    @property
    def latency_ms(self):
        return None if self.result is None else self.result.latency_ms


class Server:
    def __init__(self, device=0, memory_budget=0, max_wave_rays=0, seed=(0x1234, 0x5678)):
        self._L = _lib.load()
        h = C.c_void_p()
        o = _lib.ServerOptions(device, int(memory_budget), int(max_wave_rays), seed[0] & 0xFFFFFFFF,
                               seed[1] & 0xFFFFFFFF)
        check(self._L.hr_server_create(C.byref(h), C.byref(o)), "hr_server_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.hr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, keys, idx, materials=None, mat_idx=None, lights=None, weights=None, lambda_top=0.0,
                  lambda_leaf=0.0):
        """keys: (frames, 3, nv) or (3, nv) float32 SoA key frames; idx (nt, 3)."""
        k = np.ascontiguousarray(keys, np.float32)
        if k.ndim == 2:
            k = k[None]
        ix = np.ascontiguousarray(idx, np.uint32).reshape(-1)
        m = None if materials is None else np.ascontiguousarray(materials, np.float32).reshape(-1, 12)
        mi = None if mat_idx is None else np.ascontiguousarray(mat_idx, np.uint32)
        L = None if lights is None else np.ascontiguousarray(lights, np.float32).reshape(-1, 15)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        d = _lib.ServerScene(k.ctypes.data_as(_f32p), k.shape[0], k.shape[2], ix.ctypes.data_as(_u32p),
                             len(ix) // 3, None if m is None else m.ctypes.data_as(_f32p),
                             0 if m is None else len(m), None if mi is None else mi.ctypes.data_as(_u32p),
                             None if L is None else L.ctypes.data_as(_f32p),
                             None if w is None else w.ctypes.data_as(_f32p), 0 if L is None else len(L),
                             float(lambda_top), float(lambda_leaf))
        sid = C.c_uint32()
        check(self._L.hr_server_add_scene(self._h, C.byref(d), C.byref(sid)), "hr_server_add_scene")
        return sid.value

    def _submit(self, req, outputs, keep):
        rid = C.c_uint64()
        check(self._L.hr_server_submit(self._h, C.byref(req), C.byref(rid)), "hr_server_submit")
        return Pending(self, rid.value, req.kind, outputs, keep + [req])

    @staticmethod
    def _rays(org, dirs):
        o = np.ascontiguousarray(org, np.float32).reshape(3, -1)
        d = np.ascontiguousarray(dirs, np.float32).reshape(3, -1)
        return o, d, o.shape[1]

    def submit_closest(self, scene, org, dirs):
        o, d, n = self._rays(org, dirs)
        t = np.zeros(n, np.float32)
        p = np.zeros(n, np.uint32)
        r = _lib.Request(REQ_CLOSEST, scene, n, 0.0, 0, 0, o.ctypes.data, d.ctypes.data, 0, t.ctypes.data,
                         p.ctypes.data, None, None, 0, 0, 0.0)
        return self._submit(r, (t, p), [o, d])

    def submit_occluded(self, scene, org, dirs, tmax):
        o, d, n = self._rays(org, dirs)
        occ = np.zeros(n, np.uint8)
        r = _lib.Request(REQ_OCCLUDED, scene, n, float(tmax), 0, 0, o.ctypes.data, d.ctypes.data, 0, None,
                         None, occ.ctypes.data, None, 0, 0, 0.0)
        return self._submit(r, occ, [o, d])

    def submit_path(self, scene, org, dirs, depth=4):
        o, d, n = self._rays(org, dirs)
        rgb = np.zeros((n, 3), np.float32)
        r = _lib.Request(REQ_PATH, scene, n, 0.0, int(depth), 0, o.ctypes.data, d.ctypes.data, 0, None, None,
                         None, rgb.ctypes.data, 0, 0, 0.0)
        return self._submit(r, rgb, [o, d])

    def submit_rebuild(self, scene, k1, k2, coeff):
        r = _lib.Request(REQ_REBUILD, scene, 0, 0.0, 0, 0, None, None, 0, None, None, None, None, int(k1),
                         int(k2), float(coeff))
        return self._submit(r, None, [])

    def submit_records(self, kind, scene, records, tmax=0.0, depth=4):
        """A request whose payload is 24-byte ray records (origin xyz, direction xyz)."""
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 6)
        n = len(rec)
        if kind == REQ_CLOSEST:
            outs = (np.zeros(n, np.float32), np.zeros(n, np.uint32))
            ptrs = (outs[0].ctypes.data, outs[1].ctypes.data, None, None)
        elif kind == REQ_OCCLUDED:
            outs = np.zeros(n, np.uint8)
            ptrs = (None, None, outs.ctypes.data, None)
        elif kind == REQ_PATH:
            outs = np.zeros((n, 3), np.float32)
            ptrs = (None, None, None, outs.ctypes.data)
        else:
            raise ValueError("ray records carry closest, occluded or path requests")
        r = _lib.Request(kind, scene, n, float(tmax), int(depth), 0, rec.ctypes.data, None, 0, *ptrs, 0, 0, 0.0,
                         _lib.RAYS_RECORDS)
        return self._submit(r, outs, [rec])

    def submit_device(self, kind, scene, n, org_ptr, dir_ptr, stride=0, tmax=0.0, depth=4):
        """Rays already in device memory (SoA planes `stride` floats apart);
        outputs are computed and dropped."""
        r = _lib.Request(kind, scene, int(n), float(tmax), int(depth), 1, org_ptr, dir_ptr, int(stride), None,
                         None, None, None, 0, 0, 0.0)
        return self._submit(r, None, [])

    def stats(self):
        s = _lib.ServerStats()
        check(self._L.hr_server_get_stats(self._h, C.byref(s)), "hr_server_get_stats")
        return s
