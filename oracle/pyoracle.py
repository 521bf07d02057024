"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

GRID_UNIFORM = 0
GRID_UNIFORM_EXACT = 1
GRID_TWO_LEVEL = 2

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)


class Mesh(C.Structure):
    _fields_ = [("vx", _f32p), ("vy", _f32p), ("vz", _f32p), ("nv", C.c_uint32),
                ("idx", _u32p), ("nt", C.c_uint32), ("nx", _f32p), ("ny", _f32p),
                ("nz", _f32p), ("nn", C.c_uint32), ("nidx", _u32p)]


class Grid(C.Structure):
    _fields_ = [("kind", C.c_int), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
                ("res", C.c_int * 3), ("cell_size", C.c_float * 3),
                ("cell_size_rcp", C.c_float * 3), ("num_cells", C.c_uint32),
                ("num_leaves", C.c_uint32), ("num_refs", C.c_uint32),
                ("num_top_pairs", C.c_uint32), ("num_bbox_pairs", C.c_uint32),
                ("cells", _u32p), ("leaves", _u32p), ("prims", _u32p)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("steps", C.c_uint64), ("tests", C.c_uint64)]


class Camera(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("top_left", C.c_float * 3),
                ("step_x", C.c_float * 3), ("step_y", C.c_float * 3),
                ("res_x", C.c_int), ("res_y", C.c_int)]


class RenderParams(C.Structure):
    _fields_ = [("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32), ("jitter", C.c_int),
                ("ao_samples", C.c_int), ("ao_dist", C.c_float), ("pt_depth", C.c_int),
                ("nthreads", C.c_int), ("simple_sampler", C.c_int)]


class Instance(C.Structure):
    _fields_ = [("object", C.c_uint32), ("m", (C.c_float * 4) * 3), ("aabb_min", C.c_float * 3),
                ("aabb_max", C.c_float * 3)]


class ObjectGrid(C.Structure):
    _fields_ = [("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("res", C.c_int * 3),
                ("cell_size", C.c_float * 3), ("cell_size_rcp", C.c_float * 3), ("first_cell", C.c_uint32),
                ("num_cells", C.c_uint32)]


class Hier(C.Structure):
    _fields_ = [("top", Grid), ("nobj", C.c_uint32), ("ninst", C.c_uint32), ("objs", C.POINTER(ObjectGrid)),
                ("num_obj_cells", C.c_uint32), ("obj_cells", _u32p), ("num_obj_refs", C.c_uint32),
                ("prims", _u32p), ("num_inst_refs", C.c_uint32), ("inst_refs", _u32p),
                ("inst", C.POINTER(Instance))]


def build_lib():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build_lib()
        L = C.CDLL(_LIB_PATH)
        L.or_powf.restype = C.c_float
        L.or_powf.argtypes = [C.c_float, C.c_float]
        L.or_logf.restype = C.c_float
        L.or_logf.argtypes = [C.c_float]
        L.or_expf.restype = C.c_float
        L.or_expf.argtypes = [C.c_float]
        L.or_sincosf.argtypes = [C.c_float, _f32p, _f32p]
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_u01.restype = C.c_float
        L.or_u01.argtypes = [C.c_uint32]
        L.or_lerp_bounds.argtypes = [_f32p] * 6 + [C.c_float, C.c_uint32] + [_f32p] * 5
        L.or_frame_ids.argtypes = [C.c_float, C.c_float, C.c_uint32, _u32p, _u32p, _f32p]
        L.or_grid_resolution.argtypes = [_f32p, _f32p, C.c_uint32, C.c_float, C.POINTER(C.c_int)]
        L.or_build.argtypes = [C.POINTER(Mesh), _f32p, _f32p, C.c_int, C.c_float, C.c_float,
                               C.POINTER(Grid)]
        L.or_grid_free.argtypes = [C.POINTER(Grid)]
        L.or_prim_cell_range.argtypes = [C.POINTER(Mesh), C.c_uint32, C.POINTER(Grid),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_tri_box_overlap.argtypes = [_f32p] * 5
        L.or_shading_normal.argtypes = [C.POINTER(Mesh), C.c_uint32, _f32p, _f32p]
        L.or_primary_dirs.argtypes = [C.POINTER(Camera), C.POINTER(RenderParams), C.c_int, _f32p, _f32p, _f32p]
        L.or_path_trace_rays.argtypes = [C.POINTER(Mesh), C.POINTER(Grid)] + [_f32p] * 6 + [
            C.c_uint32, _f32p, _u32p, _f32p, _f32p, C.c_int, C.POINTER(RenderParams), _f32p, C.POINTER(Stats)]
        L.or_trace.argtypes = [C.POINTER(Mesh), C.POINTER(Grid)] + [_f32p] * 6 + [
            C.c_uint32, C.c_int, C.c_float, C.c_int, _f32p, _u32p, C.POINTER(Stats)]
        L.or_camera_init.argtypes = [C.POINTER(Camera), _f32p, _f32p, _f32p, C.c_float,
                                     C.c_int, C.c_int]
        L.or_primary.argtypes = [C.POINTER(Mesh), C.POINTER(Grid), C.POINTER(Camera),
                                 C.POINTER(RenderParams), C.c_int] + [_f32p] * 7 + [
            _u32p, C.POINTER(Stats)]
        L.or_render_ao.argtypes = [C.POINTER(Mesh), C.POINTER(Grid), C.POINTER(Camera),
                                   _f32p, _u32p, C.POINTER(RenderParams), C.c_int, _f32p,
                                   _u8p, C.POINTER(Stats)]
        L.or_render_pt.argtypes = [C.POINTER(Mesh), C.POINTER(Grid), C.POINTER(Camera),
                                   _f32p, _u32p, _f32p, _f32p, C.c_int,
                                   C.POINTER(RenderParams), C.c_int, _f32p, C.POINTER(Stats)]
        L.or_pt_glass.restype = C.c_int
        L.or_pt_glass.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float, _f32p,
                                  C.c_float, C.c_float, C.c_int]
        L.or_power_cosine_dir.argtypes = [C.c_float, C.c_float, C.c_float, _f32p, _f32p]
        L.or_cosine_dir.argtypes = [C.c_float, C.c_float, _f32p, _f32p]
        L.or_pt_glossy.restype = C.c_int
        L.or_pt_glossy.argtypes = [_f32p, _f32p, _f32p, C.c_float, _f32p, C.c_float, C.c_float, C.c_float]
        L.or_pm_merge.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_int]
        L.or_build_hier.argtypes = [C.POINTER(Mesh), _f32p, _f32p, _f32p, _f32p, _f32p, _u32p, C.c_uint32,
                                    C.POINTER(Instance), C.c_uint32, C.c_float, C.c_float, C.POINTER(Hier)]
        L.or_hier_free.argtypes = [C.POINTER(Hier)]
        L.or_trace_hier.argtypes = [C.POINTER(Mesh), C.POINTER(Hier)] + [_f32p] * 6 + [
            C.c_uint32, C.c_int, C.c_float, _f32p, _u32p, _u32p, _u8p]
        L.or_render_ao_hier.argtypes = [C.POINTER(Mesh), C.POINTER(Hier), C.POINTER(Camera), _f32p, _u32p,
                                        C.POINTER(RenderParams), C.c_int, _f32p, _u8p, _f32p, _u32p, _u32p]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _up(a):
    return a.ctypes.data_as(_u32p)


def _f3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))


class OMesh:
    """Keeps the numpy arrays alive behind an or_mesh."""

    def __init__(self, vx, vy, vz, idx, normals=None, nidx=None):
        """normals: (3, nn) SoA per-corner normals, nidx: (nt, 3) normal indices."""
        self.vx = np.ascontiguousarray(vx, dtype=np.float32)
        self.vy = np.ascontiguousarray(vy, dtype=np.float32)
        self.vz = np.ascontiguousarray(vz, dtype=np.float32)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        self.c = Mesh(_fp(self.vx), _fp(self.vy), _fp(self.vz), len(self.vx),
                      _up(self.idx), len(self.idx) // 3)
        if normals is not None:
            n = np.ascontiguousarray(normals, dtype=np.float32).reshape(3, -1)
            self.nx, self.ny, self.nz = (np.ascontiguousarray(n[k]) for k in range(3))
            self.nidx = np.ascontiguousarray(nidx, dtype=np.uint32).reshape(-1)
            assert len(self.nidx) == len(self.idx)
            assert len(self.nidx) == 0 or self.nidx.max() < n.shape[1]
            self.c.nx, self.c.ny, self.c.nz = _fp(self.nx), _fp(self.ny), _fp(self.nz)
            self.c.nn = n.shape[1]
            self.c.nidx = _up(self.nidx)

    @property
    def nt(self):
        return len(self.idx) // 3


class OGrid:
    """Owns an or_grid; exposes numpy copies of cells / leaves / prims."""

    def __init__(self, mesh, bmin, bmax, kind, lambda_top, lambda_leaf=1.0):
        self.mesh = mesh
        self.c = Grid()
        lib().or_build(C.byref(mesh.c), _fp(_f3(bmin)), _fp(_f3(bmax)), kind,
                       lambda_top, lambda_leaf, C.byref(self.c))

    def __del__(self):
        try:
            lib().or_grid_free(C.byref(self.c))
        except Exception:
            pass

    @property
    def res(self):
        return tuple(self.c.res)

    def params(self):
        g = self.c
        return {"bmin": np.array(g.bmin[:], np.float32), "bmax": np.array(g.bmax[:], np.float32),
                "res": tuple(g.res), "cell_size": np.array(g.cell_size[:], np.float32),
                "cell_size_rcp": np.array(g.cell_size_rcp[:], np.float32)}

    def cells(self):
        n = 2 * self.c.num_cells
        return np.ctypeslib.as_array(self.c.cells, shape=(n,)).copy().reshape(-1, 2)

    def leaves(self):
        n = 2 * self.c.num_leaves
        if n == 0:
            return np.zeros((0, 2), np.uint32)
        return np.ctypeslib.as_array(self.c.leaves, shape=(n,)).copy().reshape(-1, 2)

    def prims(self):
        n = self.c.num_refs
        if n == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(self.c.prims, shape=(n,)).copy()

    def prim_cell_range(self, prim):
        lo = (C.c_int * 3)()
        hi = (C.c_int * 3)()
        lib().or_prim_cell_range(C.byref(self.mesh.c), prim, C.byref(self.c), lo, hi)
        return tuple(lo), tuple(hi)


def powf(x, y):
    return lib().or_powf(x, y)


def logf(x):
    return lib().or_logf(x)


def expf(x):
    return lib().or_expf(x)


def sincosf(a):
    s = C.c_float()
    c = C.c_float()
    lib().or_sincosf(a, C.byref(s), C.byref(c))
    return s.value, c.value


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_up(c), _up(k), _up(o))
    return o


def u01(x):
    return lib().or_u01(int(x) & 0xFFFFFFFF)


def lerp_bounds(v1, v2, c):
    """v1, v2: (3, nv) float32 SoA key frames. Returns (verts (3,nv), bmin, bmax)."""
    v1 = np.ascontiguousarray(v1, np.float32)
    v2 = np.ascontiguousarray(v2, np.float32)
    nv = v1.shape[1]
    out = np.zeros((3, nv), np.float32)
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    lib().or_lerp_bounds(_fp(v1[0]), _fp(v1[1]), _fp(v1[2]), _fp(v2[0]), _fp(v2[1]),
                         _fp(v2[2]), c, nv, _fp(out[0]), _fp(out[1]), _fp(out[2]),
                         _fp(bmin), _fp(bmax))
    return out, bmin, bmax


def bounds(verts):
    _, bmin, bmax = lerp_bounds(verts, verts, 1.0)
    return bmin, bmax


def frame_ids(current, step, num_frames):
    k1 = C.c_uint32()
    k2 = C.c_uint32()
    c = C.c_float()
    lib().or_frame_ids(current, step, num_frames, C.byref(k1), C.byref(k2), C.byref(c))
    return k1.value, k2.value, c.value


def resolution(bmin, bmax, nprims, lam):
    res = (C.c_int * 3)()
    lib().or_grid_resolution(_fp(_f3(bmin)), _fp(_f3(bmax)), nprims, lam, res)
    return tuple(res)


def tri_box_overlap(v0, v1, v2, center, half):
    return bool(lib().or_tri_box_overlap(_fp(_f3(v0)), _fp(_f3(v1)), _fp(_f3(v2)),
                                         _fp(_f3(center)), _fp(_f3(half))))


def trace(grid, org, dirs, anyhit=False, tmax=3.402823466e38, nthreads=1):
    """org, dirs: (3, n) float32. Returns (t, prim, stats)."""
    org = np.ascontiguousarray(org, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = org.shape[1]
    t = np.zeros(n, np.float32)
    prim = np.zeros(n, np.uint32)
    st = Stats()
    lib().or_trace(C.byref(grid.mesh.c), C.byref(grid.c), _fp(org[0]), _fp(org[1]),
                   _fp(org[2]), _fp(dirs[0]), _fp(dirs[1]), _fp(dirs[2]), n, int(anyhit),
                   tmax, nthreads, _fp(t), _up(prim), C.byref(st))
    return t, prim, st


def camera(pos, dir, up, fov, res_x, res_y):
    cam = Camera()
    lib().or_camera_init(C.byref(cam), _fp(_f3(pos)), _fp(_f3(dir)), _fp(_f3(up)), fov,
                         res_x, res_y)
    return cam


def params(seed=(0x1234, 0x5678), jitter=1, ao_samples=4, ao_dist=2.0, pt_depth=4,
           nthreads=1, simple_sampler=0):
    return RenderParams(seed[0], seed[1], jitter, ao_samples, ao_dist, pt_depth, nthreads, simple_sampler)


def primary(grid, cam, p, image_id=0):
    n = cam.res_x * cam.res_y
    h = np.zeros((7, n), np.float32)
    hp = np.zeros(n, np.uint32)
    st = Stats()
    lib().or_primary(C.byref(grid.mesh.c), C.byref(grid.c), C.byref(cam), C.byref(p), image_id,
                     *[_fp(h[i]) for i in range(7)], _up(hp), C.byref(st))
    return h, hp, st


def _mats(materials, mat_idx):
    m = np.ascontiguousarray(materials, np.float32).reshape(-1, 12)
    mi = None if mat_idx is None else np.ascontiguousarray(mat_idx, np.uint32)
    return m, mi


def render_ao(grid, cam, p, materials, mat_idx=None, image_id=0, fb=None):
    n = cam.res_x * cam.res_y
    fb = np.zeros((n, 3), np.float32) if fb is None else np.ascontiguousarray(fb, np.float32).copy()
    m, mi = _mats(materials, mat_idx)
    occ = np.zeros(n * p.ao_samples, np.uint8)
    st = Stats()
    lib().or_render_ao(C.byref(grid.mesh.c), C.byref(grid.c), C.byref(cam), _fp(m),
                       None if mi is None else _up(mi), C.byref(p), image_id, _fp(fb),
                       occ.ctypes.data_as(_u8p), C.byref(st))
    return fb, occ, st


def render_pt(grid, cam, p, materials, mat_idx, lights, weights=None, image_id=0, fb=None):
    n = cam.res_x * cam.res_y
    fb = np.zeros((n, 3), np.float32) if fb is None else np.ascontiguousarray(fb, np.float32).copy()
    m, mi = _mats(materials, mat_idx)
    L = np.ascontiguousarray(lights, np.float32).reshape(-1, 15)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    st = Stats()
    lib().or_render_pt(C.byref(grid.mesh.c), C.byref(grid.c), C.byref(cam), _fp(m),
                       None if mi is None else _up(mi), _fp(L), None if w is None else _fp(w),
                       len(L), C.byref(p), image_id, _fp(fb), C.byref(st))
    return fb, st


def render_simple(grid, cam, p, image_id=0, fb=None):
    """renderFrame mode 0 (primary rays + simpleShade); p.simple_sampler is set here."""
    n = cam.res_x * cam.res_y
    fb = np.zeros((n, 3), np.float32) if fb is None else np.ascontiguousarray(fb, np.float32).copy()
    q = RenderParams.from_buffer_copy(p)
    q.simple_sampler = 1
    st = Stats()
    lib().or_render_simple(C.byref(grid.mesh.c), C.byref(grid.c), C.byref(cam), C.byref(q), image_id, _fp(fb),
                           C.byref(st))
    return fb, st


def render_pt_multi(grid, cam, p, materials, mat_idx, lights, weights=None, image_id=0, fb=None, state=None):
    """renderFrame mode 1 (the multi-kernel path tracer).  state: (2, npix, 3)
    new radiance and residue carried between images (zeros with a new camera).
    Returns fb, state, end planes (7, npix), end prims, importance (npix, 3), stats."""
    n = cam.res_x * cam.res_y
    fb = np.zeros((n, 3), np.float32) if fb is None else np.ascontiguousarray(fb, np.float32).copy()
    state = np.zeros((2, n, 3), np.float32) if state is None else np.ascontiguousarray(state, np.float32).copy()
    m, mi = _mats(materials, mat_idx)
    L = np.ascontiguousarray(lights, np.float32).reshape(-1, 15)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    end = np.zeros((7, n), np.float32)
    endp = np.zeros(n, np.uint32)
    imp = np.zeros((n, 3), np.float32)
    st = Stats()
    lib().or_render_pt_multi(C.byref(grid.mesh.c), C.byref(grid.c), C.byref(cam), _fp(m),
                             None if mi is None else _up(mi), _fp(L), None if w is None else _fp(w), len(L),
                             C.byref(p), image_id, _fp(fb), _fp(state), _fp(end), _up(endp), _fp(imp),
                             C.byref(st))
    return fb, state, end, endp, imp, st


def render_rays_simple(grid, records, nthreads=1):
    """(n, 3) simple shading of the closest hit of each 24-byte ray record."""
    r = np.ascontiguousarray(records, np.float32).reshape(-1, 6)
    rgb = np.zeros((len(r), 3), np.float32)
    st = Stats()
    lib().or_render_rays_simple(C.byref(grid.mesh.c), C.byref(grid.c), _fp(r), len(r), nthreads, _fp(rgb),
                                C.byref(st))
    return rgb, st


def dump_ao_rays(grid, cam, p, image_id=0):
    """(npix * (ao_samples + 1), 6) ray records as AORayExporter dumps them."""
    n = cam.res_x * cam.res_y * (p.ao_samples + 1)
    out = np.zeros((n, 6), np.float32)
    lib().or_dump_ao_rays(C.byref(grid.mesh.c), C.byref(grid.c), C.byref(cam), C.byref(p), image_id, _fp(out))
    return out


def shading_normal(mesh, prim, org):
    """Path-tracer shading normal of primitive prim seen from hit point org."""
    out = np.zeros(3, np.float32)
    lib().or_shading_normal(C.byref(mesh.c), prim, _fp(_f3(org)), _fp(out))
    return out


def path_trace_rays(grid, org, dirs, p, materials, mat_idx, lights, weights=None):
    """Radiance (n, 3) of a path from each given ray (image 0, ray i = pixel i)."""
    o = np.ascontiguousarray(org, np.float32).reshape(3, -1)
    d = np.ascontiguousarray(dirs, np.float32).reshape(3, -1)
    n = o.shape[1]
    rgb = np.zeros((n, 3), np.float32)
    m, mi = _mats(materials, mat_idx)
    L = np.ascontiguousarray(lights, np.float32).reshape(-1, 15)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    st = Stats()
    lib().or_path_trace_rays(C.byref(grid.mesh.c), C.byref(grid.c), _fp(o[0]), _fp(o[1]), _fp(o[2]),
                             _fp(d[0]), _fp(d[1]), _fp(d[2]), n, _fp(m), None if mi is None else _up(mi),
                             _fp(L), None if w is None else _fp(w), len(L), C.byref(p), _fp(rgb),
                             C.byref(st))
    return rgb, st


def pt_glass(org, dirs, att, n, cos_ne, ior, ray_t, kd, w_fres, w_tir, mode1=False):
    """One glass vertex (or_pt_glass): returns (status, org, dir, att);
    status 0 reflected, 1 refracted, 2 ended."""
    o, d, a = _f3(org).copy(), _f3(dirs).copy(), _f3(att).copy()
    r = lib().or_pt_glass(_fp(o), _fp(d), _fp(a), _fp(_f3(n)), cos_ne, ior, ray_t, _fp(_f3(kd)),
                          w_fres, w_tir, int(bool(mode1)))
    return r, o, d, a


def power_cosine_dir(w0, w1, se, upv):
    out = np.zeros(3, np.float32)
    lib().or_power_cosine_dir(w0, w1, se, _fp(_f3(upv)), _fp(out))
    return out


def cosine_dir(w0, w1, n):
    out = np.zeros(3, np.float32)
    lib().or_cosine_dir(w0, w1, _fp(_f3(n)), _fp(out))
    return out


def pt_glossy(dirs, att, n, cos_ne, ks, se, w0, w1):
    """The megakernel's glossy bounce: returns (alive, dir, att)."""
    d, a = _f3(dirs).copy(), _f3(att).copy()
    r = lib().or_pt_glossy(_fp(d), _fp(a), _fp(_f3(n)), cos_ne, _fp(_f3(ks)), se, w0, w1)
    return r, d, a


def pm_merge(imp, nrad, residue, fb, image_id, rx, ry):
    """Mode 1's merge step on (npix, 3) buffers: returns (fb, residue)."""
    fb = np.ascontiguousarray(fb, np.float32).copy()
    residue = np.ascontiguousarray(residue, np.float32).copy()
    lib().or_pm_merge(_fp(np.ascontiguousarray(imp, np.float32)), _fp(np.ascontiguousarray(nrad, np.float32)),
                      _fp(residue), _fp(fb), image_id, rx, ry)
    return fb, residue


def primary_dirs(cam, p, image_id=0):
    """(3, npix) camera ray directions of every pixel."""
    n = cam.res_x * cam.res_y
    d = np.zeros((3, n), np.float32)
    lib().or_primary_dirs(C.byref(cam), C.byref(p), image_id, _fp(d[0]), _fp(d[1]), _fp(d[2]))
    return d


# ---------------------------------------------------------------------------
# instanced hierarchy
# ---------------------------------------------------------------------------

def instances_struct(objects, matrices, aabb_min, aabb_max):
    objects = np.asarray(objects, np.uint32).reshape(-1)
    m = np.asarray(matrices, np.float32).reshape(-1, 3, 4)
    lo = np.asarray(aabb_min, np.float32).reshape(-1, 3)
    hi = np.asarray(aabb_max, np.float32).reshape(-1, 3)
    arr = (Instance * max(len(objects), 1))()
    for i in range(len(objects)):
        arr[i].object = int(objects[i])
        for r in range(3):
            for c in range(4):
                arr[i].m[r][c] = float(m[i, r, c])
            arr[i].aabb_min[r] = float(lo[i, r])
            arr[i].aabb_max[r] = float(hi[i, r])
    return arr


class OHier:
    """Owns an or_hier built over mesh (current frame) with object bounds from
    the key-frame vertices key (3, nv) and the padded scene bounds."""

    def __init__(self, mesh, key, bmin, bmax, obj_first, objects, matrices, aabb_min, aabb_max,
                 lambda_top=1.2, lambda_leaf=2.2):
        self.mesh = mesh
        self.key = np.ascontiguousarray(key, np.float32).reshape(3, -1)
        self.kx, self.ky, self.kz = (np.ascontiguousarray(self.key[i]) for i in range(3))
        self.obj_first = np.ascontiguousarray(obj_first, np.uint32)
        self.inst = instances_struct(objects, matrices, aabb_min, aabb_max)
        self.ninst = len(np.asarray(objects).reshape(-1))
        self.c = Hier()
        self.status = lib().or_build_hier(C.byref(mesh.c), _fp(self.kx), _fp(self.ky), _fp(self.kz),
                                          _fp(_f3(bmin)), _fp(_f3(bmax)), _up(self.obj_first),
                                          len(self.obj_first) - 1, self.inst, self.ninst, lambda_top,
                                          lambda_leaf, C.byref(self.c))

    def __del__(self):
        try:
            lib().or_hier_free(C.byref(self.c))
        except Exception:
            pass

    def top_params(self):
        g = self.c.top
        return {"bmin": np.array(g.bmin[:], np.float32), "bmax": np.array(g.bmax[:], np.float32),
                "res": tuple(g.res), "cell_size": np.array(g.cell_size[:], np.float32),
                "cell_size_rcp": np.array(g.cell_size_rcp[:], np.float32)}

    def top_cells(self):
        n = 2 * self.c.top.num_cells
        return np.ctypeslib.as_array(self.c.top.cells, shape=(n,)).copy().reshape(-1, 2)

    def inst_refs(self):
        n = self.c.num_inst_refs
        return np.zeros(0, np.uint32) if n == 0 else np.ctypeslib.as_array(self.c.inst_refs, shape=(n,)).copy()

    def obj_cells(self):
        n = 2 * self.c.num_obj_cells
        return np.zeros((0, 2), np.uint32) if n == 0 else \
            np.ctypeslib.as_array(self.c.obj_cells, shape=(n,)).copy().reshape(-1, 2)

    def prims(self):
        n = self.c.num_obj_refs
        return np.zeros(0, np.uint32) if n == 0 else np.ctypeslib.as_array(self.c.prims, shape=(n,)).copy()

    def objects(self):
        out = []
        for k in range(self.c.nobj):
            g = self.c.objs[k]
            out.append(dict(bmin=np.array(g.bmin[:], np.float32), bmax=np.array(g.bmax[:], np.float32),
                            res=list(g.res), cell_size=np.array(g.cell_size[:], np.float32),
                            cell_size_rcp=np.array(g.cell_size_rcp[:], np.float32),
                            first_cell=g.first_cell, num_cells=g.num_cells))
        return out


def trace_hier(h, org, dirs, anyhit=False, tmax=3.402823466e38):
    o = np.ascontiguousarray(org, np.float32).reshape(3, -1)
    d = np.ascontiguousarray(dirs, np.float32).reshape(3, -1)
    n = o.shape[1]
    ox, oy, oz = (np.ascontiguousarray(o[i]) for i in range(3))
    dx, dy, dz = (np.ascontiguousarray(d[i]) for i in range(3))
    t = np.zeros(n, np.float32)
    prim = np.zeros(n, np.uint32)
    inst = np.zeros(n, np.uint32)
    occ = np.zeros(n, np.uint8)
    lib().or_trace_hier(C.byref(h.mesh.c), C.byref(h.c), _fp(ox), _fp(oy), _fp(oz), _fp(dx), _fp(dy), _fp(dz), n,
                        1 if anyhit else 0, tmax, _fp(t), _up(prim), _up(inst), occ.ctypes.data_as(_u8p))
    return (occ, t) if anyhit else (t, prim, inst)


def render_ao_hier(h, cam, p, materials, mat_idx=None, image_id=0, fb=None):
    n = cam.res_x * cam.res_y
    fb = np.zeros((n, 3), np.float32) if fb is None else np.ascontiguousarray(fb, np.float32).copy()
    m, mi = _mats(materials, mat_idx)
    occ = np.zeros(n * p.ao_samples, np.uint8)
    hits = np.zeros((7, n), np.float32)
    hp = np.zeros(n, np.uint32)
    hi = np.zeros(n, np.uint32)
    lib().or_render_ao_hier(C.byref(h.mesh.c), C.byref(h.c), C.byref(cam), _fp(m),
                            None if mi is None else _up(mi), C.byref(p), image_id, _fp(fb),
                            occ.ctypes.data_as(_u8p), _fp(hits), _up(hp), _up(hi))
    return fb, occ, hits, hp, hi
