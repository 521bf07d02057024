"""ctypes binding of libhappyray.so (include/happyray.h).

The library is loaded from this package directory only.  When it is missing
the import fails loudly: there is no CPU fallback behind this module.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhappyray.so")

HR_OK = 0
STATUS_NAMES = {0: "HR_OK", 1: "HR_ERR_ARG", 2: "HR_ERR_DEVICE", 3: "HR_ERR_OOM", 4: "HR_ERR_IO",
                5: "HR_ERR_STATE", 6: "HR_ERR_CAPACITY", 7: "HR_ERR_PARSE", 8: "HR_ERR_BUDGET"}

GRID_UNIFORM = 0
GRID_UNIFORM_EXACT = 1
GRID_TWO_LEVEL = 2
GRID_HIERARCHY = 3
MODE_SIMPLE = 0
MODE_PATH_MULTI = 1
MODE_PATH = 2
MODE_AO = 3

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p


class Options(C.Structure):
    _fields_ = [("device", C.c_int), ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32)]


class GridInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
                ("res", C.c_int * 3), ("cell_size", C.c_float * 3), ("cell_size_rcp", C.c_float * 3),
                ("num_cells", C.c_uint32), ("num_leaves", C.c_uint32), ("num_refs", C.c_uint32),
                ("num_top_pairs", C.c_uint32), ("num_triangles", C.c_uint32),
                ("num_objects", C.c_uint32), ("num_instances", C.c_uint32)]


class Instance(C.Structure):
    """hr_instance: object id, world-to-object rows m_row_0..2, world box."""
    _fields_ = [("object", C.c_uint32), ("m", (C.c_float * 4) * 3), ("aabb_min", C.c_float * 3),
                ("aabb_max", C.c_float * 3)]


class ObjectGrid(C.Structure):
    _fields_ = [("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("res", C.c_int * 3),
                ("cell_size", C.c_float * 3), ("cell_size_rcp", C.c_float * 3), ("first_cell", C.c_uint32),
                ("num_cells", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("jitter", C.c_int), ("ao_samples", C.c_int), ("ao_dist", C.c_float),
                ("pt_depth", C.c_int), ("collect_stats", C.c_int)]


class Timing(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("render_ms", C.c_double), ("primary_ms", C.c_double),
                ("secondary_ms", C.c_double), ("shade_ms", C.c_double), ("rays_traced", C.c_uint64),
                ("primary_kernel", C.c_int), ("frames_replayed", C.c_uint64), ("replay_build_ms", C.c_double),
                ("replay_render_ms", C.c_double), ("replay_primary_ms", C.c_double),
                ("replay_secondary_ms", C.c_double), ("replay_shade_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("persistent_primary", C.c_int), ("persistent_secondary", C.c_int),
                ("refill", C.c_int), ("hier_refill", C.c_int), ("hier_bounds_blocks", C.c_int),
                ("persistent_path", C.c_int), ("path_refill", C.c_int), ("path_shade", C.c_int),
                ("fuse_units", C.c_int), ("test_batch", C.c_int), ("advance_steps", C.c_int),
                ("frame_lanes", C.c_int), ("camera_blocks", C.c_int)]


class WorkCounts(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("steps", C.c_uint64), ("tests", C.c_uint64)]


class WorkStats(C.Structure):
    _fields_ = [("primary", WorkCounts), ("secondary", WorkCounts)]


class SceneInfo(C.Structure):
    _fields_ = [("num_frames", C.c_uint32), ("num_vertices", C.c_uint32), ("num_normals", C.c_uint32),
                ("num_triangles", C.c_uint32), ("num_materials", C.c_uint32),
                ("num_lights", C.c_uint32), ("frame_step", C.c_float), ("has_camera", C.c_int),
                ("cam_pos", C.c_float * 3), ("cam_dir", C.c_float * 3), ("cam_up", C.c_float * 3),
                ("cam_fov", C.c_float), ("cam_res_x", C.c_int), ("cam_res_y", C.c_int),
                ("num_objects", C.c_uint32), ("num_instances", C.c_uint32)]


class MeshDesc(C.Structure):
    _fields_ = [("xyz", _f32p), ("nv", C.c_uint32), ("idx", _u32p), ("nt", C.c_uint32)]


class RayBatch(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("n", C.c_uint32), ("anyhit", C.c_int), ("tmax", C.c_float),
                ("org", _vp), ("dir", _vp), ("t_out", _vp), ("prim_out", _vp), ("occ_out", _vp)]


class BatchTuning(C.Structure):
    _fields_ = [("refill", C.c_int), ("blocks_per_cu", C.c_int)]


class AutotuneResult(C.Structure):
    _fields_ = [("lambda_top", C.c_float * 4), ("lambda_leaf", C.c_float * 4), ("tuning", BatchTuning),
                ("evaluations", C.c_int), ("best_build_trace_ms", C.c_float),
                ("best_trace_ms", C.c_float)]


class ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int), ("memory_budget", C.c_uint64), ("max_wave_rays", C.c_uint32),
                ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("scene", C.c_uint32), ("n", C.c_uint32), ("tmax", C.c_float),
                ("depth", C.c_int), ("on_device", C.c_int), ("org", _vp), ("dir", _vp),
                ("stride", C.c_uint32), ("t_out", _vp),
                ("prim_out", _vp), ("occ_out", _vp), ("rgb_out", _vp), ("k1", C.c_uint32),
                ("k2", C.c_uint32), ("coeff", C.c_float), ("layout", C.c_int)]


RAYS_SOA, RAYS_RECORDS = 0, 1


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("latency_ms", C.c_double), ("wave", C.c_uint64)]


class ServerScene(C.Structure):
    _fields_ = [("keys", _f32p), ("nframes", C.c_uint32), ("nv", C.c_uint32), ("idx", _u32p),
                ("nt", C.c_uint32), ("mats", _f32p), ("nmat", C.c_uint32), ("mat_idx", _u32p),
                ("lights", _f32p), ("light_weights", _f32p), ("nlights", C.c_int),
                ("lambda_top", C.c_float), ("lambda_leaf", C.c_float)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("rays", C.c_uint64), ("waves", C.c_uint64),
                ("builds", C.c_uint64), ("rebuilds", C.c_uint64), ("evictions", C.c_uint64),
                ("resident_bytes", C.c_uint64), ("peak_resident_bytes", C.c_uint64),
                ("resident_scenes", C.c_uint32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "hr_last_error": (C.c_char_p, []),
    "hr_create": (C.c_int, [C.POINTER(_vp), C.POINTER(Options)]),
    "hr_destroy": (C.c_int, [_vp]),
    "hr_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "hr_upload_mesh": (C.c_int, [_vp, _f32p, _f32p, C.c_float, C.c_uint32, _u32p, C.c_uint32]),
    "hr_upload_materials": (C.c_int, [_vp, _f32p, C.c_uint32, _u32p, C.c_uint32]),
    "hr_set_keyframes": (C.c_int, [_vp, _f32p, C.c_uint32, C.c_uint32, _u32p, C.c_uint32]),
    "hr_set_frame": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float]),
    "hr_set_grid": (C.c_int, [_vp, C.c_int, C.c_float, C.c_float]),
    "hr_build": (C.c_int, [_vp]),
    "hr_get_grid_info": (C.c_int, [_vp, C.POINTER(GridInfo)]),
    "hr_get_grid": (C.c_int, [_vp, _u32p, _u32p, _u32p]),
    "hr_set_instances": (C.c_int, [_vp, _u32p, C.c_uint32, C.POINTER(Instance), C.c_uint32]),
    "hr_get_object_grids": (C.c_int, [_vp, C.POINTER(ObjectGrid), _u32p]),
    "hr_trace_instanced": (C.c_int, [_vp, _f32p, _f32p, C.c_uint32, _f32p, _u32p, _u32p]),
    "hr_get_primary_instances": (C.c_int, [_vp, _u32p]),
    "hr_set_camera": (C.c_int, [_vp, _f32p, _f32p, _f32p, C.c_float, C.c_int, C.c_int]),
    "hr_set_lights": (C.c_int, [_vp, _f32p, _f32p, C.c_int]),
    "hr_render": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(RenderParams), _f32p]),
    "hr_get_primary_hits": (C.c_int, [_vp, _f32p, _u32p]),
    "hr_get_ao_occlusion": (C.c_int, [_vp, _u8p]),
    "hr_trace": (C.c_int, [_vp, _f32p, _f32p, C.c_uint32, _f32p, _u32p]),
    "hr_occluded": (C.c_int, [_vp, _f32p, _f32p, C.c_uint32, C.c_float, _u8p]),
    "hr_trace_dev": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp, _vp]),
    "hr_occluded_dev": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_float, _vp]),
    "hr_path_trace": (C.c_int, [_vp, _f32p, _f32p, C.c_uint32, C.c_int, _f32p]),
    "hr_path_trace_dev": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_int, _vp]),
    "hr_path_trace_dev_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_int, _vp]),
    "hr_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "hr_get_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "hr_get_timing": (C.c_int, [_vp, C.POINTER(Timing)]),
    "hr_get_work_stats": (C.c_int, [_vp, C.POINTER(WorkStats)]),
    "hr_synchronize": (C.c_int, [_vp]),
    "hr_render_frame_graph": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_int, C.c_int,
                                         C.POINTER(RenderParams), _f32p]),
    "hr_submit_frame_graph": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_int, C.c_int,
                                         C.POINTER(RenderParams)]),
    "hr_get_framebuffer": (C.c_int, [_vp, _f32p]),
    "hr_batch_create": (C.c_int, [C.POINTER(_vp), C.POINTER(Options)]),
    "hr_batch_destroy": (C.c_int, [_vp]),
    "hr_batch_get_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "hr_batch_set_scenes": (C.c_int, [_vp, C.POINTER(MeshDesc), C.c_uint32]),
    "hr_batch_set_densities": (C.c_int, [_vp, _f32p, _f32p]),
    "hr_batch_build": (C.c_int, [_vp]),
    "hr_batch_get_grid_info": (C.c_int, [_vp, C.c_uint32, C.POINTER(GridInfo)]),
    "hr_batch_get_grid": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, _u32p]),
    "hr_batch_trace_dev": (C.c_int, [_vp, C.POINTER(RayBatch), C.c_uint32]),
    "hr_batch_set_tuning": (C.c_int, [_vp, C.POINTER(BatchTuning)]),
    "hr_batch_get_tuning": (C.c_int, [_vp, C.POINTER(BatchTuning), _f32p, _f32p]),
    "hr_batch_autotune": (C.c_int, [_vp, C.POINTER(RayBatch), C.c_uint32, C.POINTER(AutotuneResult)]),
    "hr_batch_get_timing": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_uint64)]),
    "hr_batch_get_work_stats": (C.c_int, [_vp, C.POINTER(WorkStats)]),
    "hr_batch_synchronize": (C.c_int, [_vp]),
    "hr_server_create": (C.c_int, [C.POINTER(_vp), C.POINTER(ServerOptions)]),
    "hr_server_destroy": (C.c_int, [_vp]),
    "hr_server_add_scene": (C.c_int, [_vp, C.POINTER(ServerScene), C.POINTER(C.c_uint32)]),
    "hr_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "hr_server_wait": (C.c_int, [_vp, C.c_uint64, C.POINTER(RequestResult)]),
    "hr_server_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(C.c_int)]),
    "hr_server_get_stats": (C.c_int, [_vp, C.POINTER(ServerStats)]),
    "hr_set_normals": (C.c_int, [_vp, _f32p, C.c_uint32, C.c_uint32, _u32p, C.c_uint32]),
    "hr_scene_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "hr_scene_load_obj": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "hr_scene_free": (C.c_int, [_vp]),
    "hr_scene_get_info": (C.c_int, [_vp, C.POINTER(SceneInfo)]),
    "hr_scene_get_vertices": (C.c_int, [_vp, C.c_uint32, _f32p]),
    "hr_scene_get_normals": (C.c_int, [_vp, C.c_uint32, _f32p]),
    "hr_scene_get_triangles": (C.c_int, [_vp, _u32p, _u32p, _u32p]),
    "hr_scene_get_materials": (C.c_int, [_vp, _f32p]),
    "hr_scene_get_lights": (C.c_int, [_vp, _f32p, _f32p]),
    "hr_scene_get_instances": (C.c_int, [_vp, _u32p, C.POINTER(Instance)]),
    "hr_scene_upload": (C.c_int, [_vp, _vp]),
    "hr_scene_frame": (C.c_int, [_vp, C.c_float, _u32p, _u32p, _f32p]),
    "hr_scene_next_frame": (C.c_int, [_vp, _f32p]),
    "hr_trace_wire": (C.c_int, [_vp, _f32p, C.c_uint32, _f32p, _u32p]),
    "hr_occluded_wire": (C.c_int, [_vp, _f32p, C.c_uint32, C.c_float, _u8p]),
    "hr_render_rays": (C.c_int, [_vp, _f32p, C.c_uint32, _f32p]),
    "hr_dump_ao_rays": (C.c_int, [_vp, _f32p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "hr_image_to_rgb8": (C.c_int, [_f32p, C.c_int, C.c_int, C.c_float, _u8p]),
    "hr_write_png": (C.c_int, [C.c_char_p, _u8p, C.c_int, C.c_int]),
    "hr_write_screenshot": (C.c_int, [C.c_char_p, _f32p, C.c_int, C.c_int]),
}

_lib = None


def load():
    """Load libhappyray.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m happyray_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class HappyRayError(RuntimeError):
    def __init__(self, status, where, message):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {message}")


def check(status, where):
    if status != HR_OK:
        msg = load().hr_last_error()
        raise HappyRayError(status, where, msg.decode() if msg else "")
