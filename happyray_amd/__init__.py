"""happyray_amd -- MI355X (gfx950) native per-frame hot path of the happyray
ray tracer: key-frame interpolation, uniform / two-level grid builds, 3D-DDA
traversal with Moller-Trumbore intersection, ambient occlusion and path
tracing, behind the C ABI in include/happyray.h.
"""
from ._lib import (GRID_HIERARCHY, GRID_TWO_LEVEL, GRID_UNIFORM, GRID_UNIFORM_EXACT, MODE_AO, MODE_PATH, MODE_PATH_MULTI,
                   MODE_SIMPLE, HappyRayError, LIB_PATH, load)
from .engine import Engine
from .scenefile import SceneFile
from .batch import Batch
from .server import Server

__all__ = ["Engine", "Batch", "Server", "SceneFile", "HappyRayError", "load", "LIB_PATH", "GRID_UNIFORM", "GRID_UNIFORM_EXACT",
           "GRID_TWO_LEVEL", "GRID_HIERARCHY", "MODE_AO", "MODE_PATH", "MODE_PATH_MULTI", "MODE_SIMPLE"]
