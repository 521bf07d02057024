"""Framebuffers to image files over the C ABI (hr_image_to_rgb8, hr_write_png,
hr_write_screenshot): Image::gammaCorrect and Image::writePNG
(src/Utils/ImagePNG.cpp:64-72, :268-330) as SDLGLApplication::writeScreenShot
uses them (src/Application/SDLGLApplication.cpp:120-143)."""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check

_f32p = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)

CLI_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "happyray")


def _rgb(fb, width, height):
    a = np.ascontiguousarray(fb, np.float32).reshape(-1)
    if a.size != 3 * width * height:
        raise ValueError(f"framebuffer has {a.size} floats, {3 * width * height} expected")
    return a


def to_rgb8(fb, width, height, gamma=2.2):
    """(height, width, 3) uint8: powf(v, 1/gamma), clamped to [0, 1], x255, truncated."""
    a = _rgb(fb, width, height)
    out = np.zeros((height, width, 3), np.uint8)
    check(_lib.load().hr_image_to_rgb8(a.ctypes.data_as(_f32p), width, height, float(gamma),
                                       out.ctypes.data_as(_u8p)), "hr_image_to_rgb8")
    return out


def write_png(path, rgb8):
    b = np.ascontiguousarray(rgb8, np.uint8)
    h, w = b.shape[0], b.shape[1]
    check(_lib.load().hr_write_png(os.fsencode(path), b.ctypes.data_as(_u8p), w, h), "hr_write_png")


def write_screenshot(path, fb, width, height):
    """The framebuffer as SDLGLApplication::writeScreenShot writes it (gamma 2.2)."""
    a = _rgb(fb, width, height)
    check(_lib.load().hr_write_screenshot(os.fsencode(path), a.ctypes.data_as(_f32p), width, height),
          "hr_write_screenshot")
