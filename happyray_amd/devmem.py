"""Device buffers through the HIP runtime (hipMalloc / hipMemcpy / hipFree),
for callers that hand device pointers to the _dev entry points without
another framework."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipMemcpy2DAsync.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                          C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def copy_planes(dst, src, n, src_stride, planes=3, stream=None):
    """Copy `planes` rows of n floats from device address src (rows
    src_stride floats apart) to dst (rows n floats apart), on `stream`."""
    rc = hip().hipMemcpy2DAsync(dst, 4 * n, src, 4 * src_stride, 4 * n, planes, D2D, stream)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy2DAsync failed with {rc}")


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """A device allocation of `shape` elements of `dtype` (C order)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1))
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed with {rc}")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed with {rc}")

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed with {rc}")
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
