"""Device arrays allocated through the library's own C ABI (mm_device_alloc),
so one process holds exactly one HIP runtime (the one libmmhip.so links).

Layout is the reference's: scalar fields (nz, ny, nx), vector fields
(3, nz, ny, nx), contiguous.
"""
import ctypes

import numpy as np

from . import _lib

__all__ = ["DeviceArray", "to_device", "empty", "empty_like", "gpu_available"]


class DeviceArray:
    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise TypeError("DeviceArray supports float64 and float32")
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        L = _lib.device_lib()
        p = ctypes.c_void_p()
        _lib.check(L.mm_device_alloc(ctypes.c_size_t(self.nbytes), ctypes.byref(p)))
        self._ptr = p

    @property
    def ptr(self):
        return self._ptr

    def numel(self):
        return self.size

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.shape != self.shape:
            a = np.ascontiguousarray(np.broadcast_to(a, self.shape), dtype=self.dtype)
        _lib.check(_lib.load().mm_memcpy_htod(self._ptr, a.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_size_t(self.nbytes), None))
        return self

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _lib.check(_lib.load().mm_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self._ptr,
                                              ctypes.c_size_t(self.nbytes), None))
        return out

    def __del__(self):
        try:
            if self._ptr and _lib._LIB is not None:
                _lib._LIB.mm_device_free(self._ptr)
                self._ptr = None
        except Exception:
            pass


def to_device(a, dtype=np.float64):
    a = np.asarray(a)
    return DeviceArray(a.shape, dtype).upload(a)


def empty(shape, dtype=np.float64):
    return DeviceArray(shape, dtype)


def empty_like(d):
    return DeviceArray(d.shape, d.dtype)


def gpu_available():
    """True when libmmhip.so is built and its HIP runtime sees a device."""
    try:
        _lib.device_lib()
        return True
    except Exception:
        return False
