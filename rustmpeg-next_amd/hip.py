"""Minimal ctypes access to the HIP runtime librmh is linked against: device
buffers, streams and events for callers of the plan entry points from Python.

Load librmh first (``load_library``) so this binds to the same runtime
instance; a process that imports torch first instead runs librmh on torch's
copy of libamdhip64 (same soname), which works the same way.
"""
import ctypes

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def lib():
    global _hip
    if _hip is None:
        from . import load_library
        load_library()
        _hip = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        _hip.hipMemsetAsync.argtypes = [vp, ctypes.c_int, sz, vp]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
    return _hip


def check(e):
    if e != 0:
        raise RuntimeError("HIP error %d: %s" % (e, lib().hipGetErrorString(e).decode()))


class DeviceBuffer:
    """A device allocation (freed on close / garbage collection)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(p), max(int(nbytes), 1)))
        self.ptr, self.nbytes = p.value, int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D))
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        check(lib().hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, H2D))

    def download(self, dtype, count=None, offset=0):
        dtype = np.dtype(dtype)
        n = (self.nbytes - offset) // dtype.itemsize if count is None else count
        out = np.empty(n, dtype)
        check(lib().hipMemcpy(out.ctypes.data, self.ptr + offset, out.nbytes, D2H))
        return out

    def zero(self, stream=None):
        check(lib().hipMemsetAsync(self.ptr, 0, self.nbytes, stream))

    def close(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(s)))
        self.handle = s.value

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle))

    def close(self):
        if self.handle:
            lib().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        check(lib().hipEventRecord(self.handle, stream))

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle))

    def elapsed_ms(self, end):
        t = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(t), self.handle, end.handle))
        return t.value

    def close(self):
        if self.handle:
            lib().hipEventDestroy(self.handle)
            self.handle = None


def device_synchronize():
    check(lib().hipDeviceSynchronize())
