"""Device buffers for the GPU tests (plumbing: hipMalloc / hipMemcpy / hipFree through ctypes)."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        _hip = h
    return _hip


class DevBuf:
    """`nbytes` of device memory; .ptr is the device address (int)."""

    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        rc = hip().hipMalloc(ctypes.byref(self.p), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed: {rc}")

    @property
    def ptr(self):
        return self.p.value

    def get(self, dtype, count, offset_elems=0):
        out = np.empty(count, dtype=dtype)
        src = self.p.value + offset_elems * out.itemsize
        assert offset_elems * out.itemsize + out.nbytes <= self.nbytes
        rc = hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(src), out.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        rc = hip().hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
