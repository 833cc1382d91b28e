"""Device arrays for the *_dev entry points, allocated with hipMalloc through
ctypes.  The HIP functions are looked up through the library's own handle
(its dependency tree holds the HIP runtime it runs on), so the arrays come
from that runtime and no second copy of it is opened next to torch's."""
import ctypes

import numpy as np

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        from cryoim_amd import lib
        h = lib.load()
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


class DevArray:
    """float32 device buffer with the data_ptr()/numel() of a tensor."""

    def __init__(self, n: int):
        self.n = int(n)
        self.p = ctypes.c_void_p()
        rc = _hip().hipMalloc(ctypes.byref(self.p), max(4 * self.n, 4))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.size)
        if _hip().hipMemcpy(d.p, a.ctypes.data_as(ctypes.c_void_p), 4 * a.size, 1) != 0:
            raise RuntimeError("hipMemcpy to device failed")
        return d

    def numpy(self, shape=None):
        out = np.empty(self.n, np.float32)
        if _hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, 4 * self.n, 2) != 0:
            raise RuntimeError("hipMemcpy to host failed")
        return out if shape is None else out.reshape(shape)

    def data_ptr(self):
        return self.p.value

    def numel(self):
        return self.n

    @property
    def shape(self):
        return (self.n,)

    def free(self):
        if self.p:
            _hip().hipFree(self.p)
            self.p = ctypes.c_void_p()
