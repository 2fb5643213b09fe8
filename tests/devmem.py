"""Device buffers for the GPU tests through the HIP runtime libmhd runs on.

The tests hand device pointers to the _d entry points of libmhd.  The
buffers come from the same HIP runtime the library is linked against (the
process's libamdhip64.so.7, already loaded with libmhd), so no second GPU
runtime (for example the one bundled with PyTorch) is initialised beside it.
"""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import gridapmhd_amd as g
        g._lib.load()                                # libmhd first: its libamdhip64.so.7
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A 1-D device array of a numpy dtype, freed with the object."""

    def __init__(self, n, dtype):
        self.n, self.dtype = int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.copy_from(a)
        return d

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        self._check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D))

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        self._check(hip().hipDeviceSynchronize())
        self._check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H))
        return out

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise RuntimeError(f"HIP call failed ({rc})")

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = None
