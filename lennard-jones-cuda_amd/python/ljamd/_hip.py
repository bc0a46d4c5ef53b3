"""Minimal device buffers through the HIP runtime (hipMalloc/hipMemcpy), for
passing numpy data to the device-pointer entry points without torch."""
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
    return _hip


H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(self.ptr), self.nbytes) != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def download(self, dtype, shape):
        out = np.empty(shape, dtype)
        if hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
