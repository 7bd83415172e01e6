"""Device memory for tests, from the HIP runtime the library itself is linked against.

The library resolves libamdhip64.so.7 from /opt/rocm/lib; loading the same soname
here gives the same runtime instance, so addresses allocated here are valid
arguments of the library's device-pointer entry points.  (torch ships its own HIP
runtime, which is a second instance in the process.)
"""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """nbytes of device memory, freed with the object"""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.nbytes) == 0, "hipMalloc failed"

    @property
    def addr(self):
        return int(self.ptr.value)

    def upload(self, array, offset=0):
        a = np.ascontiguousarray(array)
        assert offset + a.nbytes <= self.nbytes
        assert hip().hipMemcpy(C.c_void_p(self.addr + offset), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def download(self, dtype=np.uint8, count=None, offset=0):
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count is None else count
        out = np.empty(n, dt)
        assert offset + out.nbytes <= self.nbytes
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.addr + offset), out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
