"""Device buffers for GPU tests through the HIP runtime libevoc_hip.so links
(/opt/rocm/lib/libamdhip64.so), so tests can hand the library device pointers without a
second runtime. (torch ships a HIP runtime of its own; in one process it has to initialise
the GPU before libevoc_hip does, see conftest.py.)"""
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
    return _hip


class DeviceArray:
    """A host array copied into device memory (any dtype); freed with the object."""

    def __init__(self, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype)
        self.shape, self.dtype = a.shape, a.dtype
        self.nbytes = a.nbytes
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(a.nbytes, 4)) == 0
        if a.nbytes:
            assert hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # to host
        return out

    def __del__(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None
