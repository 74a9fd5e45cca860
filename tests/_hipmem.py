"""Device buffers for GPU tests through the HIP runtime the library itself
links (libamdhip64.so.7 from ROCm): hipMalloc, hipMemcpy, hipFree by ctypes.
torch ships a HIP runtime of its own, which does not share devices with the
library's in one process, so tests that need raw device pointers use this."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import ipglasma_amd as I
        I.lib()  # the library first: its runtime is the one to share
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceArray:
    """float64 device array of a given shape"""

    def __init__(self, shape, host=None):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        if host is not None:
            self.upload(host)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1)  # host to device
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")

    def download(self):
        out = np.empty(self.shape)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)  # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
