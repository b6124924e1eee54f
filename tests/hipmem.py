"""Device buffers for GPU tests through the HIP runtime the library itself links
(/opt/rocm/lib/libamdhip64.so.7), via ctypes: hipMalloc / hipMemcpy / hipFree.  (torch ships
its own HIP runtime, which does not share devices with the library's in one process.)"""
import ctypes

import numpy as np

_H2D, _D2H = 1, 2


def _rt():
    import bem3_emf_qd2_amd as bq
    bq.lib()  # loads libamdhip64.so.7 as a dependency of the library
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DeviceArray:
    """A device copy of a numpy array (same dtype and shape); .ptr is the device address."""

    def __init__(self, a):
        self.rt = _rt()
        self.a = np.ascontiguousarray(a)
        self.ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(self.ptr), max(self.a.nbytes, 1)) == 0
        assert self.rt.hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, _H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, self.a.nbytes, _D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.rt.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
