"""Device buffers for tests that call the device-pointer entry points: hipMalloc /
hipMemcpy / hipFree of the HIP runtime that libfcg.so itself links (a second HIP
runtime in the process, such as torch's, would not see the GPU)."""
import ctypes as C

import numpy as np

from fastercap_amd import _ffi

_H2D, _D2H = 1, 2
_hip = None


def hip():
    global _hip
    if _hip is None:
        _ffi.device_lib()  # loads libamdhip64.so.7 through libfcg.so's run path
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A float64 device array of a given shape."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8))
        assert rc == 0, f"hipMalloc {self.nbytes} B: {rc}"

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        assert hip().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, _H2D) == 0
        return d

    def fill_nan(self):
        nan = np.full(self.shape, np.nan)
        assert hip().hipMemcpy(self.ptr, nan.ctypes.data_as(C.c_void_p), self.nbytes, _H2D) == 0

    def to_host(self):
        assert hip().hipDeviceSynchronize() == 0
        out = np.empty(self.shape)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, _D2H) == 0
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
