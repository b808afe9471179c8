"""Minimal ctypes access to the HIP runtime that libwacomm_hip.so links
(device buffers for the drop-in entry points, which take device pointers)."""
import ctypes as C

import numpy as np

import wacommplusplus_amd as wc

wc.load_library()                          # loads libamdhip64.so.7 first: the same runtime is reused
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, host: np.ndarray):
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.ptr), max(1, self.nbytes)) == 0
        host = np.ascontiguousarray(host)
        assert _hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert _hip.hipDeviceSynchronize() == 0
        assert _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = None


def synchronize():
    assert _hip.hipDeviceSynchronize() == 0
