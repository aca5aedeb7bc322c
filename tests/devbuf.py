"""A device buffer for tests (ctypes over libamdhip64): hipMalloc'ed bytes,
a blocking read back, hipFree on close."""
import ctypes as C

import numpy as np

_H = None


def _hip():
    global _H
    if _H is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _H = C.CDLL(name)
                break
            except OSError:
                continue
        if _H is None:
            raise RuntimeError("libamdhip64.so not found")
        _H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _H.hipFree.argtypes = [C.c_void_p]
    return _H


class DeviceArray:
    """n float64 values in device memory, filled with `fill` at creation."""

    def __init__(self, n, fill=np.nan):
        self.n = n
        self.p = C.c_void_p()
        if _hip().hipMalloc(C.byref(self.p), n * 8):
            raise RuntimeError("hipMalloc failed")
        h = np.full(n, fill, dtype=np.float64)
        if _hip().hipMemcpy(self.p, h.ctypes.data_as(C.c_void_p), n * 8, 1):   # host to device
            raise RuntimeError("hipMemcpy failed")

    @property
    def ptr(self):
        return self.p.value

    def read(self):
        h = np.zeros(self.n, dtype=np.float64)
        if _hip().hipMemcpy(h.ctypes.data_as(C.c_void_p), self.p, self.n * 8, 2):   # device to host
            raise RuntimeError("hipMemcpy failed")
        return h

    def close(self):
        if self.p:
            _hip().hipFree(self.p)
            self.p = C.c_void_p()
