"""HIP event timing on an arbitrary stream (ctypes over libamdhip64).

torch.cuda.Event only sees torch's current stream; the library launches on
its own context stream, so events are recorded on that stream directly.
"""
import ctypes as C
import os

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
        _H.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _H.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _H.hipEventSynchronize.argtypes = [C.c_void_p]
        _H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _H.hipEventDestroy.argtypes = [C.c_void_p]
    return _H


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        rc = _hip().hipEventCreate(C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate rc={rc}")

    def record(self, stream):
        rc = _hip().hipEventRecord(self.h, C.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord rc={rc}")

    def synchronize(self):
        _hip().hipEventSynchronize(self.h)

    def elapsed_ms(self, end):
        out = C.c_float(0)
        rc = _hip().hipEventElapsedTime(C.byref(out), self.h, end.h)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime rc={rc}")
        return float(out.value)

    def __del__(self):
        try:
            _hip().hipEventDestroy(self.h)
        except Exception:
            pass


def copy_probe(nbytes, reps=3):
    """Device-to-device copies of nbytes (a plain copy kernel of known size,
    used to calibrate the FETCH_SIZE / WRITE_SIZE counters)."""
    h = _hip()
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(a), nbytes) or h.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    for _ in range(reps):
        if h.hipMemcpy(b, a, nbytes, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    h.hipDeviceSynchronize()
    h.hipFree(a)
    h.hipFree(b)
