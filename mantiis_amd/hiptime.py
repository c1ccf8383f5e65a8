"""HIP events through ctypes on an arbitrary hipStream_t (the library's own
stream), for timing inside bench.py."""
import ctypes as C
import os

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = C.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("libamdhip64 not found")
        _HIP.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _HIP.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _HIP.hipEventSynchronize.argtypes = [C.c_void_p]
        _HIP.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _HIP.hipEventDestroy.argtypes = [C.c_void_p]
        _HIP.hipDeviceSynchronize.argtypes = []
    return _HIP


class StreamTimer:
    """Elapsed milliseconds between two events recorded on `stream`."""

    def __init__(self, stream):
        h = _hip()
        self.stream = C.c_void_p(stream)
        self.a, self.b = C.c_void_p(), C.c_void_p()
        assert h.hipEventCreate(C.byref(self.a)) == 0 and h.hipEventCreate(C.byref(self.b)) == 0

    def start(self):
        assert _hip().hipEventRecord(self.a, self.stream) == 0

    def stop(self):
        h = _hip()
        assert h.hipEventRecord(self.b, self.stream) == 0
        assert h.hipEventSynchronize(self.b) == 0
        ms = C.c_float()
        assert h.hipEventElapsedTime(C.byref(ms), self.a, self.b) == 0
        return float(ms.value)

    def close(self):
        h = _hip()
        h.hipEventDestroy(self.a)
        h.hipEventDestroy(self.b)


def device_synchronize():
    assert _hip().hipDeviceSynchronize() == 0
