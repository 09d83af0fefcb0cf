"""Device buffers, streams and events on the HIP runtime libkrr_hip.so uses.

A process must not hold two HIP runtimes: PyTorch wheels ship their own
libamdhip64 (same soname), and a second runtime initialised after the first
one sees no device.  These helpers bind the runtime the tracer library
itself resolved, so callers that do not need PyTorch (the tests, the
benchmark) keep a single runtime.  Callers that do use PyTorch must import
torch before loading the tracer library so that both share torch's runtime.
"""
import ctypes as C

import numpy as np

from ._abi import load_library

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3

_hip = None


def runtime():
    global _hip
    if _hip is None:
        load_library()  # resolves libamdhip64.so.7 for the process
        h = C.CDLL("libamdhip64.so.7")
        P = C.POINTER
        for name, res, args in [
            ("hipGetDeviceCount", C.c_int, [P(C.c_int)]),
            ("hipSetDevice", C.c_int, [C.c_int]),
            ("hipMalloc", C.c_int, [P(C.c_void_p), C.c_size_t]),
            ("hipFree", C.c_int, [C.c_void_p]),
            ("hipMemcpy", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
            ("hipMemcpyAsync", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]),
            ("hipMemset", C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
            ("hipDeviceSynchronize", C.c_int, []),
            ("hipStreamCreate", C.c_int, [P(C.c_void_p)]),
            ("hipStreamDestroy", C.c_int, [C.c_void_p]),
            ("hipStreamSynchronize", C.c_int, [C.c_void_p]),
            ("hipEventCreate", C.c_int, [P(C.c_void_p)]),
            ("hipEventDestroy", C.c_int, [C.c_void_p]),
            ("hipEventRecord", C.c_int, [C.c_void_p, C.c_void_p]),
            ("hipEventSynchronize", C.c_int, [C.c_void_p]),
            ("hipEventElapsedTime", C.c_int, [P(C.c_float), C.c_void_p, C.c_void_p]),
            ("hipGetErrorString", C.c_char_p, [C.c_int]),
        ]:
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _hip = h
    return _hip


def check(rc):
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}: {runtime().hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    if runtime().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def synchronize():
    check(runtime().hipDeviceSynchronize())


class DeviceBuffer:
    """hipMalloc'd bytes with numpy upload/download."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 16)))
        self._p = p

    @property
    def ptr(self):
        return self._p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(runtime().hipMemcpy(b._p, a.ctypes.data_as(C.c_void_p), a.nbytes, hipMemcpyHostToDevice))
        return b

    def copy_from_numpy(self, a):
        """Uploads a into the start of the buffer."""
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(runtime().hipMemcpy(self._p, a.ctypes.data_as(C.c_void_p), a.nbytes, hipMemcpyHostToDevice))

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._p, out.nbytes, hipMemcpyDeviceToHost))
        return out

    def zero(self):
        check(runtime().hipMemset(self._p, 0, self.nbytes))

    def free(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            runtime().hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(runtime().hipStreamCreate(C.byref(s)))
        self._s = s

    @property
    def handle(self):
        return self._s.value

    def synchronize(self):
        check(runtime().hipStreamSynchronize(self._s))

    def __del__(self):
        try:
            if self._s.value:
                runtime().hipStreamDestroy(self._s)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(runtime().hipEventCreate(C.byref(e)))
        self._e = e

    def record(self, stream=None):
        check(runtime().hipEventRecord(self._e, C.c_void_p(stream.handle if isinstance(stream, Stream) else stream)))

    def synchronize(self):
        check(runtime().hipEventSynchronize(self._e))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(runtime().hipEventElapsedTime(C.byref(ms), self._e, end._e))
        return ms.value

    def __del__(self):
        try:
            if self._e.value:
                runtime().hipEventDestroy(self._e)
        except Exception:
            pass
