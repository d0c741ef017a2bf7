"""Minimal ctypes access to the HIP runtime that libgbr.so itself links
(libamdhip64.so.7 from ROCm), for device buffers and stream-ordered event
timing without a second runtime in the process.

PyTorch wheels ship their own copy of the HIP runtime under a different
DT_NEEDED name; loading both in one process gives two runtimes that do not
see each other's device state.  Callers that want torch tensors should import
torch *before* this package so libgbr.so binds to the runtime torch already
loaded (same soname)."""
from __future__ import annotations

import ctypes

from . import _native

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _native.lib()  # make sure the runtime libgbr.so uses is the one loaded
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                          ctypes.c_void_p]
        h.hipEventDestroy.argtypes = [ctypes.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.argtypes = [ctypes.c_int]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _check(st: int, what: str) -> None:
    if st != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(st).decode()}")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        _check(hip().hipMalloc(ctypes.byref(self.ptr), nbytes), "hipMalloc")

    @property
    def value(self) -> int:
        return self.ptr.value

    def to_host(self, out) -> None:
        _check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")

    def close(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def memcpy_d2h(out, src: int) -> None:
    """Copies out.nbytes from device address src into the host array out."""
    _check(hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(src), out.nbytes, D2H),
           "hipMemcpy D2H")


def memcpy_h2d(dst: int, src) -> None:
    """Copies the host array src (src.nbytes) to device address dst."""
    _check(hip().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src.ctypes.data), src.nbytes, H2D),
           "hipMemcpy H2D")


def to_device(a) -> DeviceBuffer:
    """A device copy of the contiguous host array a."""
    b = DeviceBuffer(max(a.nbytes, 1))
    if a.nbytes:
        memcpy_h2d(b.value, a)
    return b


def memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    _check(hip().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, D2D), "hipMemcpy D2D")


class Event:
    """hipEvent recorded on an arbitrary hipStream_t (e.g. the renderer's)."""

    def __init__(self):
        self.e = ctypes.c_void_p()
        _check(hip().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream: int) -> None:
        _check(hip().hipEventRecord(self.e, ctypes.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        _check(hip().hipEventSynchronize(end.e), "hipEventSynchronize")
        ms = ctypes.c_float()
        _check(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.e:
                hip().hipEventDestroy(self.e)
        except Exception:
            pass
