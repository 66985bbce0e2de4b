"""Minimal HIP runtime plumbing over ctypes: device buffers, streams, events.

Binds the same libamdhip64.so.7 that libpbfmm.so links (so both see one HIP
runtime in the process); used by tests and benchmarks to hand resident device
vectors to the *_dev entry points and to time work on the library's stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import load as _load_pbfmm

_hip = None


def hip():
    global _hip
    if _hip is None:
        _load_pbfmm()  # make libpbfmm pull in its libamdhip64 first
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipSetDevice.argtypes = [C.c_int]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventDestroy.argtypes = [C.c_void_p]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipGetErrorString.restype = C.c_char_p
        L.hipGetErrorString.argtypes = [C.c_int]
        _hip = L
    return _hip


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


H2D, D2H, D2D = 1, 2, 3


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A float64 vector in device memory."""

    def __init__(self, n: int):
        self.n = int(n)
        self.ptr = C.c_void_p()
        _chk(hip().hipMalloc(C.byref(self.ptr), max(8 * self.n, 8)), "hipMalloc")
        _chk(hip().hipMemset(self.ptr, 0, max(8 * self.n, 8)), "hipMemset")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        _chk(hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), 8 * self.n, H2D), "hipMemcpy H2D")

    def copy_from(self, other: "DeviceArray"):
        assert other.n == self.n
        _chk(hip().hipMemcpy(self.ptr, other.ptr, 8 * self.n, D2D), "hipMemcpy D2D")

    def numpy(self):
        out = np.empty(self.n)
        _chk(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n, D2H), "hipMemcpy D2H")
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        _chk(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def synchronize(self):
        _chk(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h:
            hip().hipStreamDestroy(self.h)
            self.h = C.c_void_p()


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _chk(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream: Stream):
        _chk(hip().hipEventRecord(self.h, stream.h), "hipEventRecord")

    def synchronize(self):
        _chk(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        _chk(hip().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return float(ms.value)
