"""Minimal ctypes access to the HIP runtime for device buffers and stream events
(used by the benchmark to keep inputs resident in HBM and to time kernels on
the context's own stream)."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        P = ctypes.c_void_p
        sig = {
            "hipMalloc": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_size_t]),
            "hipFree": (ctypes.c_int, [P]),
            "hipMemcpy": (ctypes.c_int, [P, P, ctypes.c_size_t, ctypes.c_int]),
            "hipEventCreate": (ctypes.c_int, [ctypes.POINTER(P)]),
            "hipEventDestroy": (ctypes.c_int, [P]),
            "hipEventRecord": (ctypes.c_int, [P, P]),
            "hipEventSynchronize": (ctypes.c_int, [P]),
            "hipEventElapsedTime": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), P, P]),
            "hipDeviceSynchronize": (ctypes.c_int, []),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """A device allocation holding a copy of a numpy array."""

    def __init__(self, host: np.ndarray):
        self.host_dtype = host.dtype
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 8)), "hipMalloc")
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.host_dtype)
        _ok(hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape, self.host_dtype)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream):
        _ok(hip().hipEventRecord(self.e, ctypes.c_void_p(stream)), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return float(ms.value)
