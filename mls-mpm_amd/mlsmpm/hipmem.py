"""Device buffers from the HIP runtime that libmlsmpm.so is bound to (ctypes).

For callers that hand device pointers to mpm_export_device without a
framework of their own.  The runtime is looked up by its soname after the
library has loaded, so the buffers and the library's kernels share one
runtime instance.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import load

_hip = None
_D2H, _H2D, _D2D = 2, 1, 3


def runtime() -> C.CDLL:
    global _hip
    if _hip is None:
        load()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipGetErrorString.restype = C.c_char_p
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        _hip = h
    return _hip


def _ok(e: int, what: str) -> None:
    if e != 0:
        msg = runtime().hipGetErrorString(e).decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}")


class DeviceBuffer:
    """nbytes of device memory (hipMalloc); .ptr is the device address."""

    def __init__(self, nbytes: int, fill: int | None = None):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _ok(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if fill is not None:
            _ok(runtime().hipMemset(self.ptr, fill, self.nbytes), "hipMemset")
            # finished before any stream of the library touches the buffer
            _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def to_numpy(self, dtype, shape=None) -> np.ndarray:
        a = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        _ok(runtime().hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, _D2H), "hipMemcpy")
        return a if shape is None else a.reshape(shape)

    def copy_from(self, src: "DeviceBuffer") -> None:
        """Device-to-device copy of src's bytes (synchronous)."""
        n = min(self.nbytes, src.nbytes)
        _ok(runtime().hipMemcpy(self.ptr, src.ptr, n, _D2D), "hipMemcpy")

    def close(self) -> None:
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StreamTimer:
    """Milliseconds between two HIP events recorded on a given stream (the
    scene's: Scene.stream())."""

    def __init__(self, stream: int):
        self.stream = stream
        self.ev = []
        for _ in range(2):
            e = C.c_void_p()
            _ok(runtime().hipEventCreate(C.byref(e)), "hipEventCreate")
            self.ev.append(e.value)

    def start(self) -> None:
        _ok(runtime().hipEventRecord(self.ev[0], self.stream), "hipEventRecord")

    def stop(self) -> float:
        h = runtime()
        _ok(h.hipEventRecord(self.ev[1], self.stream), "hipEventRecord")
        _ok(h.hipEventSynchronize(self.ev[1]), "hipEventSynchronize")
        ms = C.c_float()
        _ok(h.hipEventElapsedTime(C.byref(ms), self.ev[0], self.ev[1]), "hipEventElapsedTime")
        return ms.value

    def close(self) -> None:
        for e in self.ev:
            runtime().hipEventDestroy(e)
        self.ev = []
