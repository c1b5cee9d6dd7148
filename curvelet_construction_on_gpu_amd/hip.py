"""Device buffers and events through the HIP runtime that libcurvelet_gpu.so
itself links (ctypes), for callers of the device-pointer entry points
(cvg_run_device, cvg_run_device_noalloc) and for timing on a context's
stream.  The same runtime matters: a second HIP runtime in the process (for
instance the one a PyTorch wheel bundles) does not share device state with it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import CurveletError, load_library

_hip = None


def runtime():
    global _hip
    if _hip is None:
        load_library()  # loads its libamdhip64 first; the soname then resolves to that copy
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                lib = C.CDLL(name)
                break
            except OSError:
                lib = None
        if lib is None:
            raise CurveletError("libamdhip64 not found")
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        lib.hipEventDestroy.argtypes = [C.c_void_p]
        lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        lib.hipEventSynchronize.argtypes = [C.c_void_p]
        lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip = lib
    return _hip


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise CurveletError(f"{what} failed (hipError {rc})", rc)


class DeviceArray:
    """A device allocation of `shape` x `dtype` (hipMalloc), freed with the object."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize if self.shape else self.dtype.itemsize
        p = C.c_void_p()
        _check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes > self.nbytes:
            raise CurveletError("upload larger than the device array")
        _check(runtime().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), a.nbytes, 1), "hipMemcpy H2D")

    def fill_bytes(self, value: int) -> None:
        _check(runtime().hipMemset(C.c_void_p(self.ptr), int(value), self.nbytes), "hipMemset")

    def get(self, rows: int | None = None) -> np.ndarray:
        """Copy to the host (the first `rows` rows of a 2-D array)."""
        shape = self.shape if rows is None else (int(rows),) + self.shape[1:]
        out = np.empty(shape, self.dtype)
        if out.nbytes:
            _check(runtime().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), out.nbytes, 2),
                   "hipMemcpy D2H")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            _hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None


class Event:
    """A HIP event with timing, recorded on a raw hipStream_t (Context.stream)."""

    def __init__(self):
        e = C.c_void_p()
        _check(runtime().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.h = e.value

    def record(self, stream: int) -> None:
        _check(runtime().hipEventRecord(C.c_void_p(self.h), C.c_void_p(stream)), "hipEventRecord")

    def synchronize(self) -> None:
        _check(runtime().hipEventSynchronize(C.c_void_p(self.h)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _check(runtime().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.h), C.c_void_p(end.h)),
               "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.hipEventDestroy(C.c_void_p(self.h))
            self.h = None
