"""Minimal ctypes access to the HIP runtime that libhalf.so itself is bound
to: device buffers, streams and events for the device-level entry points,
without a second HIP runtime in the process.

libhalf.so links libamdhip64.so.7 (/opt/rocm/lib).  When another library
has already put a HIP runtime with that SONAME into the process (torch's
bundled copy, if torch was imported first), the dynamic linker binds
libhalf.so to that one, so the runtime used here is looked up among the
loaded objects rather than opened by path."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def _loaded_runtime():
    """Path of the HIP runtime mapped into this process (the /opt/rocm one
    when two are mapped: libhalf.so, loaded first, is bound to it)."""
    try:
        with open("/proc/self/maps") as fh:
            paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln and "/" in ln}
    except OSError:
        return None
    paths = sorted(p for p in paths if os.path.basename(p).startswith("libamdhip64"))
    rocm = [p for p in paths if p.startswith("/opt/rocm")]
    return (rocm or paths or [None])[0]


def lib():
    global _hip
    if _hip is None:
        from . import load
        load()  # libhalf first: then the runtime it is bound to is loaded
        path = _loaded_runtime() or "/opt/rocm/lib/libamdhip64.so.7"
        _hip = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return _hip


def _ok(st, what):
    if st != 0:
        raise RuntimeError(f"{what} failed with hipError {st}")


def device_count() -> int:
    n = C.c_int(0)
    return n.value if lib().hipGetDeviceCount(C.byref(n)) == 0 else 0


def synchronize():
    _ok(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    """(free, total) device memory in bytes (hipMemGetInfo)."""
    free, total = C.c_size_t(0), C.c_size_t(0)
    _ok(lib().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


def memset(buf: "Buffer", value: int):
    _ok(lib().hipMemset(buf.ptr, int(value), buf.nbytes), "hipMemset")


def copy_async(dst: "Buffer", src: "Buffer", nbytes: int, stream: "Stream"):
    """Device-to-device copy on a stream."""
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy larger than the buffers")
    _ok(lib().hipMemcpyAsync(dst.ptr, src.ptr, nbytes, D2D, stream.ptr), "hipMemcpyAsync")


class Buffer:
    """Device allocation holding a numpy-typed array."""

    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(lib().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype)
        _ok(lib().hipMemcpy(b.ptr, a.ctypes.data, b.nbytes, H2D), "hipMemcpy")
        return b

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _ok(lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(lib().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.ptr = s.value

    def synchronize(self):
        _ok(lib().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ok(lib().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.ptr = e.value

    def record(self, stream: Stream | None = None):
        _ok(lib().hipEventRecord(self.ptr, stream.ptr if stream else None), "hipEventRecord")

    def elapsed_ms(self, later: "Event") -> float:
        _ok(lib().hipEventSynchronize(later.ptr), "hipEventSynchronize")
        ms = C.c_float()
        _ok(lib().hipEventElapsedTime(C.byref(ms), self.ptr, later.ptr), "hipEventElapsedTime")
        return ms.value
