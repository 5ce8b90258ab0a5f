"""Minimal ctypes binding of the HIP runtime (libamdhip64): device buffers,
copies, streams and events for the host code, tests and benchmarks.

This is plumbing only; all computation happens in liblowdin_fock.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HIP = None
H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def hip():
    global _HIP
    if _HIP is None:
        for cand in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
                     "libamdhip64.so"):
            try:
                _HIP = C.CDLL(cand)
                break
            except OSError:
                continue
        if _HIP is None:
            raise HipError("libamdhip64.so not found")
        vp = C.c_void_p
        _HIP.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        _HIP.hipFree.argtypes = [vp]
        _HIP.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        _HIP.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        _HIP.hipMemsetAsync.argtypes = [vp, C.c_int, C.c_size_t, vp]
        _HIP.hipStreamCreate.argtypes = [C.POINTER(vp)]
        _HIP.hipStreamDestroy.argtypes = [vp]
        _HIP.hipStreamSynchronize.argtypes = [vp]
        _HIP.hipDeviceSynchronize.argtypes = []
        _HIP.hipEventCreate.argtypes = [C.POINTER(vp)]
        _HIP.hipEventDestroy.argtypes = [vp]
        _HIP.hipEventRecord.argtypes = [vp, vp]
        _HIP.hipEventSynchronize.argtypes = [vp]
        _HIP.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        _HIP.hipGetErrorString.argtypes = [C.c_int]
        _HIP.hipGetErrorString.restype = C.c_char_p
        _HIP.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _HIP.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _HIP


def check(rc, what=""):
    if rc != 0:
        raise HipError("%s: %s" % (what, hip().hipGetErrorString(rc).decode()))


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def mem_info():
    free, total = C.c_size_t(0), C.c_size_t(0)
    check(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """A device buffer holding an FP64 (or other dtype) array."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if stream is None:
            check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")
        else:
            check(hip().hipMemcpyAsync(self.ptr, a.ctypes.data, self.nbytes, H2D, stream.handle), "H2D")
            stream.synchronize()

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of min(nbytes) bytes."""
        check(hip().hipMemcpy(self.ptr, other.ptr, min(self.nbytes, other.nbytes), D2D), "hipMemcpy D2D")

    def download(self, order="C") -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")  # default priority
        self.handle = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if getattr(self, "handle", None):
            hip().hipStreamSynchronize(self.handle)
            hip().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream):
        check(hip().hipEventRecord(self.handle, stream.handle), "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        check(hip().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value
