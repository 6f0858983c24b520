"""Minimal ctypes binding to the HIP runtime: device buffers, copies, streams
and events for the device-resident entry points of libthirring_hip.so.

It loads the same libamdhip64 that libthirring_hip.so links against, so one
HIP runtime serves the whole process.
"""
import ctypes
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import thirring_hip
        thirring_hip.lib()  # the library's own runtime dependency loads first
        for cand in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64 not found")
        V = ctypes.c_void_p
        S = ctypes.c_size_t
        sig = {
            "hipMalloc": ([ctypes.POINTER(V), S], ctypes.c_int),
            "hipFree": ([V], ctypes.c_int),
            "hipMemcpy": ([V, V, S, ctypes.c_int], ctypes.c_int),
            "hipMemset": ([V, ctypes.c_int, S], ctypes.c_int),
            "hipDeviceSynchronize": ([], ctypes.c_int),
            "hipStreamCreateWithFlags": ([ctypes.POINTER(V), ctypes.c_uint], ctypes.c_int),
            "hipStreamSynchronize": ([V], ctypes.c_int),
            "hipStreamDestroy": ([V], ctypes.c_int),
            "hipEventCreate": ([ctypes.POINTER(V)], ctypes.c_int),
            "hipEventRecord": ([V, V], ctypes.c_int),
            "hipEventSynchronize": ([V], ctypes.c_int),
            "hipEventElapsedTime": ([ctypes.POINTER(ctypes.c_float), V, V], ctypes.c_int),
            "hipEventDestroy": ([V], ctypes.c_int),
            "hipGetErrorString": ([ctypes.c_int], ctypes.c_char_p),
        }
        for k, (a, r) in sig.items():
            f = getattr(_hip, k)
            f.argtypes = a
            f.restype = r
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """A device allocation with a numpy dtype/shape view for copies."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    def data_ptr(self):
        return self.ptr

    @classmethod
    def from_host(cls, a):
        """Device copy of a host array, complete on return.  (A pageable hipMemcpy
        may return once the data is staged, before the DMA has landed, and
        work on non-blocking streams is not ordered after it.)"""
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(hip().hipMemcpy(ctypes.c_void_p(d.ptr), a.ctypes.data_as(ctypes.c_void_p),
                              d.nbytes, H2D), "hipMemcpy H2D")
        check(hip().hipStreamSynchronize(None), "hipStreamSynchronize")
        return d

    def to_host(self):
        a = np.empty(self.shape, dtype=self.dtype)
        check(hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                              self.nbytes, D2H), "hipMemcpy D2H")
        return a

    def head_to_host(self, k):
        """The first k entries along axis 0, copied to the host."""
        k = int(k)
        assert 0 <= k <= self.shape[0]
        a = np.empty((k,) + self.shape[1:], dtype=self.dtype)
        check(hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                              a.nbytes, D2H), "hipMemcpy D2H")
        return a

    def copy_from(self, other):
        """Device-to-device copy, complete on return."""
        assert other.nbytes == self.nbytes
        check(hip().hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), self.nbytes,
                              D2D), "hipMemcpy D2D")
        check(hip().hipStreamSynchronize(None), "hipStreamSynchronize")

    def zero(self):
        """Fill with zero bytes, complete on return."""
        check(hip().hipMemset(ctypes.c_void_p(self.ptr), 0, self.nbytes), "hipMemset")
        check(hip().hipStreamSynchronize(None), "hipStreamSynchronize")

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A non-blocking stream: no implicit ordering with the legacy null stream,
    so synchronous copies issued by other threads never wait on (or break the
    graph capture of) work queued here."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(hip().hipStreamCreateWithFlags(ctypes.byref(p), 1), "hipStreamCreateWithFlags")
        self.ptr = p.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(ctypes.c_void_p(self.ptr)), "hipStreamSynchronize")

    def destroy(self):
        if self.ptr:
            check(hip().hipStreamDestroy(ctypes.c_void_p(self.ptr)), "hipStreamDestroy")
            self.ptr = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(hip().hipEventCreate(ctypes.byref(p)), "hipEventCreate")
        self.ptr = p.value

    def record(self, stream=None):
        s = ctypes.c_void_p(stream.ptr if stream is not None else None)
        check(hip().hipEventRecord(ctypes.c_void_p(self.ptr), s), "hipEventRecord")

    def elapsed_ms(self, end):
        check(hip().hipEventSynchronize(ctypes.c_void_p(end.ptr)), "hipEventSynchronize")
        ms = ctypes.c_float()
        check(hip().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.ptr),
                                        ctypes.c_void_p(end.ptr)), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            hip().hipEventDestroy(ctypes.c_void_p(self.ptr))
        except Exception:
            pass
