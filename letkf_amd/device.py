"""Device memory, streams and events through liblekf_hip.so.

A small host-side helper so numpy callers (tests, the bench, the server) can
hand device pointers to the batched entry points without a second GPU
runtime in the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class DeviceArray:
    """A device buffer with a numpy dtype and shape (C order)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _lib.lib().lekf_dev_malloc(max(self.nbytes, 8))
        if not self.ptr:
            raise MemoryError(f"lekf_dev_malloc({self.nbytes}) failed")

    @classmethod
    def from_numpy(cls, a, dtype=None, stream=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _lib.check(_lib.lib().lekf_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0, stream), "memcpy")
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64, stream=None):
        d = cls(shape, dtype)
        _lib.check(_lib.lib().lekf_memset(d.ptr, 0, max(d.nbytes, 8), stream), "memset")
        _lib.check(_lib.lib().lekf_stream_sync(stream), "sync")
        return d

    def numpy(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _lib.check(_lib.lib().lekf_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1, stream),
                       "memcpy")
        return out

    def data_ptr(self):
        return self.ptr

    def memset(self, value, stream=None):
        _lib.check(_lib.lib().lekf_memset(self.ptr, int(value), max(self.nbytes, 8), stream), "memset")

    def copy_from_device(self, other, stream=None):
        """Asynchronous device-to-device copy of other's bytes (a streaming copy kernel
        when sizes and addresses are 16-byte aligned)."""
        n = min(self.nbytes, other.nbytes)
        _lib.check(_lib.lib().lekf_memcpy(self.ptr, other.ptr, n, 2, stream), "memcpy")

    def free(self):
        if self.ptr:
            _lib.lib().lekf_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.ptr = _lib.lib().lekf_stream_create()
        if not self.ptr:
            raise RuntimeError("lekf_stream_create failed")

    def sync(self):
        _lib.check(_lib.lib().lekf_stream_sync(self.ptr), "stream_sync")

    def __del__(self):
        try:
            if self.ptr:
                _lib.lib().lekf_stream_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class Event:
    """HIP event; ``Event.elapsed_ms(a, b)`` waits for b."""

    def __init__(self):
        self.ptr = _lib.lib().lekf_event_create()
        if not self.ptr:
            raise RuntimeError("lekf_event_create failed")

    def record(self, stream=None):
        _lib.check(_lib.lib().lekf_event_record(self.ptr, stream), "event_record")

    @staticmethod
    def elapsed_ms(a, b):
        ms = ctypes.c_float(0.0)
        _lib.check(_lib.lib().lekf_event_elapsed_ms(a.ptr, b.ptr, ctypes.byref(ms)), "elapsed")
        return float(ms.value)

    def query(self):
        """True once all work recorded before the event has completed."""
        r = _lib.lib().lekf_event_query(self.ptr)
        if r < 0:
            raise _lib.LetkfError(-r, "event_query")
        return r == 1

    def sync(self):
        _lib.check(_lib.lib().lekf_event_sync(self.ptr), "event_sync")

    def __del__(self):
        try:
            if self.ptr:
                _lib.lib().lekf_event_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


def sync(stream=None):
    _lib.check(_lib.lib().lekf_stream_sync(stream), "stream_sync")


def mem_info():
    """(free, total) device memory in bytes."""
    f = ctypes.c_size_t(0)
    t = ctypes.c_size_t(0)
    _lib.check(_lib.lib().lekf_mem_info(ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return int(f.value), int(t.value)


KTIME_NAMES = ("select", "list_cap", "cull", "solve", "gram", "eig", "eig_tri", "eig_ql",
               "eig_apply", "big", "redo_gram", "redo_tri", "redo_ql", "redo_apply")


def ktime_enable(on=True):
    """Start (clearing the record) or stop recording a HIP event pair around each
    launch of the library's hot kernels, on the launch's stream."""
    _lib.check(_lib.lib().lekf_ktime_enable(int(bool(on))), "ktime_enable")


def ktime_read(name):
    """(total ms, launches) of the recorded launches of one kernel name."""
    ms = ctypes.c_double(0.0)
    n = ctypes.c_long(0)
    _lib.check(_lib.lib().lekf_ktime_read(name.encode(), ctypes.byref(ms), ctypes.byref(n)),
               "ktime_read")
    return float(ms.value), int(n.value)


def ktime_all():
    """{name: (total ms, launches)} of every name recorded."""
    out = {}
    for n in KTIME_NAMES:
        t, c = ktime_read(n)
        if c:
            out[n] = (t, c)
    return out
