"""HBM-resident buffers, streams and events over the C ABI.

Used by callers that keep operands on the device (bench.py, the GPU tests,
the serving loop). Nothing here computes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .level3 import raise_status


class Stream:
    def __init__(self):
        h = C.c_void_p()
        raise_status(_ffi.lib().rblas_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self) -> None:
        raise_status(_ffi.lib().rblas_stream_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            raise_status(_ffi.lib().rblas_stream_destroy(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        h = C.c_void_p()
        raise_status(_ffi.lib().rblas_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream: Stream | None = None) -> None:
        raise_status(_ffi.lib().rblas_event_record(self.handle,
                                                   stream.handle if stream else None))

    def elapsed_ms(self, start: "Event") -> float:
        ms = C.c_float()
        raise_status(_ffi.lib().rblas_event_elapsed_ms(C.byref(ms), start.handle, self.handle))
        return float(ms.value)

    def close(self) -> None:
        if self.handle:
            _ffi.lib().rblas_event_destroy(self.handle)
            self.handle = None


class DeviceArray:
    """A 1-D device buffer of `size` elements of `dtype`."""

    def __init__(self, size: int, dtype):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        p = C.c_void_p()
        raise_status(_ffi.lib().rblas_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @classmethod
    def from_host(cls, x: np.ndarray, stream: Stream | None = None) -> "DeviceArray":
        x = np.ascontiguousarray(x)
        d = cls(x.size, x.dtype)
        d.upload(x, stream)
        return d

    def upload(self, x: np.ndarray, stream: Stream | None = None) -> None:
        x = np.ascontiguousarray(x, dtype=self.dtype)
        assert x.size <= self.size
        raise_status(_ffi.lib().rblas_memcpy_htod(self.ptr, x.ctypes.data, x.nbytes,
                                                  stream.handle if stream else None))

    def download(self, stream: Stream | None = None) -> np.ndarray:
        out = np.empty(self.size, self.dtype)
        raise_status(_ffi.lib().rblas_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes,
                                                  stream.handle if stream else None))
        return out

    def free(self) -> None:
        if self.ptr:
            raise_status(_ffi.lib().rblas_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    return int(_ffi.lib().rblas_device_count())


def synchronize() -> None:
    raise_status(_ffi.lib().rblas_device_synchronize())
