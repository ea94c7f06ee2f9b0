"""Level-1 BLAS on the GPU with rblas's signatures (level1.rs).

x and y are 1-D numpy arrays (float32 or float64); `incx`/`incy` are rblas's
positive increments. axpy updates y in place. Raises types.BlasPanic where
rblas panics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .level3 import _dev_prefix, _prefix, _scalar, raise_status


def dot(n, x, incx, y, incy):
    """x^T y (level1.rs:45-66)."""
    pre = _prefix(x, y)
    out = C.c_float() if pre == "s" else C.c_double()
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}dot")(n, x.ctypes.data, x.size, incx,
                                                        y.ctypes.data, y.size, incy, C.byref(out)))
    return out.value


def nrm2(n, x, incx):
    """Euclidean norm (level1.rs:96-123), overflow- and underflow-safe."""
    pre = _prefix(x)
    out = C.c_float() if pre == "s" else C.c_double()
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}nrm2")(n, x.ctypes.data, x.size, incx,
                                                         C.byref(out)))
    return out.value


def axpy(n, alpha, x, incx, y, incy) -> None:
    """y := alpha*x + y (level1.rs:17-41)."""
    pre = _prefix(x, y)
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}axpy")(n, _scalar(pre, alpha), x.ctypes.data,
                                                         x.size, incx, y.ctypes.data, y.size,
                                                         incy))


def dot_dev(n, x, incx, y, incy, result, stream=None) -> None:
    """dot on device arrays; the result goes to device array `result[0]`."""
    pre = _dev_prefix(x, y, result)
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}dot_dev")(
        n, x.ptr, x.size, incx, y.ptr, y.size, incy, result.ptr,
        stream.handle if stream is not None else None))


def nrm2_dev(n, x, incx, result, stream=None) -> None:
    pre = _dev_prefix(x, result)
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}nrm2_dev")(
        n, x.ptr, x.size, incx, result.ptr, stream.handle if stream is not None else None))


def axpy_dev(n, alpha, x, incx, y, incy, stream=None) -> None:
    pre = _dev_prefix(x, y)
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}axpy_dev")(
        n, _scalar(pre, alpha), x.ptr, x.size, incx, y.ptr, y.size, incy,
        stream.handle if stream is not None else None))
