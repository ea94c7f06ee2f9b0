"""Level-2 BLAS on the GPU with rblas's signatures: gemv and ger
(level2.rs:11-104) and the complex gemv / geru / gerc (complex2.rs:11-142).

Operands are 1-D numpy arrays (rblas slices): float32 / float64 for the real
routines, complex64 / complex128 (interleaved {re, im}) for the complex ones;
gemv dispatches on y's dtype. Raises types.BlasPanic where rblas panics.
"""
from __future__ import annotations

import numpy as np

from . import _ffi
from .level3 import _scalar, raise_status

_PRE = {np.dtype(np.float32): "s", np.dtype(np.float64): "d",
        np.dtype(np.complex64): "c", np.dtype(np.complex128): "z"}


def _pre(*arrs) -> str:
    dt = arrs[-1].dtype
    for x in arrs:
        if x.dtype != dt:
            raise TypeError("all operands must share one dtype")
        if isinstance(x, np.ndarray) and (x.ndim != 1 or not x.flags.c_contiguous):
            raise TypeError("operands are 1-D contiguous arrays (rblas slices)")
    if dt not in _PRE:
        raise TypeError(f"unsupported dtype {dt}")
    return _PRE[dt]


def _sc(pre, v):
    if pre in "sd":
        return _scalar(pre, v)
    v = complex(v)
    return (_ffi.Complex32 if pre == "c" else _ffi.Complex64)(v.real, v.imag)


def _ptr(x):
    return x.ctypes.data if isinstance(x, np.ndarray) else x.ptr


def _stream(stream):
    return stream.handle if stream is not None else None


def gemv(layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy) -> None:
    """y := alpha op(A) x + beta y in place (level2.rs:11-61; complex2.rs:11-76
    for complex arrays, ConjTrans conjugating A)."""
    pre = _pre(a, x, y)
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}gemv")(
        int(layout), int(trans), m, n, _sc(pre, alpha), _ptr(a), a.size, lda, _ptr(x), x.size,
        incx, _sc(pre, beta), _ptr(y), y.size, incy))


def gemv_dev(layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream=None) -> None:
    """gemv on HBM-resident operands (device.DeviceArray)."""
    pre = _pre(a, x, y)
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}gemv_dev")(
        int(layout), int(trans), m, n, _sc(pre, alpha), _ptr(a), a.size, lda, _ptr(x), x.size,
        incx, _sc(pre, beta), _ptr(y), y.size, incy, _stream(stream)))


def _rank1(name, layout, m, n, alpha, x, incx, y, incy, a, lda, stream=None, dev=False):
    pre = _pre(x, y, a)
    args = [int(layout), m, n, _sc(pre, alpha), _ptr(x), x.size, incx, _ptr(y), y.size, incy,
            _ptr(a), a.size, lda]
    if dev:
        args.append(_stream(stream))
    raise_status(getattr(_ffi.lib(), f"rblas_{pre}{name}{'_dev' if dev else ''}")(*args))


def ger(layout, m, n, alpha, x, incx, y, incy, a, lda) -> None:
    """A := alpha x y^T + A (level2.rs:65-88), real."""
    _rank1("ger", layout, m, n, alpha, x, incx, y, incy, a, lda)


def geru(layout, m, n, alpha, x, incx, y, incy, a, lda) -> None:
    """A := alpha x y^T + A (complex2.rs:80-93), complex, unconjugated."""
    _rank1("geru", layout, m, n, alpha, x, incx, y, incy, a, lda)


def gerc(layout, m, n, alpha, x, incx, y, incy, a, lda) -> None:
    """A := alpha x y^H + A (complex2.rs:97-110), complex, y conjugated."""
    _rank1("gerc", layout, m, n, alpha, x, incx, y, incy, a, lda)


def ger_dev(layout, m, n, alpha, x, incx, y, incy, a, lda, stream=None) -> None:
    _rank1("ger", layout, m, n, alpha, x, incx, y, incy, a, lda, stream, True)


def geru_dev(layout, m, n, alpha, x, incx, y, incy, a, lda, stream=None) -> None:
    _rank1("geru", layout, m, n, alpha, x, incx, y, incy, a, lda, stream, True)


def gerc_dev(layout, m, n, alpha, x, incx, y, incy, a, lda, stream=None) -> None:
    _rank1("gerc", layout, m, n, alpha, x, incx, y, incy, a, lda, stream, True)
