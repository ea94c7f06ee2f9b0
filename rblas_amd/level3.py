"""Level-3 BLAS on the GPU, with rblas's signatures (level3.rs).

Operands are 1-D numpy arrays ("slices"), float32 or float64, exactly as
rblas takes `&[T]` / `&mut [T]`; `c` (or `b` for trsm) is updated in place.
Every call goes through librblas_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .types import (BAD_ARG, BAD_ENUM, BUFFER_TOO_SMALL, INVALID_LEADING_DIM, OK, PANIC,
                    ZERO_STRIDE, BlasError, BlasPanic, DeviceError)

_KIND = {INVALID_LEADING_DIM: "InvalidLeadingDim", BUFFER_TOO_SMALL: "BufferTooSmall",
         ZERO_STRIDE: "ZeroStride"}


def raise_status(status: int, err: _ffi.RblasError | None = None) -> None:
    """Turn a C-ABI status into the Python exception rblas's API implies."""
    if status == OK:
        return
    if status in _KIND:
        which = err.which.decode() if err is not None else "?"
        raise BlasError(_KIND[status], which, int(err.got) if err else 0,
                        int(err.bound) if err else 0)
    msg = _ffi.last_error()
    if status == PANIC:
        raise BlasPanic(msg)
    if status in (BAD_ENUM, BAD_ARG):
        raise ValueError(msg)
    raise DeviceError(f"status {status}: {msg}")


def _prefix(*arrs: np.ndarray) -> str:
    dt = arrs[-1].dtype
    for x in arrs:
        if x.dtype != dt:
            raise TypeError("all operands must share one dtype")
        if x.ndim != 1 or not x.flags.c_contiguous:
            raise TypeError("operands are 1-D contiguous arrays (rblas slices)")
    if dt == np.float32:
        return "s"
    if dt == np.float64:
        return "d"
    raise TypeError(f"unsupported dtype {dt}")


def _scalar(pre: str, v):
    return C.c_float(v) if pre == "s" else C.c_double(v)


def gemm(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """C := alpha*op(A)*op(B) + beta*C (level3.rs:32-62). Raises BlasPanic
    where rblas's gemm panics."""
    pre = _prefix(a, b, c)
    fn = getattr(_ffi.lib(), f"rblas_{pre}gemm")
    st = fn(int(layout), int(transa), int(transb), m, n, k, _scalar(pre, alpha),
            a.ctypes.data, a.size, lda, b.ctypes.data, b.size, ldb, _scalar(pre, beta),
            c.ctypes.data, c.size, ldc)
    raise_status(st)


def try_gemm(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """Checked gemm (level3.rs:68-103): raises BlasError for the first bad
    operand in the order a, b, c."""
    pre = _prefix(a, b, c)
    fn = getattr(_ffi.lib(), f"rblas_{pre}gemm_try")
    err = _ffi.RblasError()
    st = fn(int(layout), int(transa), int(transb), m, n, k, _scalar(pre, alpha),
            a.ctypes.data, a.size, lda, b.ctypes.data, b.size, ldb, _scalar(pre, beta),
            c.ctypes.data, c.size, ldc, C.byref(err))
    raise_status(st, err)


def _dev_prefix(*arrs) -> str:
    dt = arrs[-1].dtype
    if any(x.dtype != dt for x in arrs):
        raise TypeError("all operands must share one dtype")
    if dt == np.float32:
        return "s"
    if dt == np.float64:
        return "d"
    raise TypeError(f"unsupported dtype {dt}")


def gemm_dev(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
             stream=None) -> None:
    """gemm on HBM-resident operands (device.DeviceArray); enqueued on
    `stream`, not synchronised."""
    pre = _dev_prefix(a, b, c)
    fn = getattr(_ffi.lib(), f"rblas_{pre}gemm_dev")
    st = fn(int(layout), int(transa), int(transb), m, n, k, _scalar(pre, alpha),
            a.ptr, a.size, lda, b.ptr, b.size, ldb, _scalar(pre, beta), c.ptr, c.size, ldc,
            stream.handle if stream is not None else None)
    raise_status(st)


def dgemm_dev_noalloc(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                      stream=None) -> None:
    """f64 gemm_dev that allocates no device memory (its k-split workspace
    is a fixed device array shared by every call on the GPU: calls must not
    overlap on the device)."""
    if _dev_prefix(a, b, c) != "d":
        raise TypeError("dgemm_dev_noalloc takes float64 operands")
    st = _ffi.lib().rblas_dgemm_dev_noalloc(
        int(layout), int(transa), int(transb), m, n, k, float(alpha), a.ptr, a.size, lda, b.ptr,
        b.size, ldb, float(beta), c.ptr, c.size, ldc, stream.handle if stream is not None else None)
    raise_status(st)


def sgemm_ex(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
             compute) -> None:
    """f32 gemm with a chosen arithmetic (types.Compute); host slices."""
    if _prefix(a, b, c) != "s":
        raise TypeError("sgemm_ex takes float32 operands")
    st = _ffi.lib().rblas_sgemm_ex(int(layout), int(transa), int(transb), m, n, k,
                                   C.c_float(alpha), a.ctypes.data, a.size, lda, b.ctypes.data,
                                   b.size, ldb, C.c_float(beta), c.ctypes.data, c.size, ldc,
                                   int(compute))
    raise_status(st)


def sgemm_ex_dev(layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                 compute, stream=None) -> None:
    """sgemm_ex on HBM-resident operands, enqueued on `stream`."""
    if _dev_prefix(a, b, c) != "s":
        raise TypeError("sgemm_ex_dev takes float32 operands")
    st = _ffi.lib().rblas_sgemm_ex_dev(int(layout), int(transa), int(transb), m, n, k,
                                       C.c_float(alpha), a.ptr, a.size, lda, b.ptr, b.size, ldb,
                                       C.c_float(beta), c.ptr, c.size, ldc, int(compute),
                                       stream.handle if stream is not None else None)
    raise_status(st)


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc) -> None:
    """C_uplo := alpha*op(A)*op(A)^T + beta*C_uplo (level3.rs:236-294)."""
    pre = _prefix(a, c)
    st = getattr(_ffi.lib(), f"rblas_{pre}syrk")(
        int(uplo), int(trans), n, k, _scalar(pre, alpha), a.ctypes.data, a.size, lda,
        _scalar(pre, beta), c.ctypes.data, c.size, ldc)
    raise_status(st)


def syrk_dev(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, stream=None) -> None:
    """syrk on HBM-resident operands, enqueued on `stream`."""
    pre = _dev_prefix(a, c)
    st = getattr(_ffi.lib(), f"rblas_{pre}syrk_dev")(
        int(uplo), int(trans), n, k, _scalar(pre, alpha), a.ptr, a.size, lda,
        _scalar(pre, beta), c.ptr, c.size, ldc, stream.handle if stream is not None else None)
    raise_status(st)


def symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """C := alpha*A*B + beta*C (Left) or alpha*B*A + beta*C (Right), A
    symmetric with its `uplo` triangle stored (level3.rs:301-381)."""
    pre = _prefix(a, b, c)
    st = getattr(_ffi.lib(), f"rblas_{pre}symm")(
        int(side), int(uplo), m, n, _scalar(pre, alpha), a.ctypes.data, a.size, lda,
        b.ctypes.data, b.size, ldb, _scalar(pre, beta), c.ctypes.data, c.size, ldc)
    raise_status(st)


def symm_dev(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream=None) -> None:
    """symm on HBM-resident operands, enqueued on `stream`."""
    pre = _dev_prefix(a, b, c)
    st = getattr(_ffi.lib(), f"rblas_{pre}symm_dev")(
        int(side), int(uplo), m, n, _scalar(pre, alpha), a.ptr, a.size, lda, b.ptr, b.size, ldb,
        _scalar(pre, beta), c.ptr, c.size, ldc, stream.handle if stream is not None else None)
    raise_status(st)


def trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb) -> None:
    """Solve op(A) X = alpha B or X op(A) = alpha B in place in b
    (level3.rs:390-520)."""
    pre = _prefix(a, b)
    st = getattr(_ffi.lib(), f"rblas_{pre}trsm")(
        int(side), int(uplo), int(trans), int(diag), m, n, _scalar(pre, alpha), a.ctypes.data,
        a.size, lda, b.ctypes.data, b.size, ldb)
    raise_status(st)


def trsm_dev(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb, stream=None) -> None:
    """trsm on HBM-resident operands, enqueued on `stream`."""
    pre = _dev_prefix(a, b)
    st = getattr(_ffi.lib(), f"rblas_{pre}trsm_dev")(
        int(side), int(uplo), int(trans), int(diag), m, n, _scalar(pre, alpha), a.ptr, a.size,
        lda, b.ptr, b.size, ldb, stream.handle if stream is not None else None)
    raise_status(st)
