"""Complex Level-3 on the GPU with rblas's signature (complex3.rs).

a, b, c are 1-D numpy complex64 / complex128 arrays (interleaved {re, im},
rblas's Complex<T> layout), column-major; c is updated in place.
"""
from __future__ import annotations

import numpy as np

from . import _ffi
from .level3 import raise_status


def _kind(*arrs):
    dt = arrs[-1].dtype
    for x in arrs:
        if x.dtype != dt:
            raise TypeError("all operands must share one dtype")
        if x.ndim != 1 or not x.flags.c_contiguous:
            raise TypeError("operands are 1-D contiguous arrays (rblas slices)")
    if dt == np.complex64:
        return "c", _ffi.Complex32
    if dt == np.complex128:
        return "z", _ffi.Complex64
    raise TypeError(f"unsupported dtype {dt}")


def gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """C := alpha*op(A)*op(B) + beta*C by 4M (complex3.rs:27-74)."""
    pre, cs = _kind(a, b, c)
    al, be = complex(alpha), complex(beta)
    st = getattr(_ffi.lib(), f"rblas_{pre}gemm")(
        int(transa), int(transb), m, n, k, cs(al.real, al.imag), a.ctypes.data, a.size, lda,
        b.ctypes.data, b.size, ldb, cs(be.real, be.imag), c.ctypes.data, c.size, ldc)
    raise_status(st)


def gemm_dev(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None) -> None:
    """gemm on HBM-resident complex operands (device.DeviceArray)."""
    pre = {np.dtype(np.complex64): "c", np.dtype(np.complex128): "z"}[c.dtype]
    cs = _ffi.Complex32 if pre == "c" else _ffi.Complex64
    al, be = complex(alpha), complex(beta)
    st = getattr(_ffi.lib(), f"rblas_{pre}gemm_dev")(
        int(transa), int(transb), m, n, k, cs(al.real, al.imag), a.ptr, a.size, lda, b.ptr,
        b.size, ldb, cs(be.real, be.imag), c.ptr, c.size, ldc,
        stream.handle if stream is not None else None)
    raise_status(st)


def _hesy(name, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    pre, cs = _kind(a, b, c)
    al, be = complex(alpha), complex(beta)
    st = getattr(_ffi.lib(), f"rblas_{pre}{name}")(
        int(side), int(uplo), m, n, cs(al.real, al.imag), a.ctypes.data, a.size, lda,
        b.ctypes.data, b.size, ldb, cs(be.real, be.imag), c.ctypes.data, c.size, ldc)
    raise_status(st)


def _hesy_dev(name, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream) -> None:
    pre = {np.dtype(np.complex64): "c", np.dtype(np.complex128): "z"}[c.dtype]
    cs = _ffi.Complex32 if pre == "c" else _ffi.Complex64
    al, be = complex(alpha), complex(beta)
    st = getattr(_ffi.lib(), f"rblas_{pre}{name}_dev")(
        int(side), int(uplo), m, n, cs(al.real, al.imag), a.ptr, a.size, lda, b.ptr, b.size, ldb,
        cs(be.real, be.imag), c.ptr, c.size, ldc, stream.handle if stream is not None else None)
    raise_status(st)


def symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """Complex symmetric multiply, A = A^T with its `uplo` triangle stored
    (complex3.rs:127-146)."""
    _hesy("symm", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc)


def hemm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """Complex Hermitian multiply, A = A^H with its `uplo` triangle stored;
    the diagonal's imaginary part is taken as zero (complex3.rs:148-168)."""
    _hesy("hemm", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc)


def symm_dev(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream=None) -> None:
    """symm on HBM-resident complex operands (device.DeviceArray)."""
    _hesy_dev("symm", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream)


def hemm_dev(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream=None) -> None:
    """hemm on HBM-resident complex operands (device.DeviceArray)."""
    _hesy_dev("hemm", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream)


def trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb) -> None:
    """Solve op(A) X = alpha B or X op(A) = alpha B in place in b, op with
    conjugation for ConjTrans (complex3.rs:252-349)."""
    pre, cs = _kind(a, b)
    al = complex(alpha)
    st = getattr(_ffi.lib(), f"rblas_{pre}trsm")(
        int(side), int(uplo), int(trans), int(diag), m, n, cs(al.real, al.imag), a.ctypes.data,
        a.size, lda, b.ctypes.data, b.size, ldb)
    raise_status(st)


def trsm_dev(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb, stream=None) -> None:
    """Complex trsm on HBM-resident operands (device.DeviceArray)."""
    pre = {np.dtype(np.complex64): "c", np.dtype(np.complex128): "z"}[b.dtype]
    cs = _ffi.Complex32 if pre == "c" else _ffi.Complex64
    al = complex(alpha)
    st = getattr(_ffi.lib(), f"rblas_{pre}trsm_dev")(
        int(side), int(uplo), int(trans), int(diag), m, n, cs(al.real, al.imag), a.ptr, a.size,
        lda, b.ptr, b.size, ldb, stream.handle if stream is not None else None)
    raise_status(st)
