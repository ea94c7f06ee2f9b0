"""Grouped plans: many rblas gemm (or trsm) problems in one persistent launch.

No rblas counterpart (rblas runs one `level3::gemm`, level3.rs:32-62, or
`level3::trsm`, level3.rs:390-520, per call); every problem keeps that
function's arguments and meaning. Operands are
HBM-resident (device.DeviceArray); a plan captures their addresses and
reruns on whatever they hold.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .level3 import raise_status


class GemmProblem:
    """One problem: the arguments of level3.gemm, with DeviceArray operands."""

    __slots__ = ("layout", "transa", "transb", "m", "n", "k", "alpha", "a", "lda", "b", "ldb",
                 "beta", "c", "ldc")

    def __init__(self, layout, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        self.layout, self.transa, self.transb = int(layout), int(transa), int(transb)
        self.m, self.n, self.k = int(m), int(n), int(k)
        self.alpha, self.beta = alpha, beta
        self.a, self.lda, self.b, self.ldb, self.c, self.ldc = a, int(lda), b, int(ldb), c, int(ldc)


class TrsmProblem:
    """One problem: the arguments of level3.trsm, with DeviceArray operands."""

    __slots__ = ("side", "uplo", "trans", "diag", "m", "n", "alpha", "a", "lda", "b", "ldb")

    def __init__(self, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb):
        self.side, self.uplo, self.trans, self.diag = int(side), int(uplo), int(trans), int(diag)
        self.m, self.n, self.alpha = int(m), int(n), alpha
        self.a, self.lda, self.b, self.ldb = a, int(lda), b, int(ldb)


class _Plan:
    """Shared handle logic of grouped plans (rblas_grouped_t)."""

    def _create(self, create, arr, count):
        h = C.c_void_p()
        bad = C.c_size_t()
        self.bad_index = None
        status = create(C.byref(h), arr, count, C.byref(bad))
        if status != 0:
            self.bad_index = int(bad.value)
        raise_status(status)
        self.handle = h.value

    def run(self, stream=None) -> None:
        raise_status(_ffi.lib().rblas_grouped_run(self.handle,
                                                  stream.handle if stream else None))

    def run_graph(self, stream) -> None:
        raise_status(_ffi.lib().rblas_grouped_run_graph(self.handle, stream.handle))

    def info(self) -> dict:
        inf = _ffi.GroupedInfo()
        raise_status(_ffi.lib().rblas_grouped_info(self.handle, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in _ffi.GroupedInfo._fields_ if f != "pad"}

    def close(self) -> None:
        if getattr(self, "handle", None):
            raise_status(_ffi.lib().rblas_grouped_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GroupedTrsm(_Plan):
    """A plan over trsm `problems` (all float32 or all float64)."""

    def __init__(self, problems: list[TrsmProblem], dtype):
        dt = np.dtype(dtype)
        if dt == np.float32:
            st, create = _ffi.StrsmProblem, _ffi.lib().rblas_strsm_grouped_create
        elif dt == np.float64:
            st, create = _ffi.DtrsmProblem, _ffi.lib().rblas_dtrsm_grouped_create
        else:
            raise TypeError(f"unsupported dtype {dt}")
        arr = (st * max(1, len(problems)))()
        for i, p in enumerate(problems):
            for x in (p.a, p.b):
                if x.dtype != dt:
                    raise TypeError("all operands must match the plan dtype")
            arr[i] = st(p.side, p.uplo, p.trans, p.diag, p.m, p.n, p.alpha, p.a.ptr, p.a.size,
                        p.lda, p.b.ptr, p.b.size, p.ldb)
        self._keep = list(problems)
        self.dtype = dt
        self._create(create, arr, len(problems))


class GroupedGemm(_Plan):
    """A plan over `problems` (all float32 or all float64)."""

    def __init__(self, problems: list[GemmProblem], dtype):
        dt = np.dtype(dtype)
        if dt == np.float32:
            st, create = _ffi.SgemmProblem, _ffi.lib().rblas_sgemm_grouped_create
        elif dt == np.float64:
            st, create = _ffi.DgemmProblem, _ffi.lib().rblas_dgemm_grouped_create
        else:
            raise TypeError(f"unsupported dtype {dt}")
        arr = (st * max(1, len(problems)))()
        for i, p in enumerate(problems):
            for x in (p.a, p.b, p.c):
                if x.dtype != dt:
                    raise TypeError("all operands must match the plan dtype")
            arr[i] = st(p.layout, p.transa, p.transb, p.m, p.n, p.k, p.alpha, p.a.ptr, p.a.size,
                        p.lda, p.b.ptr, p.b.size, p.ldb, p.beta, p.c.ptr, p.c.size, p.ldc)
        self._keep = list(problems)
        self.dtype = dt
        self._create(create, arr, len(problems))


def get_tuning(dtype) -> list[float]:
    """Planner cost model {c0_64, c1_64, c0_128, c1_128, unit_cap_us}."""
    d = 0 if np.dtype(dtype) == np.float32 else 1
    buf = (C.c_double * 5)()
    raise_status(_ffi.lib().rblas_grouped_get_tuning(d, buf))
    return list(buf)


def set_tuning(dtype, coeffs) -> None:
    d = 0 if np.dtype(dtype) == np.float32 else 1
    buf = (C.c_double * 5)(*[float(x) for x in coeffs])
    raise_status(_ffi.lib().rblas_grouped_set_tuning(d, buf))
