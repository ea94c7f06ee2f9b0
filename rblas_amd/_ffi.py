"""ctypes binding of librblas_hip.so (declared in include/rblas_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librblas_hip.so")

sz, i32, vp = C.c_size_t, C.c_int32, C.c_void_p
f32, f64 = C.c_float, C.c_double


class RblasError(C.Structure):
    """`rblas_error_t` — a BlasError (error.rs:15-38) as a C struct."""
    _fields_ = [("kind", C.c_int32), ("which", C.c_char), ("got", C.c_uint64),
                ("bound", C.c_uint64)]


class Complex32(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class Complex64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _problem_fields(t):
    return [("layout", C.c_int32), ("transa", C.c_int32), ("transb", C.c_int32),
            ("m", sz), ("n", sz), ("k", sz), ("alpha", t), ("a", vp), ("a_len", sz),
            ("lda", sz), ("b", vp), ("b_len", sz), ("ldb", sz), ("beta", t), ("c", vp),
            ("c_len", sz), ("ldc", sz)]


class SgemmProblem(C.Structure):
    """`rblas_sgemm_problem_t`: one gemm of a grouped plan (f32)."""
    _fields_ = _problem_fields(C.c_float)


class DgemmProblem(C.Structure):
    """`rblas_dgemm_problem_t`: one gemm of a grouped plan (f64)."""
    _fields_ = _problem_fields(C.c_double)


def _trsm_problem_fields(t):
    return [("side", C.c_int32), ("uplo", C.c_int32), ("trans", C.c_int32), ("diag", C.c_int32),
            ("m", sz), ("n", sz), ("alpha", t), ("a", vp), ("a_len", sz), ("lda", sz),
            ("b", vp), ("b_len", sz), ("ldb", sz)]


class StrsmProblem(C.Structure):
    """`rblas_strsm_problem_t`: one trsm of a grouped plan (f32)."""
    _fields_ = _trsm_problem_fields(C.c_float)


class DtrsmProblem(C.Structure):
    """`rblas_dtrsm_problem_t`: one trsm of a grouped plan (f64)."""
    _fields_ = _trsm_problem_fields(C.c_double)


class GroupedInfo(C.Structure):
    """`rblas_grouped_info_t`."""
    _fields_ = [("problems", sz), ("units", sz), ("grid", C.c_int32), ("tiles64", C.c_int32),
                ("tiles128", C.c_int32), ("split_problems", C.c_int32),
                ("scale_problems", C.c_int32), ("pad", C.c_int32), ("flops", C.c_double),
                ("est_us", C.c_double), ("workspace_bytes", sz)]


class ServeConfig(C.Structure):
    """`rblas_serve_config_t`."""
    _fields_ = [("mode", C.c_int32), ("max_inflight", C.c_int32), ("max_batch", C.c_int32),
                ("capture_count", C.c_int32), ("hash_col_step", C.c_int32), ("pad", C.c_int32),
                ("hbm_cap_bytes", C.c_double),
                ("capture_ids", C.POINTER(C.c_int32)), ("capture_out", C.POINTER(C.c_void_p))]


class ServeStats(C.Structure):
    """`rblas_serve_stats_t`."""
    _fields_ = [("wall_s", C.c_double), ("flops", C.c_double), ("p50_s", C.c_double),
                ("p99_s", C.c_double), ("peak_footprint_bytes", C.c_double),
                ("batches", C.c_int32), ("max_inflight_seen", C.c_int32)]


def _gemm_args(t):
    return [i32, i32, i32, sz, sz, sz, t, vp, sz, sz, vp, sz, sz, t, vp, sz, sz]


def _gemv_args(t):
    return [i32, i32, sz, sz, t, vp, sz, sz, vp, sz, sz, t, vp, sz, sz]


def _ger_args(t):
    return [i32, sz, sz, t, vp, sz, sz, vp, sz, sz, vp, sz, sz]


# name -> (argtypes, restype). Every symbol here is declared in the header;
# tests check that the library exports each of them.
SIGNATURES: dict[str, tuple[list, object]] = {
    "rblas_last_error_message": ([], C.c_char_p),
    "rblas_sgemm": (_gemm_args(f32), i32),
    "rblas_dgemm": (_gemm_args(f64), i32),
    "rblas_sgemm_try": (_gemm_args(f32) + [C.POINTER(RblasError)], i32),
    "rblas_dgemm_try": (_gemm_args(f64) + [C.POINTER(RblasError)], i32),
    "rblas_sgemm_dev": (_gemm_args(f32) + [vp], i32),
    "rblas_sgemm_ex": (_gemm_args(f32) + [i32], i32),
    "rblas_sgemm_ex_dev": (_gemm_args(f32) + [i32, vp], i32),
    "rblas_dgemm_dev": (_gemm_args(f64) + [vp], i32),
    "rblas_dgemm_dev_noalloc": (_gemm_args(f64) + [vp], i32),
    "rblas_cgemm": ([i32, i32, sz, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, Complex32, vp, sz,
                     sz], i32),
    "rblas_zgemm": ([i32, i32, sz, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, Complex64, vp, sz,
                     sz], i32),
    "rblas_cgemm_dev": ([i32, i32, sz, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, Complex32, vp,
                         sz, sz, vp], i32),
    "rblas_zgemm_dev": ([i32, i32, sz, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, Complex64, vp,
                         sz, sz, vp], i32),
    "rblas_ssyrk": ([i32, i32, sz, sz, f32, vp, sz, sz, f32, vp, sz, sz], i32),
    "rblas_dsyrk": ([i32, i32, sz, sz, f64, vp, sz, sz, f64, vp, sz, sz], i32),
    "rblas_ssyrk_dev": ([i32, i32, sz, sz, f32, vp, sz, sz, f32, vp, sz, sz, vp], i32),
    "rblas_dsyrk_dev": ([i32, i32, sz, sz, f64, vp, sz, sz, f64, vp, sz, sz, vp], i32),
    "rblas_strsm": ([i32, i32, i32, i32, sz, sz, f32, vp, sz, sz, vp, sz, sz], i32),
    "rblas_dtrsm": ([i32, i32, i32, i32, sz, sz, f64, vp, sz, sz, vp, sz, sz], i32),
    "rblas_strsm_dev": ([i32, i32, i32, i32, sz, sz, f32, vp, sz, sz, vp, sz, sz, vp], i32),
    "rblas_dtrsm_dev": ([i32, i32, i32, i32, sz, sz, f64, vp, sz, sz, vp, sz, sz, vp], i32),
    "rblas_ssymm": ([i32, i32, sz, sz, f32, vp, sz, sz, vp, sz, sz, f32, vp, sz, sz], i32),
    "rblas_dsymm": ([i32, i32, sz, sz, f64, vp, sz, sz, vp, sz, sz, f64, vp, sz, sz], i32),
    "rblas_ssymm_dev": ([i32, i32, sz, sz, f32, vp, sz, sz, vp, sz, sz, f32, vp, sz, sz, vp], i32),
    "rblas_dsymm_dev": ([i32, i32, sz, sz, f64, vp, sz, sz, vp, sz, sz, f64, vp, sz, sz, vp], i32),
    "rblas_csymm": ([i32, i32, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, Complex32, vp, sz, sz],
                    i32),
    "rblas_zsymm": ([i32, i32, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, Complex64, vp, sz, sz],
                    i32),
    "rblas_chemm": ([i32, i32, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, Complex32, vp, sz, sz],
                    i32),
    "rblas_zhemm": ([i32, i32, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, Complex64, vp, sz, sz],
                    i32),
    "rblas_csymm_dev": ([i32, i32, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, Complex32, vp, sz,
                         sz, vp], i32),
    "rblas_zsymm_dev": ([i32, i32, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, Complex64, vp, sz,
                         sz, vp], i32),
    "rblas_chemm_dev": ([i32, i32, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, Complex32, vp, sz,
                         sz, vp], i32),
    "rblas_zhemm_dev": ([i32, i32, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, Complex64, vp, sz,
                         sz, vp], i32),
    "rblas_ctrsm": ([i32, i32, i32, i32, sz, sz, Complex32, vp, sz, sz, vp, sz, sz], i32),
    "rblas_ztrsm": ([i32, i32, i32, i32, sz, sz, Complex64, vp, sz, sz, vp, sz, sz], i32),
    "rblas_ctrsm_dev": ([i32, i32, i32, i32, sz, sz, Complex32, vp, sz, sz, vp, sz, sz, vp], i32),
    "rblas_ztrsm_dev": ([i32, i32, i32, i32, sz, sz, Complex64, vp, sz, sz, vp, sz, sz, vp], i32),
    "rblas_sgemv": (_gemv_args(f32), i32),
    "rblas_dgemv": (_gemv_args(f64), i32),
    "rblas_cgemv": (_gemv_args(Complex32), i32),
    "rblas_zgemv": (_gemv_args(Complex64), i32),
    "rblas_sgemv_dev": (_gemv_args(f32) + [vp], i32),
    "rblas_dgemv_dev": (_gemv_args(f64) + [vp], i32),
    "rblas_cgemv_dev": (_gemv_args(Complex32) + [vp], i32),
    "rblas_zgemv_dev": (_gemv_args(Complex64) + [vp], i32),
    "rblas_sger": (_ger_args(f32), i32),
    "rblas_dger": (_ger_args(f64), i32),
    "rblas_cgeru": (_ger_args(Complex32), i32),
    "rblas_zgeru": (_ger_args(Complex64), i32),
    "rblas_cgerc": (_ger_args(Complex32), i32),
    "rblas_zgerc": (_ger_args(Complex64), i32),
    "rblas_sger_dev": (_ger_args(f32) + [vp], i32),
    "rblas_dger_dev": (_ger_args(f64) + [vp], i32),
    "rblas_cgeru_dev": (_ger_args(Complex32) + [vp], i32),
    "rblas_zgeru_dev": (_ger_args(Complex64) + [vp], i32),
    "rblas_cgerc_dev": (_ger_args(Complex32) + [vp], i32),
    "rblas_zgerc_dev": (_ger_args(Complex64) + [vp], i32),
    "rblas_sdot": ([sz, vp, sz, sz, vp, sz, sz, C.POINTER(f32)], i32),
    "rblas_ddot": ([sz, vp, sz, sz, vp, sz, sz, C.POINTER(f64)], i32),
    "rblas_snrm2": ([sz, vp, sz, sz, C.POINTER(f32)], i32),
    "rblas_dnrm2": ([sz, vp, sz, sz, C.POINTER(f64)], i32),
    "rblas_saxpy": ([sz, f32, vp, sz, sz, vp, sz, sz], i32),
    "rblas_daxpy": ([sz, f64, vp, sz, sz, vp, sz, sz], i32),
    "rblas_sdot_dev": ([sz, vp, sz, sz, vp, sz, sz, vp, vp], i32),
    "rblas_ddot_dev": ([sz, vp, sz, sz, vp, sz, sz, vp, vp], i32),
    "rblas_snrm2_dev": ([sz, vp, sz, sz, vp, vp], i32),
    "rblas_dnrm2_dev": ([sz, vp, sz, sz, vp, vp], i32),
    "rblas_saxpy_dev": ([sz, f32, vp, sz, sz, vp, sz, sz, vp], i32),
    "rblas_daxpy_dev": ([sz, f64, vp, sz, sz, vp, sz, sz, vp], i32),
    "rblas_sgemm_grouped_create": ([C.POINTER(vp), C.POINTER(SgemmProblem), sz,
                                    C.POINTER(sz)], i32),
    "rblas_dgemm_grouped_create": ([C.POINTER(vp), C.POINTER(DgemmProblem), sz,
                                    C.POINTER(sz)], i32),
    "rblas_strsm_grouped_create": ([C.POINTER(vp), C.POINTER(StrsmProblem), sz,
                                    C.POINTER(sz)], i32),
    "rblas_dtrsm_grouped_create": ([C.POINTER(vp), C.POINTER(DtrsmProblem), sz,
                                    C.POINTER(sz)], i32),
    "rblas_grouped_run": ([vp, vp], i32),
    "rblas_grouped_run_graph": ([vp, vp], i32),
    "rblas_grouped_info": ([vp, C.POINTER(GroupedInfo)], i32),
    "rblas_grouped_destroy": ([vp], i32),
    "rblas_grouped_set_tuning": ([i32, C.POINTER(f64)], i32),
    "rblas_grouped_get_tuning": ([i32, C.POINTER(f64)], i32),
    "rblas_serve_run": ([vp, sz, C.POINTER(ServeConfig), vp, C.POINTER(ServeStats)], i32),
    "rblas_malloc": ([C.POINTER(vp), sz], i32),
    "rblas_free": ([vp], i32),
    "rblas_memcpy_htod": ([vp, vp, sz, vp], i32),
    "rblas_memcpy_dtoh": ([vp, vp, sz, vp], i32),
    "rblas_memcpy_dtod": ([vp, vp, sz, vp], i32),
    "rblas_stream_create": ([C.POINTER(vp)], i32),
    "rblas_stream_destroy": ([vp], i32),
    "rblas_stream_synchronize": ([vp], i32),
    "rblas_event_create": ([C.POINTER(vp)], i32),
    "rblas_event_destroy": ([vp], i32),
    "rblas_event_record": ([vp, vp], i32),
    "rblas_event_elapsed_ms": ([C.POINTER(C.c_float), vp, vp], i32),
    "rblas_device_synchronize": ([], i32),
    "rblas_device_count": ([], i32),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m rblas_amd.build` "
                "(the HIP library is the only implementation; there is no fallback)")
        lb = C.CDLL(LIB_PATH)
        for name, (argtypes, restype) in SIGNATURES.items():
            fn = getattr(lb, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = lb
    return _lib


def last_error() -> str:
    return lib().rblas_last_error_message().decode("utf-8", "replace")
