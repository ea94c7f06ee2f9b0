"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg). Never imported by the product package.

Python face of oracle/liboracle.so (a C++ restatement of rblas's CPU
algorithms, see oracle.cpp) plus restatements of rblas's argument checks:

  check_matrix / try_gemm_check   level3.rs:68-140
  gemm_panics                     level3.rs:143-222 + kernel/gemm.rs:198-205,322-338
  syrk_panics / trsm_panics       level3.rs:236-294 / 390-520 (index walks)
  symm / hesy and their panics    level3.rs:301-381, complex3.rs:127-247
  ctrsm (C++)                     complex3.rs:252-349 (panics: trsm_panics, same loops)
  gemv / ger / geru / gerc (C++)  level2.rs:11-104, complex2.rs:11-142
  gemv_panics / ger_panics        their index walks

All arrays are numpy, 1-D "slices" exactly as rblas takes them.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ROW_MAJOR, COL_MAJOR = 0, 1
NO_TRANS, TRANS, CONJ_TRANS = 0, 1, 2

_lib = None


def build() -> str:
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        sz, i = C.c_size_t, C.c_int
        f32p, f64p = C.POINTER(C.c_float), C.POINTER(C.c_double)
        for name, t, tp in (("oracle_sgemm", C.c_float, f32p), ("oracle_dgemm", C.c_double, f64p)):
            fn = getattr(_lib, name)
            fn.argtypes = [i, i, i, sz, sz, sz, t, tp, sz, tp, sz, t, tp, sz, i]
            fn.restype = None
        for name, t, tp in (("oracle_ssyrk", C.c_float, f32p), ("oracle_dsyrk", C.c_double, f64p)):
            fn = getattr(_lib, name)
            fn.argtypes = [i, i, sz, sz, t, tp, sz, t, tp, sz, i]
            fn.restype = None
        for name, t, tp in (("oracle_strsm", C.c_float, f32p), ("oracle_dtrsm", C.c_double, f64p)):
            fn = getattr(_lib, name)
            fn.argtypes = [i, i, i, i, sz, sz, t, tp, sz, tp, sz]
            fn.restype = None
        for name, t, tp in (("oracle_cgemm", C.c_float, f32p), ("oracle_zgemm", C.c_double, f64p)):
            fn = getattr(_lib, name)
            fn.argtypes = [i, i, sz, sz, sz, t, t, tp, sz, tp, sz, t, t, tp, sz, i]
            fn.restype = None
        for name, t, tp in (("oracle_ctrsm", C.c_float, f32p), ("oracle_ztrsm", C.c_double, f64p)):
            fn = getattr(_lib, name)
            fn.argtypes = [i, i, i, i, sz, sz, t, t, tp, sz, tp, sz]
            fn.restype = None
        for pre, t, tp in (("s", C.c_float, f32p), ("d", C.c_double, f64p)):
            fn = getattr(_lib, f"oracle_{pre}dot")
            fn.argtypes = [sz, tp, sz, tp, sz]
            fn.restype = t
            fn = getattr(_lib, f"oracle_{pre}axpy")
            fn.argtypes = [sz, t, tp, sz, tp, sz]
            fn.restype = None
            fn = getattr(_lib, f"oracle_{pre}nrm2")
            fn.argtypes = [sz, tp, sz]
            fn.restype = t
        for pre, t, tp in (("s", C.c_float, f32p), ("d", C.c_double, f64p)):
            fn = getattr(_lib, f"oracle_{pre}gemv")
            fn.argtypes = [i, i, sz, sz, t, tp, sz, tp, sz, t, tp, sz]
            fn.restype = None
            fn = getattr(_lib, f"oracle_{pre}ger")
            fn.argtypes = [i, sz, sz, t, tp, sz, tp, sz, tp, sz]
            fn.restype = None
        for pre, t, tp in (("c", C.c_float, f32p), ("z", C.c_double, f64p)):
            fn = getattr(_lib, f"oracle_{pre}gemv")
            fn.argtypes = [i, i, sz, sz, t, t, tp, sz, tp, sz, t, t, tp, sz]
            fn.restype = None
            fn = getattr(_lib, f"oracle_{pre}rank1")
            fn.argtypes = [i, sz, sz, t, t, tp, sz, tp, sz, tp, sz, i]
            fn.restype = None
        _lib.oracle_has_avx2.restype = i
    return _lib


def _ptr(x: np.ndarray):
    ct = C.c_float if x.dtype == np.float32 else C.c_double
    return x.ctypes.data_as(C.POINTER(ct))


def _pre(x: np.ndarray) -> str:
    if x.dtype == np.float32:
        return "s"
    if x.dtype == np.float64:
        return "d"
    raise TypeError(f"oracle: unsupported dtype {x.dtype}")


# ----------------------------------------------------------------- compute
def gemm(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, threads=1):
    """level3::gemm restated (in place on c). Caller guarantees in-bounds."""
    p = _pre(c)
    getattr(lib(), f"oracle_{p}gemm")(layout, ta, tb, m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb,
                                      beta, _ptr(c), ldc, threads)


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, threads=1):
    getattr(lib(), f"oracle_{_pre(c)}syrk")(uplo, trans, n, k, alpha, _ptr(a), lda, beta, _ptr(c),
                                           ldc, threads)


def trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb):
    getattr(lib(), f"oracle_{_pre(b)}trsm")(side, uplo, trans, diag, m, n, alpha, _ptr(a), lda,
                                           _ptr(b), ldb)


def cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, threads=1):
    """complex3::gemm (4M) restated. a/b/c are complex64/complex128 arrays."""
    real = np.float32 if c.dtype == np.complex64 else np.float64
    fn = lib().oracle_cgemm if real == np.float32 else lib().oracle_zgemm
    av, bv, cv = (x.view(real) for x in (a, b, c))
    fn(ta, tb, m, n, k, complex(alpha).real, complex(alpha).imag, _ptr(av), lda, _ptr(bv), ldb,
       complex(beta).real, complex(beta).imag, _ptr(cv), ldc, threads)


def ctrsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb):
    """complex3::trsm (complex3.rs:252-349) restated; in place on b
    (complex64 / complex128 arrays)."""
    real = np.float32 if b.dtype == np.complex64 else np.float64
    fn = lib().oracle_ctrsm if real == np.float32 else lib().oracle_ztrsm
    av, bv = (x.view(real) for x in (a, b))
    fn(side, uplo, trans, diag, m, n, complex(alpha).real, complex(alpha).imag, _ptr(av), lda,
       _ptr(bv), ldb)


def dot(n, x, incx, y, incy):
    return getattr(lib(), f"oracle_{_pre(x)}dot")(n, _ptr(x), incx, _ptr(y), incy)


def axpy(n, alpha, x, incx, y, incy):
    getattr(lib(), f"oracle_{_pre(y)}axpy")(n, alpha, _ptr(x), incx, _ptr(y), incy)


def nrm2(n, x, incx):
    return getattr(lib(), f"oracle_{_pre(x)}nrm2")(n, _ptr(x), incx)


def _cpre(x: np.ndarray) -> tuple[str, type]:
    if x.dtype == np.complex64:
        return "c", np.float32
    if x.dtype == np.complex128:
        return "z", np.float64
    raise TypeError(f"oracle: unsupported dtype {x.dtype}")


def gemv(layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    """level2::gemv (level2.rs:11-61) or complex2::gemv (complex2.rs:11-76)
    by y's dtype; in place on y."""
    if np.iscomplexobj(y):
        pre, real = _cpre(y)
        al, be = complex(alpha), complex(beta)
        getattr(lib(), f"oracle_{pre}gemv")(layout, trans, m, n, al.real, al.imag,
                                            _ptr(a.view(real)), lda, _ptr(x.view(real)), incx,
                                            be.real, be.imag, _ptr(y.view(real)), incy)
        return
    getattr(lib(), f"oracle_{_pre(y)}gemv")(layout, trans, m, n, alpha, _ptr(a), lda, _ptr(x),
                                            incx, beta, _ptr(y), incy)


def ger(layout, m, n, alpha, x, incx, y, incy, a, lda):
    """level2::ger (level2.rs:65-88); in place on a."""
    getattr(lib(), f"oracle_{_pre(a)}ger")(layout, m, n, alpha, _ptr(x), incx, _ptr(y), incy,
                                           _ptr(a), lda)


def _rank1(layout, m, n, alpha, x, incx, y, incy, a, lda, conj):
    pre, real = _cpre(a)
    al = complex(alpha)
    getattr(lib(), f"oracle_{pre}rank1")(layout, m, n, al.real, al.imag, _ptr(x.view(real)), incx,
                                         _ptr(y.view(real)), incy, _ptr(a.view(real)), lda,
                                         int(conj))


def geru(layout, m, n, alpha, x, incx, y, incy, a, lda):
    """complex2::geru (complex2.rs:80-93); in place on a."""
    _rank1(layout, m, n, alpha, x, incx, y, incy, a, lda, False)


def gerc(layout, m, n, alpha, x, incx, y, incy, a, lda):
    """complex2::gerc (complex2.rs:97-110); in place on a."""
    _rank1(layout, m, n, alpha, x, incx, y, incy, a, lda, True)


def has_avx2() -> bool:
    return bool(lib().oracle_has_avx2())


# ------------------------------------------------------------------ checks
_MASK = (1 << 64) - 1  # usize arithmetic wraps (release build)


def check_matrix(layout, which, rows, cols, lead, length):
    """level3.rs:107-140. Returns None or a tuple (kind, which, got, bound)
    with kind 'InvalidLeadingDim' or 'BufferTooSmall'."""
    if rows == 0 or cols == 0:
        return None
    if layout == ROW_MAJOR:
        min_lead, need = cols, ((rows - 1) * lead + cols) & _MASK
    else:
        min_lead, need = rows, ((cols - 1) * lead + rows) & _MASK
    if lead < min_lead:
        return ("InvalidLeadingDim", which, lead, min_lead)
    if length < need:
        return ("BufferTooSmall", which, length, need)
    return None


def try_gemm_check(layout, ta, tb, m, n, k, a_len, lda, b_len, ldb, c_len, ldc):
    """level3.rs:84-98: a, then b, then c."""
    at, bt = ta != NO_TRANS, tb != NO_TRANS
    for which, rows, cols, ld, ln in (("a", k if at else m, m if at else k, lda, a_len),
                                      ("b", n if bt else k, k if bt else n, ldb, b_len),
                                      ("c", m, n, ldc, c_len)):
        e = check_matrix(layout, which, rows, cols, ld, ln)
        if e:
            return e
    return None


def gemm_panics(layout, ta, tb, m, n, k, a_len, lda, b_len, ldb, c_len, ldc, alpha, beta):
    """True where rblas's unchecked gemm panics (or writes C out of bounds)."""
    at, bt = ta != NO_TRANS, tb != NO_TRANS
    if layout == ROW_MAJOR:  # level3.rs:56-60
        at, bt, m, n = bt, at, n, m
        a_len, lda, b_len, ldb = b_len, ldb, a_len, lda
    if not at and not bt:  # level3.rs:164-166
        if not (a_len >= (((k - 1) * lda + m) & _MASK) or k == 0):
            return True
        if not (b_len >= (((n - 1) * ldb + k) & _MASK) or n == 0):
            return True
        if not (c_len >= (((n - 1) * ldc + m) & _MASK) or n == 0):
            return True
        return False
    # materialize (level3.rs:193-220): largest gathered index must be < len
    if m > 0 and k > 0:
        hi = (k - 1) + (m - 1) * lda if at else (m - 1) + (k - 1) * lda
        if hi >= a_len:
            return True
    if n > 0 and k > 0:
        hi = (n - 1) + (k - 1) * ldb if bt else (k - 1) + (n - 1) * ldb
        if hi >= b_len:
            return True
    if m == 0 or n == 0:
        return False
    touches_c = beta != 1 or (k != 0 and alpha != 0)
    return touches_c and (m - 1) + (n - 1) * ldc >= c_len


def axpy_panics(n, alpha, x_len, incx, y_len, incy):
    """level1.rs:17-41: no-op (no checks) when n == 0 or alpha == 0."""
    if n == 0 or alpha == 0:
        return False
    if incx == 1 and incy == 1:
        return not (x_len >= n and y_len >= n)
    if not (incx > 0 and incy > 0):
        return True
    return not ((n - 1) * incx < x_len and (n - 1) * incy < y_len)


def dot_panics(n, x_len, incx, y_len, incy):
    """level1.rs:45-66."""
    if n == 0:
        return False
    if incx == 1 and incy == 1:
        return not (x_len >= n and y_len >= n)
    if not (incx > 0 and incy > 0):
        return True
    return not ((n - 1) * incx < x_len and (n - 1) * incy < y_len)


def nrm2_panics(n, x_len, incx):
    """level1.rs:96-101."""
    if n == 0:
        return False
    return not (incx > 0) or not ((n - 1) * incx < x_len)


def syrk_panics(uplo, trans, n, k, a_len, lda, c_len, ldc):
    """level3.rs:236-294 walked literally (small sizes): n == 0 returns; the
    inner gemm is never NN, so materialize gathers op(A) (n x k) and op(A)^T
    (k x n) from `a` (level3.rs:193-220); then the fold indexes c[i + j*ldc]
    over the uplo triangle. True if any index reaches past its slice."""
    if n == 0:
        return False
    t = trans != NO_TRANS
    for i in range(n):          # op(A)(i, p)
        for p in range(k):
            if (p + i * lda if t else i + p * lda) >= a_len:
                return True
    for p in range(k):          # op(A)^T(p, j) = op(A)(j, p)
        for j in range(n):
            if (p + j * lda if t else j + p * lda) >= a_len:
                return True
    for j in range(n):
        lo, hi = (0, j + 1) if uplo == 0 else (j, n)
        for i in range(lo, hi):
            if i + j * ldc >= c_len:
                return True
    return False


def trsm_panics(side, uplo, trans, diag, m, n, alpha, a_len, lda, b_len, ldb):
    """level3.rs:390-520 walked literally (small sizes): the alpha pre-scale,
    then solve_left per column or solve_right per row, recording every a/b
    index (finish's a_ii argument is evaluated for Unit too). True if any
    index reaches past its slice."""
    t = trans != NO_TRANS
    idx_a = (lambda r, c: c + r * lda) if t else (lambda r, c: r + c * lda)
    eff_upper = (uplo == 0) != t
    bad = [False]

    def ra(r, c):
        if idx_a(r, c) >= a_len:
            bad[0] = True

    def rb(i):
        if i >= b_len:
            bad[0] = True

    if alpha != 1:
        for j in range(n):
            for i in range(m):
                rb(i + j * ldb)
    if side == 0:
        for j in range(n):
            col = j * ldb
            rows = range(m - 1, -1, -1) if eff_upper else range(m)
            for i in rows:
                rb(i + col)
                ks = range(i + 1, m) if eff_upper else range(i)
                for kk in ks:
                    ra(i, kk)
                    rb(kk + col)
                ra(i, i)
    else:
        order = range(n) if eff_upper else range(n - 1, -1, -1)
        for i in range(m):
            for j in order:
                rb(i + j * ldb)
                ks = range(j) if eff_upper else range(j + 1, n)
                for kk in ks:
                    rb(i + kk * ldb)
                    ra(kk, j)
                ra(j, j)
    return bad[0]


# ------------------------------------------------------------ symm / hemm
def full_symmetric(uplo, q, a, lda, herm=False):
    """level3.rs:358-381 (real) / complex3.rs:188-213 (he_or_sy), walked
    literally: the dense q x q column-major matrix from the stored `uplo`
    triangle; herm conjugates the mirrored half and zeroes the diagonal's
    imaginary part."""
    full = np.zeros(q * q, a.dtype)
    for j in range(q):
        for i in range(q):
            here = i <= j if uplo == 0 else i >= j
            if here:
                v = a[i + j * lda]
                if herm and i == j:
                    v = v.real + 0j
            else:
                v = a[j + i * lda]
                if herm:
                    v = np.conj(v)
            full[i + j * q] = v
    return full


def symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    """level3::symm (level3.rs:301-356): full_symmetric, then gemm(ColMajor,
    N, N) with A left (side 0) or right (side 1). In place on c."""
    q = m if side == 0 else n
    full = full_symmetric(uplo, q, a, lda)
    if side == 0:
        gemm(COL_MAJOR, NO_TRANS, NO_TRANS, m, n, m, alpha, full, q, b, ldb, beta, c, ldc)
    else:
        gemm(COL_MAJOR, NO_TRANS, NO_TRANS, m, n, n, alpha, b, ldb, full, q, beta, c, ldc)


def hesy(herm, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    """complex3::symm / hemm (complex3.rs:127-247, he_or_sy): nothing when m
    or n is 0; full_symmetric (Hermitian when herm), then complex3::gemm(N,
    N). In place on c."""
    if m == 0 or n == 0:
        return
    q = m if side == 0 else n
    full = full_symmetric(uplo, q, a, lda, herm)
    if side == 0:
        cgemm(NO_TRANS, NO_TRANS, m, n, m, alpha, full, q, b, ldb, beta, c, ldc)
    else:
        cgemm(NO_TRANS, NO_TRANS, m, n, n, alpha, b, ldb, full, q, beta, c, ldc)


def symm_panics(side, uplo, m, n, alpha, a_len, lda, b_len, ldb, beta, c_len, ldc):
    """level3::symm: full_symmetric indexes a over the stored triangle (all of
    its indices, walked), then the gemm on the dense q x q copy."""
    q = m if side == 0 else n
    for j in range(q):
        for i in range(q):
            here = i <= j if uplo == 0 else i >= j
            if (i + j * lda if here else j + i * lda) >= a_len:
                return True
    if side == 0:
        return gemm_panics(COL_MAJOR, NO_TRANS, NO_TRANS, m, n, m, q * q, q, b_len, ldb, c_len,
                           ldc, alpha, beta)
    return gemm_panics(COL_MAJOR, NO_TRANS, NO_TRANS, m, n, n, b_len, ldb, q * q, q, c_len, ldc,
                       alpha, beta)


def hesy_panics(side, uplo, m, n, a_len, lda, b_len, ldb, c_len, ldc):
    """complex3::symm / hemm: nothing when m or n is 0; full_symmetric's
    indices of a, then complex3::gemm(N, N)'s: split_op gathers op(A) and
    op(B) (complex3.rs:111-118) and the combine updates C (:61-73)."""
    if m == 0 or n == 0:
        return False
    q = m if side == 0 else n
    for j in range(q):
        for i in range(q):
            here = i <= j if uplo == 0 else i >= j
            if (i + j * lda if here else j + i * lda) >= a_len:
                return True
    k = q
    la, lda2, lb, ldb2 = (q * q, q, b_len, ldb) if side == 0 else (b_len, ldb, q * q, q)
    for p in range(k):          # op(A)(i, p) of the m x k left operand
        for i in range(m):
            if i + p * lda2 >= la:
                return True
    for j in range(n):          # op(B)(p, j) of the k x n right operand
        for p in range(k):
            if p + j * ldb2 >= lb:
                return True
    for j in range(n):
        for i in range(m):
            if i + j * ldc >= c_len:
                return True
    return False


def gemv_panics(layout, trans, m, n, alpha, beta, a_len, lda, x_len, incx, y_len, incy,
                complex_=False):
    """level2.rs:11-104 / complex2.rs:11-76 walked literally (small sizes):
    every a / x / y index the loops touch, in order. True if any reaches past
    its slice. Real: y scaled only when beta != 1; complex: always."""
    rs, cs = (lda, 1) if layout == ROW_MAJOR else (1, lda)
    tr = trans != NO_TRANS
    leny = n if tr else m
    bad = [False]

    def touch(i, length):
        if i >= length:
            bad[0] = True

    if complex_ or beta != 1:
        for t in range(leny):
            touch(t * incy, y_len)
    if alpha == 0:
        return bad[0]
    if not tr:
        for i in range(m):
            for j in range(n):
                touch(i * rs + j * cs, a_len)
                touch(j * incx, x_len)
            touch(i * incy, y_len)
    elif complex_:
        for j in range(n):
            for i in range(m):
                touch(i * rs + j * cs, a_len)
                touch(i * incx, x_len)
            touch(j * incy, y_len)
    else:
        for i in range(m):
            touch(i * incx, x_len)
            for j in range(n):
                touch(i * rs + j * cs, a_len)
                touch(j * incy, y_len)
    return bad[0]


def ger_panics(layout, m, n, alpha, x_len, incx, y_len, incy, a_len, lda):
    """level2.rs:65-88 / complex2.rs:112-142 walked literally (small sizes):
    x_i for every row (also when n == 0), then y_j and a_ij."""
    if alpha == 0:
        return False
    rs, cs = (lda, 1) if layout == ROW_MAJOR else (1, lda)
    for i in range(m):
        if i * incx >= x_len:
            return True
        for j in range(n):
            if j * incy >= y_len or i * rs + j * cs >= a_len:
                return True
    return False
