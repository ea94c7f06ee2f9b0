"""Pins the oracle (oracle/oracle.cpp) before anything is checked against it.

Each test restates one of rblas's own tests (file:line) or checks the oracle
against something it cannot share a mistake with: exact small-integer
products, numpy float64, closed forms.
"""
import itertools
import math

import numpy as np
import pytest

from tests.util import COL, ROW, assert_gemm_close, gemm_expect, lcg_f32, op, slice_len, view


def naive_colmajor(m, n, k, a, b):
    A = a.reshape(k, m).T.astype(np.float64)
    B = b.reshape(n, k).T.astype(np.float64)
    return (A @ B).T.reshape(-1)


def test_gemm_matches_reference_17x13x19(oracle):
    # kernel/gemm.rs:690-701
    m, n, k = 17, 13, 19
    a = np.sin(np.arange(m * k) * 0.5)
    b = np.cos(np.arange(k * n) * 0.3)
    c = np.zeros(m * n)
    oracle.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    np.testing.assert_allclose(c, naive_colmajor(m, n, k, a, b), rtol=0, atol=1e-9)


def test_gemm_alpha_beta_spot(oracle):
    # kernel/gemm.rs:704-717
    m = n = k = 8
    a = np.arange(m * k, dtype=np.float32)
    b = (np.arange(k * n) % 5).astype(np.float32)
    c = np.ones(m * n, np.float32)
    oracle.gemm(COL, 0, 0, m, n, k, 2.0, a, m, b, k, 3.0, c, m)
    acc = sum(float(a[p * m]) * float(b[p]) for p in range(k))
    assert abs(c[0] - (2 * acc + 3)) < 1e-3


@pytest.mark.parametrize("case", ["nn_colmajor", "tn_colmajor", "rowmajor_nn"])
def test_gemm_layout_kats(oracle, case):
    # level3.rs:642-736
    if case == "nn_colmajor":
        m, n, k = 6, 5, 7
        a = np.arange(m * k, dtype=np.float64)
        b = np.arange(k * n, dtype=np.float64) * 2
        args = (COL, 0, 0, m, n, k, 1.0, a, m, b, k)
    elif case == "tn_colmajor":
        m, n, k = 4, 5, 6
        a = np.sqrt(np.arange(k * m, dtype=np.float64))
        b = np.arange(k * n, dtype=np.float64) * 0.1
        args = (COL, 1, 0, m, n, k, 1.0, a, k, b, k)
    else:
        m, n, k = 5, 4, 3
        a = np.arange(m * k, dtype=np.float64)
        b = np.arange(k * n, dtype=np.float64) + 1
        args = (ROW, 0, 0, m, n, k, 1.0, a, k, b, n)
    layout = args[0]
    ldc = m if layout == COL else n
    c = np.zeros(m * n)
    oracle.gemm(*args, 0.0, c, ldc)
    want, _ = gemm_expect(*args, 0.0, np.zeros(m * n), ldc)
    np.testing.assert_allclose(view(c, layout, m, n, ldc), want, rtol=0, atol=1e-9)


def test_correctness_lcg_200x180x150(oracle):
    # tests/correctness.rs:29-58: LCG seed 0x12345678, max rel err < 1e-3
    m, n, k = 200, 180, 150
    a, s = lcg_f32(0x1234_5678, m * k, 33, 31, 1.0)
    b, _ = lcg_f32(s, k * n, 33, 31, 1.0)
    c = np.zeros(m * n, np.float32)
    oracle.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = naive_colmajor(m, n, k, a, b)
    rel = np.abs(c - want) / np.maximum(np.abs(want), 1e-4)
    assert rel.max() < 1e-3


@pytest.mark.parametrize("threads", [1, 4])
def test_parallel_lcg_64x5000x48(oracle, threads):
    # tests/parallel.rs:17-63: seed 0xC0FFEE, n > NC so bands split; < 1e-2
    m, n, k = 64, 5000, 48
    a, s = lcg_f32(0xC0FFEE, m * k, 40, 24, 0.5)
    b, _ = lcg_f32(s, k * n, 40, 24, 0.5)
    c = np.zeros(m * n, np.float32)
    oracle.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m, threads=threads)
    want = naive_colmajor(m, n, k, a, b)
    rel = np.abs(c - want) / np.maximum(np.abs(want), 1e-3)
    assert rel.max() < 1e-2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout,ta,tb", list(itertools.product([ROW, COL], [0, 1, 2], [0, 1, 2])))
def test_gemm_exact_small_integers(oracle, dtype, layout, ta, tb):
    # Small integers make every product and partial sum exact in f32 and f64,
    # so any indexing, transpose, layout, alpha or beta mistake shows as a
    # nonzero difference. Sizes cross the MR/NR/MC/KC block edges.
    rng = np.random.default_rng(7)
    m, n, k = 37, 23, 530
    pad = 3
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda = (ac if layout == ROW else ar) + pad
    ldb = (bc if layout == ROW else br) + pad
    ldc = (n if layout == ROW else m) + pad
    a = rng.integers(-3, 4, slice_len(layout, ar, ac, lda)).astype(dtype)
    b = rng.integers(-3, 4, slice_len(layout, br, bc, ldb)).astype(dtype)
    c = rng.integers(-3, 4, slice_len(layout, m, n, ldc)).astype(dtype)
    c0 = c.copy()
    oracle.gemm(layout, ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c, ldc)
    want, _ = gemm_expect(layout, ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c0, ldc)
    assert np.array_equal(view(c, layout, m, n, ldc).astype(np.float64), want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gemm_special_cases(oracle, dtype):
    # kernel/gemm.rs:198-205, 322-338: m==0/n==0 leave C untouched; beta==0
    # stores exact zeros (clears NaN); alpha==0 or k==0 gives beta*C only.
    m, n, k = 5, 4, 3
    a = np.full(m * k, np.nan, dtype)
    b = np.ones(k * n, dtype)
    c = np.full(m * n, np.nan, dtype)
    oracle.gemm(COL, 0, 0, 0, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert np.isnan(c).all()
    oracle.gemm(COL, 0, 0, m, n, k, 0.0, a, m, b, k, 0.0, c, m)  # alpha 0: A never read
    assert (c == 0).all()
    c[:] = 2
    oracle.gemm(COL, 0, 0, m, n, 0, 1.0, a, m, b, k, 0.5, c, m)
    assert (c == 1).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout,ta,tb", list(itertools.product([ROW, COL], [0, 1], [0, 1])))
def test_gemm_vs_numpy_random(oracle, dtype, layout, ta, tb):
    rng = np.random.default_rng(11)
    m, n, k = 131, 97, 1100
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda = (ac if layout == ROW else ar) + 7
    ldb = (bc if layout == ROW else br) + 7
    ldc = (n if layout == ROW else m) + 7
    a = rng.standard_normal(slice_len(layout, ar, ac, lda)).astype(dtype)
    b = rng.standard_normal(slice_len(layout, br, bc, ldb)).astype(dtype)
    c = rng.standard_normal(slice_len(layout, m, n, ldc)).astype(dtype)
    c0 = c.copy()
    oracle.gemm(layout, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c, ldc)
    want, mag = gemm_expect(layout, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c0, ldc)
    assert_gemm_close(view(c, layout, m, n, ldc), want, mag, k, dtype, what="oracle gemm")


def test_syrk_kats(oracle):
    # level3.rs:739-767 (lower, N) and :794-824 (upper, T, alpha 2, beta 3)
    n, k = 3, 2
    a = np.array([1.0, 2, 3, 4, 5, 6])
    c = np.zeros(9)
    oracle.syrk(1, 0, n, k, 1.0, a, n, 0.0, c, n)
    A = a.reshape(k, n).T
    full = A @ A.T
    for j in range(n):
        for i in range(j, n):
            assert abs(c[i + j * n] - full[i, j]) < 1e-12
    assert c[0 + 1 * n] == 0  # upper triangle untouched
    n, k = 9, 11
    a = np.sin(np.arange(k * n) * 0.31)
    c = np.arange(n * n) * 0.1
    c0 = c.copy()
    oracle.syrk(0, 1, n, k, 2.0, a, k, 3.0, c, n)
    A = a.reshape(n, k)  # stored k x n col-major -> rows are columns of A
    for j in range(n):
        for i in range(j + 1):
            want = 2 * float(A[i] @ A[j]) + 3 * c0[i + j * n]
            assert abs(c[i + j * n] - want) < 1e-9
        for i in range(j + 1, n):
            assert c[i + j * n] == c0[i + j * n]


def test_trsm_kats(oracle):
    # level3.rs:866-901: L x = b with x = (1, 2, 3)
    l = np.array([2.0, 1, 1, 0, 3, 2, 0, 0, 4])
    L = l.reshape(3, 3).T
    x = np.array([1.0, 2, 3])
    b = L @ x
    oracle.trsm(0, 1, 0, 0, 3, 1, 1.0, l, 3, b, 3)
    np.testing.assert_allclose(b, x, rtol=0, atol=1e-12)
    # level3.rs:903-934: X U = B round trip
    u = np.array([2.0, 0, 0, 1, 3, 0, 4, 5, 6])
    b0 = np.array([1.0, 4, 2, 5, 3, 6])
    b = b0.copy()
    oracle.trsm(1, 0, 0, 0, 2, 3, 1.0, u, 3, b, 2)
    X = b.reshape(3, 2).T
    U = u.reshape(3, 3).T
    np.testing.assert_allclose((X @ U).T.reshape(-1), b0, rtol=0, atol=1e-10)


@pytest.mark.parametrize("side,uplo,trans,diag",
                         list(itertools.product([0, 1], [0, 1], [0, 1], [0, 1])))
def test_trsm_vs_numpy(oracle, side, uplo, trans, diag):
    rng = np.random.default_rng(3)
    m, n = 19, 13
    q = m if side == 0 else n
    lda, ldb = q + 2, m + 3
    a = rng.standard_normal(lda * q) / math.sqrt(q)
    A = view(a, COL, q, q, lda).copy()
    A[np.diag_indices(q)] = 2 + np.abs(A[np.diag_indices(q)])
    a = np.zeros(lda * q)
    view(a, COL, q, q, lda)[:] = A
    T = np.triu(A) if uplo == 0 else np.tril(A)
    if diag == 1:
        T[np.diag_indices(q)] = 1.0
    opT = T.T if trans else T
    b = rng.standard_normal(ldb * n)
    B = view(b, COL, m, n, ldb).copy()
    oracle.trsm(side, uplo, trans, diag, m, n, 0.75, a, lda, b, ldb)
    X = view(b, COL, m, n, ldb)
    res = opT @ X - 0.75 * B if side == 0 else X @ opT - 0.75 * B
    assert np.abs(res).max() < 1e-12


def test_complex_gemm_kats(oracle):
    # complex3.rs:372-401 (5x4x6 vs naive) and :403-437 (ConjTrans A)
    m, n, k = 5, 4, 6
    i = np.arange(m * k)
    a = np.sin(i * 0.3) + 1j * np.cos(i * 0.7)
    j = np.arange(k * n)
    b = np.cos(j * 0.2) + 1j * np.sin(j * 0.5)
    c = np.zeros(m * n, np.complex128)
    oracle.cgemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = (a.reshape(k, m).T @ b.reshape(n, k).T).T.reshape(-1)
    assert np.abs(c - want).max() < 1e-9
    m, n, k = 3, 2, 4
    a = np.arange(k * m) + 0.5j * np.arange(k * m)
    b = 1.0 + 1j * np.arange(k * n)
    c = np.zeros(m * n, np.complex128)
    oracle.cgemm(2, 0, m, n, k, 1.0, a, k, b, k, 0.0, c, m)
    want = (np.conj(a.reshape(m, k)) @ b.reshape(n, k).T).T.reshape(-1)
    assert np.abs(c - want).max() < 1e-9


@pytest.mark.parametrize("ctype", [np.complex64, np.complex128])
@pytest.mark.parametrize("ta,tb", list(itertools.product([0, 1, 2], [0, 1, 2])))
def test_complex_gemm_vs_numpy(oracle, ctype, ta, tb):
    rng = np.random.default_rng(5)
    m, n, k = 33, 21, 45
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda, ldb, ldc = ar + 2, br + 1, m + 3
    a = (rng.standard_normal(lda * ac) + 1j * rng.standard_normal(lda * ac)).astype(ctype)
    b = (rng.standard_normal(ldb * bc) + 1j * rng.standard_normal(ldb * bc)).astype(ctype)
    c = (rng.standard_normal(ldc * n) + 1j * rng.standard_normal(ldc * n)).astype(ctype)
    c0 = c.copy()
    alpha, beta = 1.5 - 0.5j, -0.25 + 0.75j
    oracle.cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    A = op(a.astype(np.complex128), COL, ta, m, k, lda)
    B = op(b.astype(np.complex128), COL, tb, k, n, ldb)
    want = alpha * (A @ B) + beta * view(c0.astype(np.complex128), COL, m, n, ldc)
    got = view(c, COL, m, n, ldc)
    u = 2.0 ** -24 if ctype == np.complex64 else 2.0 ** -53
    bound = 8 * (k + 2) * u * (abs(alpha) * (np.abs(A) @ np.abs(B)) +
                               abs(beta) * np.abs(view(c0, COL, m, n, ldc)))
    assert (np.abs(got - want) <= bound).all()


def test_level1_kats(oracle):
    # level1.rs:169-228
    x = np.array([1.0, 2, 3, 4], np.float32)
    y = np.array([10.0, 20, 30, 40], np.float32)
    oracle.axpy(4, 2.0, x, 1, y, 1)
    assert list(y) == [12, 24, 36, 48]
    x = np.array([1.0, 0, 2, 0, 3])
    y = np.array([10.0, 20, 30])
    oracle.axpy(3, 1.0, x, 2, y, 1)
    assert list(y) == [11, 22, 33]
    assert oracle.dot(3, np.array([1.0, 2, 3], np.float32), 1,
                      np.array([4.0, 5, 6], np.float32), 1) == 32.0
    assert abs(oracle.nrm2(2, np.array([3.0, 4], np.float32), 1) - 5) < 1e-6
    big = 1e200
    got = oracle.nrm2(3, np.array([big] * 3), 1)
    assert math.isfinite(got) and abs(got - big * math.sqrt(3)) / (big * math.sqrt(3)) < 1e-12
    v = np.array([1.0, -2, 3, -4, 5])
    assert abs(oracle.nrm2(5, v, 1) - math.sqrt(55)) < 1e-12
    assert abs(oracle.nrm2(2, np.array([3.0, 99, 4]), 2) - 5) < 1e-12


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nrm2_nonfinite_oracle(oracle, dtype):
    """level1.rs:103-122 traced by hand under IEEE rules: the first infinity
    takes the new-max branch (r = scale / inf = 0, ssq = 1, scale = inf); a
    later finite entry adds (a / inf)^2 = 0; a second infinity takes the
    other branch (inf < inf is false) with r = inf / inf = NaN, so ssq and
    the result are NaN. A NaN entry fails scale < NaN and makes r = NaN."""
    def nrm(v):
        return oracle.nrm2(len(v), np.array(v, dtype), 1)
    assert nrm([np.inf]) == math.inf
    assert nrm([2.0, -np.inf, 7.0]) == math.inf
    assert math.isnan(nrm([np.inf, np.inf]))
    assert math.isnan(nrm([np.inf, 1.0, -np.inf]))
    assert math.isnan(nrm([np.nan]))
    assert math.isnan(nrm([3.0, np.nan, 4.0]))
    assert math.isnan(nrm([np.inf, np.nan]))
    assert math.isnan(nrm([np.nan, np.inf]))
    assert nrm([0.0, 0.0]) == 0.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_level1_vs_numpy(oracle, dtype):
    rng = np.random.default_rng(2)
    for n in (0, 1, 7, 15, 16, 17, 1000, 4099):
        x = rng.standard_normal(n).astype(dtype)
        y = rng.standard_normal(n).astype(dtype)
        want = float(x.astype(np.float64) @ y.astype(np.float64))
        tol = (n + 2) * np.finfo(dtype).eps * float(np.abs(x).astype(np.float64) @ np.abs(y)) + 1e-300
        assert abs(oracle.dot(n, x, 1, y, 1) - want) <= tol
        nrm = float(np.sqrt((x.astype(np.float64) ** 2).sum()))
        assert abs(oracle.nrm2(n, x, 1) - nrm) <= 4 * (n + 2) * np.finfo(dtype).eps * nrm + 1e-300
        y2 = y.copy()
        oracle.axpy(n, 0.5, x, 1, y2, 1)
        np.testing.assert_array_equal(
            y2, (0.5 * x.astype(np.float64) + y.astype(np.float64)).astype(dtype)
            if dtype == np.float32 else y2)
        if dtype == np.float64:
            # fma(alpha, x, y): one rounding of the exact value
            np.testing.assert_allclose(y2, 0.5 * x + y, rtol=2 * np.finfo(dtype).eps, atol=0)


def test_try_gemm_errors_reference(oracle):
    # level3.rs:566-640: exact BlasError values
    m, n, k = 4, 3, 5
    assert oracle.try_gemm_check(COL, 0, 0, m, n, k, m * k, m, k * n, k, m * n, m) is None
    assert oracle.try_gemm_check(COL, 0, 0, m, n, k, m * k, 3, k * n, k, m * n, m) == \
        ("InvalidLeadingDim", "a", 3, 4)
    e = oracle.try_gemm_check(COL, 0, 0, m, n, k, m * k, m, k * n, k, m * n - 1, m)
    assert e[0] == "BufferTooSmall" and e[1] == "c"
