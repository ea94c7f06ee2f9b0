"""Level-2 (gemv, ger; complex gemv, geru, gerc) through the C ABI.

CPU part: the oracle's restatement pinned against rblas's own unit tests
(level2.rs:110-187, complex2.rs:150-182), numpy over every layout, transpose,
stride and padding, and the special cases rblas's code fixes (beta 1 / 0,
alpha 0, the complex beta multiply that always runs); its index walks
against the library's closed-form panic checks.

GPU part, tolerances written here: gemv within
  |y - y64| <= c (k+2) u (|alpha| |op(A)| |x| + |beta| |y0|) + tiny
(k the summed length, u the unit roundoff, c = 2 real, 8 complex) against
float64 / complex128 and against the oracle; exact on small integers; ger /
geru / gerc bit for bit against the oracle (each element is the same one or
two roundings rblas makes).
"""
import itertools

import numpy as np
import pytest

from rblas_amd import level2
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic, DeviceError

ROW, COL = 0, 1
REAL = (np.float32, np.float64)
CPLX = (np.complex64, np.complex128)


def _u(dt):
    return 2.0 ** -24 if dt in (np.float32, np.complex64) else 2.0 ** -53


def _rand(rng, dt, size):
    v = rng.standard_normal(size)
    if np.dtype(dt).kind == "c":
        v = v + 1j * rng.standard_normal(size)
    return v.astype(dt)


def _mat(a, layout, m, n, lda):
    """Logical m x n A from the slice (types.rs:20-25 strides)."""
    rs, cs = (lda, 1) if layout == ROW else (1, lda)
    i = np.arange(m)[:, None] * rs + np.arange(n)[None, :] * cs
    return a[i]


def _problem(rng, dt, layout, trans, m, n, incx, incy, pad=2):
    lda = (n if layout == ROW else m) + pad
    a = _rand(rng, dt, (m - 1) * (lda if layout == ROW else 1) + (n - 1) * (1 if layout == ROW else lda) + 1
               if m and n else 0)
    lenx, leny = (m, n) if trans else (n, m)
    x = _rand(rng, dt, max(0, (lenx - 1) * incx + 1))
    y = _rand(rng, dt, max(0, (leny - 1) * incy + 1))
    return a, lda, x, y, lenx, leny


def _want(a, layout, trans, m, n, lda, alpha, x, incx, beta, y, incy, lenx, leny):
    A = _mat(a.astype(np.complex128), layout, m, n, lda) if m and n else np.zeros((m, n))
    op = A if trans == 0 else (A.T if trans == 1 else A.conj().T)
    xv = x.astype(np.complex128)[np.arange(lenx) * incx] if lenx else np.zeros(0)
    yv = y.astype(np.complex128)[np.arange(leny) * incy] if leny else np.zeros(0)
    want = alpha * (op @ xv) + beta * yv
    mag = abs(alpha) * (np.abs(op) @ np.abs(xv)) + abs(beta) * np.abs(yv)
    return want, mag


def _status(fn):
    try:
        fn()
    except BlasPanic:
        return "panic"
    except ValueError:
        return "bad"
    except DeviceError:
        return "ran"
    return "ran"


# ------------------------------------------------------------------ CPU side
def test_oracle_reference_unit_tests(oracle):
    # level2.rs:110-187
    a = np.array([1.0, 4, 2, 5, 3, 6])
    y = np.zeros(2)
    oracle.gemv(COL, 0, 2, 3, 1.0, a, 2, np.ones(3), 1, 0.0, y, 1)
    assert list(y) == [6.0, 15.0]
    y = np.zeros(3)
    oracle.gemv(COL, 1, 2, 3, 1.0, a, 2, np.ones(2), 1, 0.0, y, 1)
    assert list(y) == [5.0, 7.0, 9.0]
    y = np.zeros(2)
    oracle.gemv(ROW, 0, 2, 3, 1.0, np.array([1.0, 2, 3, 4, 5, 6]), 3, np.array([1.0, 0, 1]), 1,
                0.0, y, 1)
    assert list(y) == [4.0, 10.0]
    A = np.zeros(4)
    oracle.ger(COL, 2, 2, 1.0, np.array([1.0, 2]), 1, np.array([3.0, 4]), 1, A, 2)
    assert list(A) == [3.0, 6.0, 4.0, 8.0]
    # complex2.rs:150-182
    y = np.zeros(1, np.complex128)
    oracle.gemv(COL, 0, 1, 2, 1, np.array([1, 1j]), 1, np.ones(2, np.complex128), 1, 0, y, 1)
    assert y[0] == 1 + 1j
    A = np.zeros(1, np.complex128)
    oracle.gerc(COL, 1, 1, 1, np.ones(1, np.complex128), 1, np.array([1j]), 1, A, 1)
    assert A[0] == -1j


@pytest.mark.parametrize("dt", REAL + CPLX)
def test_oracle_gemv_pinned_against_numpy(oracle, dt):
    rng = np.random.default_rng(3)
    cplx = np.dtype(dt).kind == "c"
    alpha, beta = (0.75 - 0.5j, -0.25 + 1j) if cplx else (0.75, -0.25)
    for layout, trans, (m, n), incx, incy in itertools.product(
            (ROW, COL), (0, 1, 2) if cplx else (0, 1), ((7, 5), (1, 9), (9, 1)), (1, 3), (1, 2)):
        a, lda, x, y, lenx, leny = _problem(rng, dt, layout, trans, m, n, incx, incy)
        want, mag = _want(a, layout, trans, m, n, lda, alpha, x, incx, beta, y, incy, lenx, leny)
        yo = y.copy()
        oracle.gemv(layout, trans, m, n, alpha, a, lda, x, incx, beta, yo, incy)
        got = yo.astype(np.complex128)[np.arange(leny) * incy]
        c = 8 if cplx else 2
        assert (np.abs(got - want) <= c * (lenx + 2) * _u(dt) * mag + 1e-300).all(), \
            (layout, trans, m, n, incx, incy)
        gaps = np.ones(y.size, bool)
        gaps[np.arange(leny) * incy] = False
        assert np.array_equal(yo[gaps], y[gaps])


@pytest.mark.parametrize("dt", REAL + CPLX)
def test_oracle_rank1_pinned_against_numpy(oracle, dt):
    rng = np.random.default_rng(4)
    cplx = np.dtype(dt).kind == "c"
    alpha = 1.5 - 0.5j if cplx else 1.5
    for layout, (m, n), incx, incy, conj in itertools.product(
            (ROW, COL), ((6, 4), (1, 5), (5, 1)), (1, 2), (1, 3), (False, True) if cplx else (False,)):
        lda = (n if layout == ROW else m) + 3
        a = _rand(rng, dt, (m if layout == ROW else n) * lda)
        x = _rand(rng, dt, (m - 1) * incx + 1)
        y = _rand(rng, dt, (n - 1) * incy + 1)
        ao = a.copy()
        if not cplx:
            oracle.ger(layout, m, n, alpha, x, incx, y, incy, ao, lda)
        else:
            (oracle.gerc if conj else oracle.geru)(layout, m, n, alpha, x, incx, y, incy, ao, lda)
        xv = x.astype(np.complex128)[np.arange(m) * incx]
        yv = y.astype(np.complex128)[np.arange(n) * incy]
        if conj:
            yv = yv.conj()
        A0 = _mat(a.astype(np.complex128), layout, m, n, lda)
        want = A0 + alpha * np.outer(xv, yv)
        got = _mat(ao.astype(np.complex128), layout, m, n, lda)
        mag = np.abs(A0) + abs(alpha) * np.outer(np.abs(xv), np.abs(yv))
        assert (np.abs(got - want) <= 8 * _u(dt) * mag).all()
        rs, cs = (lda, 1) if layout == ROW else (1, lda)
        touched = np.zeros(a.size, bool)
        touched[(np.arange(m)[:, None] * rs + np.arange(n)[None, :] * cs).ravel()] = True
        assert np.array_equal(ao[~touched], a[~touched])


def test_oracle_special_cases(oracle):
    """What rblas's code fixes beyond the arithmetic: real beta == 1 leaves y
    alone (NaN and -0 kept), beta == 0 stores 0 (NaN cleared), alpha == 0
    never reads A or x; complex scales y even for beta == 1 (0 * inf in the
    complex multiply turns (inf, 0) into (inf, nan)) and clears it for
    beta == 0."""
    nan, inf = np.nan, np.inf
    a = np.array([nan, nan, nan, nan])
    y = np.array([nan, -0.0])
    oracle.gemv(COL, 0, 2, 2, 0.0, a, 2, np.ones(2), 1, 1.0, y, 1)
    assert np.isnan(y[0]) and np.signbit(y[1])
    oracle.gemv(COL, 1, 2, 2, 0.0, a, 2, np.ones(2), 1, 0.0, y, 1)
    assert list(y) == [0.0, 0.0]
    y = np.array([inf + 0j, 2 - 1j])
    oracle.gemv(COL, 0, 2, 2, 0, np.full(4, nan + 0j), 2, np.ones(2, complex), 1, 1, y, 1)
    assert y[0].real == inf and np.isnan(y[0].imag) and y[1] == 2 - 1j
    oracle.gemv(COL, 2, 2, 2, 0, np.full(4, nan + 0j), 2, np.ones(2, complex), 1, 0, y, 1)
    assert (y == 0).all()
    A = np.full(4, nan)
    oracle.ger(COL, 2, 2, 0.0, np.full(2, nan), 1, np.full(2, nan), 1, A, 2)
    assert np.isnan(A).all()


def test_gemv_panics_match_oracle(oracle):
    checked = panics = 0
    for cplx, layout, trans, m, n, alpha, beta in itertools.product(
            (False, True), (ROW, COL), (0, 1, 2), (0, 1, 3), (0, 2, 3), (0.0, 2.0), (1.0, 0.5)):
        if trans == 2 and not cplx:
            continue
        dt = np.complex128 if cplx else np.float64
        for lda, incx, incy in itertools.product((1, 3, 4), (1, 2), (1, 3)):
            rs, cs = (lda, 1) if layout == ROW else (1, lda)
            need_a = (m - 1) * rs + (n - 1) * cs + 1 if m and n else 0
            lenx, leny = (m, n) if trans else (n, m)
            for a_len in sorted({0, max(need_a - 1, 0), need_a}):
                for x_len in sorted({0, max((lenx - 1) * incx, 0), max((lenx - 1) * incx + 1, 0)}):
                    for y_len in sorted({0, max((leny - 1) * incy, 0), max((leny - 1) * incy + 1, 0)}):
                        want = oracle.gemv_panics(layout, trans, m, n, alpha, beta, a_len, lda,
                                                  x_len, incx, y_len, incy, complex_=cplx)
                        a, x, y = (np.ones(k, dt) for k in (a_len, x_len, y_len))
                        got = _status(lambda: level2.gemv(layout, trans, m, n, alpha, a, lda, x,
                                                          incx, beta, y, incy))
                        if got == "bad":
                            continue
                        assert (got == "panic") == want, (cplx, layout, trans, m, n, alpha, beta,
                                                          lda, incx, incy, a_len, x_len, y_len)
                        checked += 1
                        panics += want
    assert checked > 5000 and panics > 1500


def test_ger_panics_match_oracle(oracle):
    checked = panics = 0
    for name, dt in (("ger", np.float64), ("geru", np.complex128), ("gerc", np.complex64)):
        fn = getattr(level2, name)
        for layout, m, n, alpha, lda, incx, incy in itertools.product(
                (ROW, COL), (0, 1, 3), (0, 2, 3), (0.0, 1.5), (1, 3, 4), (1, 2), (1, 3)):
            rs, cs = (lda, 1) if layout == ROW else (1, lda)
            need_a = (m - 1) * rs + (n - 1) * cs + 1 if m and n else 0
            for a_len in sorted({0, max(need_a - 1, 0), need_a}):
                for x_len in sorted({0, max((m - 1) * incx, 0), (m - 1) * incx + 1 if m else 0}):
                    for y_len in sorted({0, max((n - 1) * incy, 0), (n - 1) * incy + 1 if n else 0}):
                        want = oracle.ger_panics(layout, m, n, alpha, x_len, incx, y_len, incy,
                                                 a_len, lda)
                        a, x, y = (np.ones(k, dt) for k in (a_len, x_len, y_len))
                        got = _status(lambda: fn(layout, m, n, alpha, x, incx, y, incy, a, lda))
                        if got == "bad":
                            continue
                        assert (got == "panic") == want, (name, layout, m, n, alpha, lda, incx,
                                                          incy, a_len, x_len, y_len)
                        checked += 1
                        panics += want
    assert checked > 2000 and panics > 500


def test_bad_enums_and_device_limits():
    y = np.ones(4)
    with pytest.raises(ValueError):
        level2.gemv(2, 0, 2, 2, 1.0, np.ones(4), 2, np.ones(2), 1, 0.0, y, 1)
    with pytest.raises(ValueError):
        level2.gemv(COL, 3, 2, 2, 1.0, np.ones(4), 2, np.ones(2), 1, 0.0, y, 1)
    # incy == 0 with two outputs: rblas accumulates both into y[0]; refused.
    with pytest.raises(ValueError):
        level2.gemv(COL, 0, 2, 2, 1.0, np.ones(4), 2, np.ones(2), 1, 0.0, y, 0)
    # ger writes that alias (lda < m, ColMajor): refused.
    with pytest.raises(ValueError):
        level2.ger(COL, 3, 2, 1.0, np.ones(3), 1, np.ones(2), 1, np.ones(8), 2)
    # Nothing to do: no device needed, no error.
    level2.gemv(COL, 0, 0, 5, 1.0, np.ones(0), 1, np.ones(5), 1, 0.0, np.ones(0), 1)
    level2.gemv(COL, 0, 3, 2, 0.0, np.ones(6), 3, np.ones(2), 1, 1.0, y, 1)
    level2.ger(COL, 3, 0, 1.0, np.ones(3), 1, np.ones(0), 1, np.ones(0), 3)
    level2.geru(ROW, 3, 3, 0.0, np.ones(0, complex), 1, np.ones(0, complex), 1,
                np.ones(0, complex), 3)


# ------------------------------------------------------------------ GPU side
@pytest.mark.gpu
def test_reference_unit_tests_on_gpu(gpu):
    a = np.array([1.0, 4, 2, 5, 3, 6])
    y = np.zeros(2)
    level2.gemv(COL, 0, 2, 3, 1.0, a, 2, np.ones(3), 1, 0.0, y, 1)
    assert list(y) == [6.0, 15.0]
    y = np.zeros(3)
    level2.gemv(COL, 1, 2, 3, 1.0, a, 2, np.ones(2), 1, 0.0, y, 1)
    assert list(y) == [5.0, 7.0, 9.0]
    y = np.zeros(2)
    level2.gemv(ROW, 0, 2, 3, 1.0, np.array([1.0, 2, 3, 4, 5, 6]), 3, np.array([1.0, 0, 1]), 1,
                0.0, y, 1)
    assert list(y) == [4.0, 10.0]
    A = np.zeros(4)
    level2.ger(COL, 2, 2, 1.0, np.array([1.0, 2]), 1, np.array([3.0, 4]), 1, A, 2)
    assert list(A) == [3.0, 6.0, 4.0, 8.0]
    y = np.zeros(1, np.complex128)
    level2.gemv(COL, 0, 1, 2, 1, np.array([1, 1j]), 1, np.ones(2, np.complex128), 1, 0, y, 1)
    assert y[0] == 1 + 1j
    A = np.zeros(1, np.complex128)
    level2.gerc(COL, 1, 1, 1, np.ones(1, np.complex128), 1, np.array([1j]), 1, A, 1)
    assert A[0] == -1j


# Shapes: tiny, one split (many outputs), split sums (few outputs, long sums).
SHAPES = [(7, 5), (1, 300), (300, 1), (2000, 3000), (70, 40000), (40000, 70)]


# ConjTrans only for complex operands (for real ones it is Trans).
GEMV_CASES = [(dt, layout, trans) for dt in REAL + CPLX for layout in (ROW, COL)
              for trans in ((0, 1, 2) if np.dtype(dt).kind == "c" else (0, 1))]


@pytest.mark.gpu
@pytest.mark.parametrize("dt,layout,trans", GEMV_CASES)
def test_gemv_vs_float64_and_oracle(gpu, oracle, dt, layout, trans):
    cplx = np.dtype(dt).kind == "c"
    rng = np.random.default_rng(10 + 3 * layout + trans)
    alpha, beta = (0.75 - 0.5j, -0.25 + 1j) if cplx else (0.75, -0.25)
    c = 8 if cplx else 2
    for (m, n), incx, incy in zip(SHAPES, (1, 2, 1, 1, 3, 1), (1, 1, 3, 2, 1, 1)):
        a, lda, x, y, lenx, leny = _problem(rng, dt, layout, trans, m, n, incx, incy)
        want, mag = _want(a, layout, trans, m, n, lda, alpha, x, incx, beta, y, incy, lenx, leny)
        yg, yo = y.copy(), y.copy()
        level2.gemv(layout, trans, m, n, alpha, a, lda, x, incx, beta, yg, incy)
        oracle.gemv(layout, trans, m, n, alpha, a, lda, x, incx, beta, yo, incy)
        idx = np.arange(leny) * incy
        for what, got in (("gpu", yg), ("oracle", yo)):
            err = np.abs(got.astype(np.complex128)[idx] - want)
            assert (err <= c * (lenx + 2) * _u(dt) * mag + 1e-300).all(), (what, m, n)
        gaps = np.ones(y.size, bool)
        gaps[idx] = False
        assert np.array_equal(yg[gaps], y[gaps])


@pytest.mark.gpu
@pytest.mark.parametrize("dt", REAL + CPLX)
def test_gemv_exact_integers(gpu, dt):
    """Small-integer operands: every sum is exact, so any order gives the
    float64 result bit for bit."""
    rng = np.random.default_rng(5)
    cplx = np.dtype(dt).kind == "c"
    for layout, trans in itertools.product((ROW, COL), (0, 1, 2) if cplx else (0, 1)):
        for m, n in ((33, 2000), (3000, 17), (513, 257)):
            lda = (n if layout == ROW else m) + 1
            size = (m if layout == ROW else n) * lda
            a = rng.integers(-3, 4, size).astype(dt)
            if cplx:
                a = (a + 1j * rng.integers(-3, 4, size)).astype(dt)
            lenx, leny = (m, n) if trans else (n, m)
            x = rng.integers(-3, 4, lenx).astype(dt)
            y = rng.integers(-3, 4, leny).astype(dt)
            alpha, beta = (2 - 1j, 1 + 1j) if cplx else (2.0, -1.0)
            want, _ = _want(a, layout, trans, m, n, lda, alpha, x, 1, beta, y, 1, lenx, leny)
            level2.gemv(layout, trans, m, n, alpha, a, lda, x, 1, beta, y, 1)
            assert np.array_equal(y.astype(np.complex128), want), (layout, trans, m, n)


@pytest.mark.gpu
def test_gemv_special_cases(gpu, oracle):
    nan, inf = np.nan, np.inf
    a = np.full(4, nan)
    y = np.array([nan, -0.0])
    level2.gemv(COL, 0, 2, 2, 0.0, a, 2, np.ones(2), 1, 1.0, y, 1)
    assert np.isnan(y[0]) and np.signbit(y[1])
    level2.gemv(COL, 1, 2, 2, 0.0, a, 2, np.ones(2), 1, 0.0, y, 1)
    assert list(y) == [0.0, 0.0] and not np.signbit(y).any()
    y = np.array([nan, 1.0])
    level2.gemv(COL, 0, 2, 2, 1.0, np.ones(4), 2, np.ones(2), 1, 0.0, y, 1)
    assert list(y) == [2.0, 2.0]
    for beta in (1, 0, 0.5 + 0.5j):
        for alpha in (0, 1j):
            y = np.array([inf + 0j, 2 - 1j, nan + 0j])
            yo = y.copy()
            av = np.arange(6, dtype=np.complex128) if alpha else np.full(6, nan + 0j)
            level2.gemv(COL, 2, 2, 3, alpha, av, 2, np.ones(2, complex), 1, beta, y, 1)
            oracle.gemv(COL, 2, 2, 3, alpha, av, 2, np.ones(2, complex), 1, beta, yo, 1)
            assert np.array_equal(np.isnan(y.real), np.isnan(yo.real))
            assert np.array_equal(np.isnan(y.imag), np.isnan(yo.imag))
            fin = np.isfinite(yo)
            assert np.allclose(y[fin], yo[fin], rtol=1e-15, atol=0)
    # incx == 0 broadcasts x[0]; incy == 0 is fine for one output.
    a = np.arange(12, dtype=np.float64)
    y = np.zeros(3)
    level2.gemv(COL, 0, 3, 4, 1.0, a, 3, np.array([2.0]), 0, 0.0, y, 1)
    assert list(y) == list(2 * _mat(a, COL, 3, 4, 3).sum(axis=1))
    y = np.array([1.0])
    level2.gemv(COL, 1, 3, 1, 1.0, a, 3, np.ones(3), 1, 1.0, y, 0)
    assert y[0] == 1.0 + 0 + 1 + 2


@pytest.mark.gpu
@pytest.mark.parametrize("name,dt", [("ger", np.float32), ("ger", np.float64),
                                     ("geru", np.complex64), ("geru", np.complex128),
                                     ("gerc", np.complex64), ("gerc", np.complex128)])
def test_rank1_bit_exact_vs_oracle(gpu, oracle, name, dt):
    rng = np.random.default_rng(8)
    cplx = np.dtype(dt).kind == "c"
    alpha = 1.25 - 0.75j if cplx else -1.25
    for layout, (m, n), incx, incy in itertools.product(
            (ROW, COL), ((5, 3), (1, 700), (700, 1), (1500, 1300), (64, 20000)), (1, 2), (1, 3)):
        lda = (n if layout == ROW else m) + 5
        a = _rand(rng, dt, (m if layout == ROW else n) * lda)
        x = _rand(rng, dt, (m - 1) * incx + 1)
        y = _rand(rng, dt, (n - 1) * incy + 1)
        ag, ao = a.copy(), a.copy()
        getattr(level2, name)(layout, m, n, alpha, x, incx, y, incy, ag, lda)
        getattr(oracle, name)(layout, m, n, alpha, x, incx, y, incy, ao, lda)
        assert np.array_equal(ag.view(np.uint8), ao.view(np.uint8)), (layout, m, n, incx, incy)


@pytest.mark.gpu
def test_device_entries_match_host(gpu):
    rng = np.random.default_rng(21)
    s = Stream()
    for dt, layout, trans in ((np.float64, COL, 1), (np.float32, ROW, 0), (np.complex128, ROW, 2),
                              (np.complex64, COL, 0)):
        m, n = 900, 1700
        a, lda, x, y, _, _ = _problem(rng, dt, layout, trans, m, n, 1, 1)
        alpha, beta = (1.5, 0.5) if np.dtype(dt).kind == "f" else (1.5 - 1j, 0.5j)
        yh = y.copy()
        level2.gemv(layout, trans, m, n, alpha, a, lda, x, 1, beta, yh, 1)
        da, dx, dy = (DeviceArray.from_host(v, s) for v in (a, x, y))
        level2.gemv_dev(layout, trans, m, n, alpha, da, lda, dx, 1, beta, dy, 1, s)
        s.synchronize()
        assert np.array_equal(dy.download(s), yh)
        for v in (da, dx, dy):
            v.free()
    for name, dt in (("ger", np.float64), ("gerc", np.complex64)):
        m, n, lda = 800, 600, 803
        a = _rand(rng, dt, n * lda)
        x, y = _rand(rng, dt, m), _rand(rng, dt, n)
        ah = a.copy()
        getattr(level2, name)(COL, m, n, 0.5, x, 1, y, 1, ah, lda)
        da, dx, dy = (DeviceArray.from_host(v, s) for v in (a, x, y))
        getattr(level2, name + "_dev")(COL, m, n, 0.5, dx, 1, dy, 1, da, lda, s)
        s.synchronize()
        assert np.array_equal(da.download(s), ah)
        for v in (da, dx, dy):
            v.free()
    s.close()
