"""The C ABI without a GPU: the library loads, exports every symbol the
header declares, and reproduces rblas's error contract (try_gemm's exact
BlasError values and gemm's panic conditions) before touching the device."""
import itertools
import os
import re

import numpy as np
import pytest

from rblas_amd import _ffi, complex3, grouped, level3
from rblas_amd.types import BlasError, BlasPanic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(rblas_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    syms = header_symbols()
    assert len(syms) >= 10
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert set(syms) <= set(_ffi.SIGNATURES), set(syms) - set(_ffi.SIGNATURES)


def test_try_gemm_reference_errors():
    # level3.rs:566-640
    m, n, k = 4, 3, 5
    a = np.ones(m * k)
    b = np.ones(k * n)
    c = np.zeros(m * n)
    with pytest.raises(BlasError) as e:
        level3.try_gemm(1, 0, 0, m, n, k, 1.0, a, 3, b, k, 0.0, c, m)
    assert e.value == BlasError("InvalidLeadingDim", "a", 3, 4)
    small = np.zeros(m * n - 1)
    with pytest.raises(BlasError) as e:
        level3.try_gemm(1, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, small, m)
    assert e.value.kind == "BufferTooSmall" and e.value.which == "c"
    assert e.value.got == m * n - 1 and e.value.bound == m * n


def _sweep():
    rng = np.random.default_rng(1234)
    for _ in range(400):
        layout = int(rng.integers(0, 2))
        ta, tb = (int(x) for x in rng.integers(0, 3, 2))
        m, n, k = (int(x) for x in rng.integers(0, 6, 3))
        lda, ldb, ldc = (int(x) for x in rng.integers(0, 8, 3))
        a_len, b_len, c_len = (int(x) for x in rng.integers(0, 40, 3))
        yield layout, ta, tb, m, n, k, lda, ldb, ldc, a_len, b_len, c_len


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_try_gemm_errors_match_oracle(oracle, dtype):
    """Every failing try_gemm returns exactly the oracle's BlasError; no
    device is needed because validation precedes any device work."""
    checked = 0
    for layout, ta, tb, m, n, k, lda, ldb, ldc, al, bl, cl in _sweep():
        want = oracle.try_gemm_check(layout, ta, tb, m, n, k, al, lda, bl, ldb, cl, ldc)
        if want is None:
            continue
        a, b, c = np.zeros(al, dtype), np.zeros(bl, dtype), np.zeros(cl, dtype)
        with pytest.raises(BlasError) as e:
            level3.try_gemm(layout, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)
        assert (e.value.kind, e.value.which, e.value.got, e.value.bound) == want
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gemm_panics_match_oracle(oracle, dtype):
    checked = 0
    for layout, ta, tb, m, n, k, lda, ldb, ldc, al, bl, cl in _sweep():
        for alpha, beta in ((1.0, 0.0), (0.0, 1.0)):
            if not oracle.gemm_panics(layout, ta, tb, m, n, k, al, lda, bl, ldb, cl, ldc,
                                      alpha, beta):
                continue
            a, b, c = np.zeros(al, dtype), np.zeros(bl, dtype), np.zeros(cl, dtype)
            with pytest.raises(BlasPanic):
                level3.gemm(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
            checked += 1
    assert checked > 100


def test_bad_enums_rejected():
    a = np.zeros(4)
    for layout, ta, tb in ((2, 0, 0), (1, 3, 0), (1, 0, -1)):
        with pytest.raises(ValueError):
            level3.gemm(layout, ta, tb, 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a, 2)


def test_gemm_zero_dims_need_no_device():
    # m == 0 or n == 0 returns before any device work (kernel/gemm.rs:198-200)
    for layout, ta, tb in itertools.product((0, 1), (0, 1), (0, 1)):
        c = np.full(6, 7.0)
        level3.gemm(layout, ta, tb, 0, 3, 2, 1.0, np.zeros(64), 3, np.zeros(64), 3, 0.0, c, 3)
        assert (c == 7).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gemm_overlapping_c_columns_refused(dtype):
    """rblas updates C's columns one after another, so with ldc < m (column
    major) or ldc < n (row major) overlapping columns see earlier columns'
    results; a device cannot follow that order, so the call is refused with
    RBLAS_BAD_ARG before any device work. A call that leaves C untouched
    (beta 1 and alpha 0) stays a no-op."""
    m, n, k = 6, 4, 3
    a, b = np.ones(64, dtype), np.ones(64, dtype)
    for layout, ldc in ((1, m - 1), (0, n - 1)):
        lda = m if layout == 1 else k
        ldb = k if layout == 1 else n
        c = np.zeros(64, dtype)
        for alpha, beta in ((1.0, 0.0), (0.0, 0.5), (1.0, 1.0)):
            with pytest.raises(ValueError, match="overlapping columns"):
                level3.gemm(layout, 0, 0, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        level3.gemm(layout, 0, 0, m, n, k, 0.0, a, lda, b, ldb, 1.0, c, ldc)
        assert (c == 0).all()
    # one column (n == 1) cannot overlap
    with pytest.raises(Exception) as e:
        level3.gemm(1, 0, 0, m, 1, k, 1.0, a, m, b, k, 0.0, np.zeros(64, dtype), 1)
    assert "overlapping" not in str(e.value)


def test_complex_gemm_overlapping_c_columns_refused():
    m, n, k = 5, 3, 2
    for dt in (np.complex64, np.complex128):
        a, b, c = np.ones(32, dt), np.ones(32, dt), np.zeros(32, dt)
        with pytest.raises(ValueError, match="overlapping columns"):
            complex3.gemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m - 2)


def test_gemm_large_ldc_refused_before_device():
    """A 128-row tile of C must span under 2 GiB for the epilogue's 31-bit
    offsets: ldc * 128 * 8 >= 2^31 is refused with RBLAS_BAD_ARG (not a HIP
    error) before any device work."""
    m, n, k = 4, 2, 3
    ldc = (1 << 31) // (128 * 8) + 1
    c = np.zeros(ldc + m)
    with pytest.raises(ValueError, match="ldc"):
        level3.gemm(1, 0, 0, m, n, k, 1.0, np.ones(m * k), m, np.ones(k * n), k, 0.0, c, ldc)


def test_grouped_overlapping_c_columns_refused():
    class _Host:
        def __init__(self, size):
            self.ptr, self.size, self.dtype = 0, size, np.dtype(np.float64)

    ok = grouped.GemmProblem(1, 0, 0, 4, 4, 4, 1.0, _Host(16), 4, _Host(16), 4, 0.0, _Host(16), 4)
    over = grouped.GemmProblem(1, 0, 0, 4, 4, 4, 1.0, _Host(16), 4, _Host(16), 4, 0.0,
                               _Host(16), 3)
    with pytest.raises(ValueError) as e:
        grouped.GroupedGemm([ok, over], np.float64)
    assert "problem 1" in str(e.value) and "overlapping" in str(e.value)
