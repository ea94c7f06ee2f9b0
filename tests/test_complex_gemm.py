"""Complex gemm (4M) through the C ABI.

Tolerance (elementwise, written here): |dC| <= 8 (k+2) u (|alpha| |A||B| +
|beta| |C0|), u the real unit roundoff: each 4M entry is two length-2k real
dot products plus the complex alpha/beta update.
"""
import itertools

import numpy as np
import pytest

from rblas_amd import complex3
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic
from tests.util import COL, op, view


def make(ta, tb, m, n, k, ctype, seed, pad=3):
    rng = np.random.default_rng(seed)
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda, ldb, ldc = ar + pad, br + pad, m + pad
    def rnd(sz):
        return (rng.standard_normal(sz) + 1j * rng.standard_normal(sz)).astype(ctype)
    return rnd((ac - 1) * lda + ar), lda, rnd((bc - 1) * ldb + br), ldb, rnd((n - 1) * ldc + m), ldc


def expect(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc):
    A = op(a.astype(np.complex128), COL, ta, m, k, lda)
    B = op(b.astype(np.complex128), COL, tb, k, n, ldb)
    C0 = view(c0.astype(np.complex128), COL, m, n, ldc)
    want = alpha * (A @ B) + (0 if beta == 0 else beta * C0)
    mag = abs(alpha) * (np.abs(A) @ np.abs(B)) + abs(beta) * np.abs(C0)
    return want, mag


def test_cgemm_panics_without_device():
    # split_op indexes A past its slice (complex3.rs:111-118)
    a = np.zeros(3, np.complex128)
    b = np.zeros(16, np.complex128)
    c = np.zeros(16, np.complex128)
    with pytest.raises(BlasPanic):
        complex3.gemm(0, 0, 2, 2, 2, 1.0, a, 2, b, 2, 0.0, c, 2)
    with pytest.raises(BlasPanic):  # C too short for the final update
        complex3.gemm(0, 0, 2, 2, 2, 1.0, b, 2, b, 2, 0.0, c[:3].copy(), 2)


@pytest.mark.gpu
def test_cgemm_reference_cases(gpu):
    # complex3.rs:372-401: 5x4x6 vs naive
    m, n, k = 5, 4, 6
    i = np.arange(m * k)
    a = np.sin(i * 0.3) + 1j * np.cos(i * 0.7)
    j = np.arange(k * n)
    b = np.cos(j * 0.2) + 1j * np.sin(j * 0.5)
    c = np.zeros(m * n, np.complex128)
    complex3.gemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = (a.reshape(k, m).T @ b.reshape(n, k).T).T.reshape(-1)
    assert np.abs(c - want).max() < 1e-9
    # complex3.rs:403-437: C = A^H B
    m, n, k = 3, 2, 4
    a = np.arange(k * m) + 0.5j * np.arange(k * m)
    b = 1.0 + 1j * np.arange(k * n)
    c = np.zeros(m * n, np.complex128)
    complex3.gemm(2, 0, m, n, k, 1.0, a, k, b, k, 0.0, c, m)
    want = (np.conj(a.reshape(m, k)) @ b.reshape(n, k).T).T.reshape(-1)
    assert np.abs(c - want).max() < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("ctype", [np.complex64, np.complex128])
@pytest.mark.parametrize("ta,tb", list(itertools.product([0, 1, 2], [0, 1, 2])))
def test_cgemm_vs_oracle(gpu, oracle, ctype, ta, tb):
    m, n, k = 133, 71, 259
    a, lda, b, ldb, c, ldc = make(ta, tb, m, n, k, ctype, 10 + 3 * ta + tb)
    c0, c_ref = c.copy(), c.copy()
    alpha, beta = 1.5 - 0.5j, -0.25 + 0.75j
    complex3.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    oracle.cgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c_ref, ldc)
    want, mag = expect(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc)
    u = 2.0 ** -24 if ctype == np.complex64 else 2.0 ** -53
    tol = 8 * (k + 2) * u * mag + 1e-300
    assert (np.abs(view(c, COL, m, n, ldc) - want) <= tol).all()
    assert (np.abs(view(c_ref, COL, m, n, ldc) - want) <= tol).all()
    mask = np.ones(c.size, bool)
    view(mask, COL, m, n, ldc)[:] = False
    assert np.array_equal(c[mask], c0[mask])


@pytest.mark.gpu
def test_cgemm_special_cases(gpu):
    m, n, k = 9, 7, 5
    a, lda, b, ldb, c, ldc = make(0, 0, m, n, k, np.complex128, 3, pad=0)
    c[:] = np.nan
    complex3.gemm(0, 0, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc)  # beta 0 never reads C
    assert np.isfinite(c).all()
    c2 = np.full(m * n, 2 + 1j)
    complex3.gemm(0, 0, m, n, 0, 1.0, a, lda, b, ldb, 0.5j, c2, m)  # k == 0: beta*C
    assert np.allclose(c2, 0.5j * (2 + 1j))


@pytest.mark.gpu
def test_cgemm_device_entry(gpu):
    m, n, k = 300, 257, 130
    a, lda, b, ldb, c, ldc = make(1, 2, m, n, k, np.complex128, 8)
    ch = c.copy()
    complex3.gemm(1, 2, m, n, k, 0.5 + 2j, a, lda, b, ldb, 1.0, ch, ldc)
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    complex3.gemm_dev(1, 2, m, n, k, 0.5 + 2j, da, lda, db, ldb, 1.0, dc, ldc, s)
    s.synchronize()
    assert np.array_equal(dc.download(s), ch)
    s.close()
