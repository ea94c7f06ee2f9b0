"""syrk and trsm through the C ABI.

CPU part: where rblas panics (oracle.syrk_panics / trsm_panics walk the
reference loops index by index) against the library's closed-form checks.

GPU part, tolerances written here:
  syrk: the gemm bound on the triangle, |dC| <= 2 (k+2) u (|alpha| |op(A)||op(A)^T|
        + |beta| |C0|); outside the triangle C is untouched, bit for bit.
  trsm: componentwise residual |op(A) X - alpha B| <= 16 q u (|op(A)||X| + |alpha B|)
        (q the order of A), for the GPU and for the oracle alike, on matrices
        of the batching workload's family (diag 2+|N(0,1)|, off-diagonal
        N(0,1)/sqrt(q)); exact-integer cases are bit-exact.
"""
import itertools

import numpy as np
import pytest

from rblas_amd import level3
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic, DeviceError
from tests.util import COL, assert_gemm_close, view


def _call_status(fn):
    """'panic', 'bad' (BAD_ARG/BAD_ENUM), or 'ran' (OK, or no device here)."""
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
def test_syrk_panics_match_oracle(oracle):
    checked = panics = 0
    for uplo, trans, n, k, lda, ldc in itertools.product((0, 1), (0, 1), (0, 1, 3), (0, 1, 2),
                                                         (1, 3, 4), (2, 3, 4)):
        rows = k if trans else n
        for a_len in sorted({0, max(0, rows * lda - 2), rows * lda + 3}):
            for c_len in sorted({0, (n - 1) * (ldc + 1) if n else 0, n * ldc}):
                want = oracle.syrk_panics(uplo, trans, n, k, a_len, lda, c_len, ldc)
                a, c = np.zeros(a_len), np.zeros(c_len)
                got = _call_status(lambda: level3.syrk(uplo, trans, n, k, 1.0, a, lda, 0.5, c,
                                                       ldc))
                assert (got == "panic") == want, (uplo, trans, n, k, lda, ldc, a_len, c_len, got)
                checked += 1
                panics += want
    assert checked > 300 and panics > 100


def test_trsm_panics_match_oracle(oracle):
    checked = panics = 0
    for side, uplo, trans, diag in itertools.product((0, 1), (0, 1), (0, 1), (0, 1)):
        for m, n, lda, ldb, alpha in itertools.product((0, 1, 3), (0, 2, 3), (1, 3, 4), (3, 5),
                                                       (1.0, 2.0)):
            q = m if side == 0 else n
            for a_len in sorted({0, max(0, (q - 1) * (lda + 1)), q * lda + 1}):
                for b_len in sorted({0, max(0, (m - 1) + (n - 1) * ldb), n * ldb}):
                    want = oracle.trsm_panics(side, uplo, trans, diag, m, n, alpha, a_len, lda,
                                              b_len, ldb)
                    a, b = np.ones(a_len), np.ones(b_len)
                    got = _call_status(lambda: level3.trsm(side, uplo, trans, diag, m, n, alpha,
                                                           a, lda, b, ldb))
                    assert (got == "panic") == want, (side, uplo, trans, diag, m, n, lda, ldb,
                                                      a_len, b_len, got)
                    checked += 1
                    panics += want
    assert checked > 1000 and panics > 300


def test_syrk_trsm_bad_enums_and_empty():
    a = np.zeros(16)
    for uplo, trans in ((2, 0), (0, 3), (-1, 0)):
        with pytest.raises(ValueError):
            level3.syrk(uplo, trans, 2, 2, 1.0, a, 2, 0.0, a, 2)
    for args in ((2, 0, 0, 0), (0, 2, 0, 0), (0, 0, 3, 0), (0, 0, 0, 2)):
        with pytest.raises(ValueError):
            level3.trsm(*args, 2, 2, 1.0, a, 2, a, 2)
    # n == 0 (syrk), m == 0 or n == 0 (trsm): return before any device work.
    c = np.full(4, 3.0)
    level3.syrk(0, 0, 0, 5, 1.0, np.zeros(0), 1, 0.0, c, 2)
    level3.trsm(0, 1, 0, 0, 0, 4, 2.0, np.zeros(0), 1, c, 1)
    level3.trsm(1, 1, 0, 0, 4, 0, 2.0, np.zeros(0), 1, c, 4)
    assert (c == 3).all()


def test_overlapping_outputs_rejected():
    # ldc < n (syrk) / ldb < m (trsm) with several columns would overlap the
    # columns the device writes in parallel: BAD_ARG, not a wrong answer.
    with pytest.raises(ValueError):
        level3.syrk(1, 0, 3, 2, 1.0, np.ones(6), 3, 0.0, np.zeros(9), 2)
    with pytest.raises(ValueError):
        level3.trsm(0, 1, 0, 0, 3, 2, 1.0, np.eye(3).reshape(-1), 3, np.ones(6), 2)


# ------------------------------------------------------------------ GPU side
def _tri_mask(n, uplo):
    i, j = np.indices((n, n))
    return i <= j if uplo == 0 else i >= j


@pytest.mark.gpu
def test_syrk_reference_kats(gpu):
    # level3.rs:739-766: lower, no transpose, A 3x2 = [1..6], tol 1e-12.
    n, k = 3, 2
    a = np.arange(1.0, 7.0)
    c = np.zeros(n * n)
    level3.syrk(1, 0, n, k, 1.0, a, n, 0.0, c, n)
    A = a.reshape(k, n).T
    full = A @ A.T
    got = c.reshape(n, n).T
    assert np.abs(got - full)[_tri_mask(n, 1)].max() < 1e-12
    assert (got[~_tri_mask(n, 1)] == 0).all()
    # level3.rs:795-824: upper, Trans, 9x11, alpha 2, beta 3, tol 1e-9.
    n, k = 9, 11
    a = np.sin(np.arange(k * n) * 0.31)
    c = np.arange(n * n) * 0.1
    c0 = c.copy()
    level3.syrk(0, 1, n, k, 2.0, a, k, 3.0, c, n)
    A = a.reshape(n, k)          # A stored k x n column-major -> row i is column i
    want = 2.0 * (A @ A.T) + 3.0 * c0.reshape(n, n).T
    got = c.reshape(n, n).T
    assert np.abs(got - want)[_tri_mask(n, 0)].max() < 1e-9
    assert np.array_equal(got[~_tri_mask(n, 0)], c0.reshape(n, n).T[~_tri_mask(n, 0)])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("uplo,trans", list(itertools.product((0, 1), (0, 1, 2))))
def test_syrk_vs_oracle(gpu, oracle, dtype, uplo, trans):
    n, k, pad = 300, 77, 5
    rng = np.random.default_rng(7 + 3 * uplo + trans)
    rows, cols = (k, n) if trans else (n, k)
    lda, ldc = rows + pad, n + pad
    a = rng.standard_normal((cols - 1) * lda + rows).astype(dtype)
    c = rng.standard_normal((n - 1) * ldc + n).astype(dtype)
    c0, c_ref = c.copy(), c.copy()
    level3.syrk(uplo, trans, n, k, 1.5, a, lda, -0.25, c, ldc)
    oracle.syrk(uplo, min(trans, 1), n, k, 1.5, a, lda, -0.25, c_ref, ldc)
    A = view(a, COL, rows, cols, lda).astype(np.float64)
    opA = A.T if trans else A
    C0 = view(c0, COL, n, n, ldc).astype(np.float64)
    want = 1.5 * (opA @ opA.T) - 0.25 * C0
    mag = 1.5 * (np.abs(opA) @ np.abs(opA.T)) + 0.25 * np.abs(C0)
    tri = _tri_mask(n, uplo)
    g, r = view(c, COL, n, n, ldc), view(c_ref, COL, n, n, ldc)
    assert_gemm_close(np.where(tri, g, want), want, mag, k, dtype, what="gpu syrk")
    assert_gemm_close(np.where(tri, r, want), want, mag, k, dtype, what="oracle syrk")
    # The other triangle and the padding rows stay bit-identical.
    keep = np.ones(c.size, bool)
    view(keep, COL, n, n, ldc)[:] = ~tri
    assert np.array_equal(c[keep], c0[keep])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_syrk_exact_and_special(gpu, dtype):
    n, k = 261, 130
    rng = np.random.default_rng(3)
    a = rng.integers(-3, 4, n * k).astype(dtype)
    for uplo in (0, 1):
        c = rng.integers(-8, 9, n * n).astype(dtype)
        c0 = c.copy()
        level3.syrk(uplo, 0, n, k, 2.0, a, n, 0.5, c, n)
        A = a.reshape(k, n).T.astype(np.float64)
        want = 2.0 * (A @ A.T) + 0.5 * c0.reshape(n, n).T
        tri = _tri_mask(n, uplo)
        got = c.reshape(n, n).T
        assert np.array_equal(got[tri], want[tri])
        # beta == 0: the triangle is stored without reading C (NaN cleared)
        c = np.full(n * n, np.nan, dtype)
        level3.syrk(uplo, 0, n, k, 1.0, a, n, 0.0, c, n)
        got = c.reshape(n, n).T
        assert np.isfinite(got[tri]).all() and np.isnan(got[~tri]).all()
        # k == 0 and alpha == 0: C_tri = beta * C_tri (0 stores zeros)
        for kk, alpha, beta in ((0, 1.0, 0.5), (k, 0.0, -2.0), (0, 1.0, 0.0)):
            c = rng.integers(-8, 9, n * n).astype(dtype)
            c[5] = np.nan
            c0 = c.copy()
            level3.syrk(uplo, 1, n, kk, alpha, a, max(kk, 1), beta, c, n)
            got, ref = c.reshape(n, n).T, c0.reshape(n, n).T
            exp = np.zeros_like(ref) if beta == 0 else beta * ref
            assert np.array_equal(got[tri], exp[tri], equal_nan=True)
            assert np.array_equal(got[~tri], ref[~tri], equal_nan=True)


@pytest.mark.gpu
def test_syrk_device_entry(gpu):
    n, k = 517, 260
    rng = np.random.default_rng(11)
    a = rng.standard_normal(n * k)
    c = rng.standard_normal(n * n)
    ch = c.copy()
    level3.syrk(1, 0, n, k, 0.5, a, n, 2.0, ch, n)
    s = Stream()
    da, dc = DeviceArray.from_host(a, s), DeviceArray.from_host(c, s)
    level3.syrk_dev(1, 0, n, k, 0.5, da, n, 2.0, dc, n, s)
    s.synchronize()
    assert np.array_equal(dc.download(s), ch)
    s.close()


def _tri_system(q, uplo, trans, diag, lda, dtype, rng):
    """Stored q x q A (ld lda) whose op(A) is triangular with the batching
    workload's entries: diag 2+|N(0,1)|, off-diagonal N(0,1)/sqrt(q). The
    unused triangle holds garbage that must not be read."""
    full = rng.standard_normal((q, q)) / np.sqrt(q)
    np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
    stored = np.triu(full) if uplo == 0 else np.tril(full)
    junk = 1e3 * (~(np.triu(np.ones((q, q), bool)) if uplo == 0 else
                    np.tril(np.ones((q, q), bool))))
    stored = stored + junk
    if diag == 1:
        np.fill_diagonal(stored, 7e3)  # the stored diagonal is not read
    a = np.full((q - 1) * lda + q, -5e2)
    view(a, COL, q, q, lda)[:] = stored
    a = a.astype(dtype)
    S = view(a, COL, q, q, lda).astype(np.float64)
    S = np.triu(S) if uplo == 0 else np.tril(S)
    if diag == 1:
        np.fill_diagonal(S, 1.0)
    return a, (S.T if trans else S)


def _residual_ok(side, opA, X, alpha, B0, q, dtype):
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
    if side == 0:
        r = np.abs(opA @ X - alpha * B0)
        bound = 16 * q * u * (np.abs(opA) @ np.abs(X) + np.abs(alpha * B0))
    else:
        r = np.abs(X @ opA - alpha * B0)
        bound = 16 * q * u * (np.abs(X) @ np.abs(opA) + np.abs(alpha * B0))
    return float((r / np.maximum(bound, 1e-300)).max())


@pytest.mark.gpu
def test_trsm_reference_kats(gpu):
    # level3.rs:867-901: L X = B with x = (1, 2, 3), tol 1e-12.
    l = np.array([2.0, 1.0, 1.0, 0.0, 3.0, 2.0, 0.0, 0.0, 4.0])
    x = np.array([1.0, 2.0, 3.0])
    b = l.reshape(3, 3).T @ x
    level3.trsm(0, 1, 0, 0, 3, 1, 1.0, l, 3, b, 3)
    assert np.abs(b - x).max() < 1e-12
    # level3.rs:903-934: X U = B round trip, tol 1e-10.
    u = np.array([2.0, 0.0, 0.0, 1.0, 3.0, 0.0, 4.0, 5.0, 6.0])
    b0 = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0])
    b = b0.copy()
    level3.trsm(1, 0, 0, 0, 2, 3, 1.0, u, 3, b, 2)
    X, U = b.reshape(3, 2).T, u.reshape(3, 3).T
    assert np.abs(X @ U - b0.reshape(3, 2).T).max() < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("side,uplo,trans,diag",
                         list(itertools.product((0, 1), (0, 1), (0, 1), (0, 1))))
def test_trsm_vs_oracle(gpu, oracle, dtype, side, uplo, trans, diag):
    m, n = (200, 75) if side == 0 else (75, 200)
    q = m if side == 0 else n
    rng = np.random.default_rng(100 + 8 * side + 4 * uplo + 2 * trans + diag)
    lda, ldb = q + 3, m + 4
    a, opA = _tri_system(q, uplo, trans, diag, lda, dtype, rng)
    b = rng.standard_normal((n - 1) * ldb + m).astype(dtype)
    b0, b_ref = b.copy(), b.copy()
    alpha = 0.75
    level3.trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb)
    oracle.trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b_ref, ldb)
    B0 = view(b0, COL, m, n, ldb).astype(np.float64)
    for what, got in (("gpu", b), ("oracle", b_ref)):
        X = view(got, COL, m, n, ldb).astype(np.float64)
        ratio = _residual_ok(side, opA, X, alpha, B0, q, dtype)
        assert ratio <= 1.0, f"{what} residual over bound ({ratio:.3g})"
    keep = np.ones(b.size, bool)
    view(keep, COL, m, n, ldb)[:] = False
    assert np.array_equal(b[keep], b0[keep])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("side,uplo,trans", list(itertools.product((0, 1), (0, 1), (0, 1))))
def test_trsm_exact_integers(gpu, dtype, side, uplo, trans):
    """Bidiagonal op(A) with entries in {-1, 0, 1} and diagonal +-1: the
    inverse of every diagonal block holds only -1, 0, 1, so every step is
    exact and X must equal the integer solution bit for bit."""
    m, n = (150, 67) if side == 0 else (67, 150)
    q = m if side == 0 else n
    rng = np.random.default_rng(5 + 4 * side + 2 * uplo + trans)
    S = np.diag(rng.choice([-1.0, 1.0], q))
    off = rng.integers(-1, 2, q - 1).astype(np.float64)
    eff_upper = (uplo == 0) != bool(trans)
    opA = S + (np.diag(off, 1) if eff_upper else np.diag(off, -1))
    stored = opA.T if trans else opA
    X = rng.integers(-4, 5, (m, n)).astype(np.float64)
    B = opA @ X if side == 0 else X @ opA
    a = stored.T.reshape(-1).astype(dtype)
    b = B.T.reshape(-1).astype(dtype)
    level3.trsm(side, uplo, trans, 0, m, n, 1.0, a, q, b, m)
    assert np.array_equal(b.reshape(n, m).T, X)


@pytest.mark.gpu
def test_trsm_special_cases(gpu, oracle):
    rng = np.random.default_rng(9)
    for q in (1, 63, 64, 65, 129):
        a, opA = _tri_system(q, 1, 0, 0, q, np.float64, rng)
        b = rng.standard_normal(q * 3)
        b_ref = b.copy()
        level3.trsm(0, 1, 0, 0, q, 3, -2.0, a, q, b, q)
        oracle.trsm(0, 1, 0, 0, q, 3, -2.0, a, q, b_ref, q)
        assert np.allclose(b, b_ref, rtol=1e-12, atol=1e-14)
    # alpha == 0: B becomes 0 * B (NaN stays NaN), then X = 0.
    a, _ = _tri_system(70, 0, 0, 0, 70, np.float64, rng)
    b = rng.standard_normal(70 * 2)
    b[3] = np.nan
    level3.trsm(0, 0, 0, 0, 70, 2, 0.0, a, 70, b, 70)
    b_ref = rng.standard_normal(70 * 2)
    b_ref[3] = np.nan
    oracle.trsm(0, 0, 0, 0, 70, 2, 0.0, a, 70, b_ref, 70)
    assert np.array_equal(np.isnan(b), np.isnan(b_ref))
    assert (b[~np.isnan(b)] == 0).all()
    # One column with ldb < m is allowed (no overlap).
    a, opA = _tri_system(5, 1, 0, 0, 5, np.float64, rng)
    b = rng.standard_normal(5)
    b0 = b.copy()
    level3.trsm(0, 1, 0, 0, 5, 1, 1.0, a, 5, b, 2)
    assert np.allclose(opA @ b, b0, rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("side", [0, 1])
def test_trsm_device_entry_and_large(gpu, side):
    m, n = (1000, 300) if side == 0 else (300, 1000)
    q = m if side == 0 else n
    rng = np.random.default_rng(21 + side)
    a, opA = _tri_system(q, 1, 1, 0, q, np.float64, rng)
    b = rng.standard_normal(m * n)
    bh = b.copy()
    level3.trsm(side, 1, 1, 0, m, n, 1.25, a, q, bh, m)
    X = bh.reshape(n, m).T
    assert _residual_ok(side, opA, X, 1.25, b.reshape(n, m).T, q, np.float64) <= 1.0
    s = Stream()
    da, db = DeviceArray.from_host(a, s), DeviceArray.from_host(b, s)
    level3.trsm_dev(side, 1, 1, 0, m, n, 1.25, da, q, db, m, s)
    s.synchronize()
    assert np.array_equal(db.download(s), bh)
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("q", [1024, 1025, 2048, 2049, 2600])
@pytest.mark.parametrize("side,uplo,trans", [(0, 1, 0), (0, 0, 1), (0, 0, 0), (0, 1, 1),
                                             (1, 0, 0), (1, 1, 1), (1, 1, 0), (1, 0, 1)])
def test_trsm_order_switches(gpu, q, side, uplo, trans):
    """Orders on both sides of the 16/8 right-hand-side width switch (1024)
    and of the LDS-resident/recursive switch (2048), f32 and f64, forward
    (op(A) lower on the Left) and backward substitution on both sides."""
    nrhs = 37
    m, n = (q, nrhs) if side == 0 else (nrhs, q)
    for dtype in (np.float32, np.float64):
        rng = np.random.default_rng(q + 10 * side + 3 * uplo + trans)
        lda, ldb = q + 1, m + 2
        a, opA = _tri_system(q, uplo, trans, 0, lda, dtype, rng)
        b = rng.standard_normal((n - 1) * ldb + m).astype(dtype)
        b0 = b.copy()
        level3.trsm(side, uplo, trans, 0, m, n, 0.5, a, lda, b, ldb)
        X = view(b, COL, m, n, ldb).astype(np.float64)
        B0 = view(b0, COL, m, n, ldb).astype(np.float64)
        ratio = _residual_ok(side, opA, X, 0.5, B0, q, dtype)
        assert ratio <= 1.0, f"{dtype.__name__} residual over bound ({ratio:.3g})"
        keep = np.ones(b.size, bool)
        view(keep, COL, m, n, ldb)[:] = False
        assert np.array_equal(b[keep], b0[keep])


@pytest.mark.gpu
@pytest.mark.parametrize("q", [2048, 2200])
@pytest.mark.parametrize("side", [0, 1])
def test_trsm_exact_integers_large(gpu, q, side):
    """The bidiagonal exact-integer system at orders taken by each path."""
    nrhs = 20
    m, n = (q, nrhs) if side == 0 else (nrhs, q)
    rng = np.random.default_rng(77 + q + side)
    S = np.diag(rng.choice([-1.0, 1.0], q))
    off = rng.integers(-1, 2, q - 1).astype(np.float64)
    opA = S + np.diag(off, -1)
    X = rng.integers(-4, 5, (m, n)).astype(np.float64)
    B = opA @ X if side == 0 else X @ opA
    a = opA.T.reshape(-1).copy()
    b = B.T.reshape(-1).copy()
    level3.trsm(side, 1, 0, 0, m, n, 1.0, a, q, b, m)
    assert np.array_equal(b.reshape(n, m).T, X)


def _ill_conditioned_lower(kind, q, rng):
    """Lower triangular q x q (float64): 'unit' has a unit diagonal and N(0,1)
    off-diagonals (condition number growing like 2^q); 'spread' has
    diagonals log-uniform over [1e-3, 1e3] and N(0,1)/sqrt(q) below."""
    if kind == "unit":
        return np.tril(rng.standard_normal((q, q)), -1) + np.eye(q)
    L = np.tril(rng.standard_normal((q, q)), -1) / np.sqrt(q)
    np.fill_diagonal(L, np.exp(rng.uniform(np.log(1e-3), np.log(1e3), q)))
    return L


def _solve_longdouble(T, C, lower):
    """T Y = C by substitution in 80-bit extended precision (x87 long
    double): the forward-error reference."""
    q = T.shape[0]
    Tl = T.astype(np.longdouble)
    Y = C.astype(np.longdouble).copy()
    order = range(q) if lower else range(q - 1, -1, -1)
    for i in order:
        rest = slice(0, i) if lower else slice(i + 1, q)
        Y[i] = (Y[i] - Tl[i, rest] @ Y[rest]) / Tl[i, i]
    return Y


ILL_CASES = [(np.float64, q, kind) for q in (64, 200, 1000) for kind in ("unit", "spread")] + \
    [(np.float32, 64, "unit"), (np.float32, 200, "unit"), (np.float32, 64, "spread")]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,q,kind", ILL_CASES)
@pytest.mark.parametrize("side", [0, 1])
def test_trsm_ill_conditioned_vs_substitution(gpu, oracle, dtype, q, kind, side):
    """Ill-conditioned triangular factors, where an explicit inverse of a
    diagonal block is less stable than rblas's substitution
    (level3.rs:440-520, unfused s - a x, then / a_ii). The GPU must stay
    within the residual bound 16 q u (|op(A)||X| + |B|) and within 16x the
    oracle's forward error (floor q u) against an extended-precision solve.
    Left: L X = B (forward substitution); Right: X L = B, i.e. L^T X^T =
    B^T (backward). float32 runs the sizes whose solutions stay finite in
    float32 (the 'unit' solution reaches ~1e114 at q = 1000, the 'spread'
    one ~1e58 at q = 200)."""
    rng = np.random.default_rng(q + (kind == "unit") * 7 + side * 13)
    L = _ill_conditioned_lower(kind, q, rng).astype(dtype)
    nrhs = 8
    m, n = (q, nrhs) if side == 0 else (nrhs, q)
    B = rng.standard_normal((m, n)).astype(dtype)
    a = L.T.reshape(-1).copy()  # column-major, lda = q, uplo Lower
    b = B.T.reshape(-1).copy()  # column-major, ldb = m
    b_ref = b.copy()
    level3.trsm(side, 1, 0, 0, m, n, 1.0, a, q, b, m)
    oracle.trsm(side, 1, 0, 0, m, n, 1.0, a, q, b_ref, m)
    L64 = L.astype(np.float64)
    B64 = B.astype(np.float64)
    if side == 0:
        X_true = _solve_longdouble(L64, B64, lower=True)
    else:
        X_true = _solve_longdouble(L64.T, B64.T, lower=False).T
    scale = float(np.abs(X_true).max())
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
    out = {}
    for what, got in (("gpu", b), ("oracle", b_ref)):
        X = got.reshape(n, m).T.astype(np.float64)
        assert np.isfinite(X).all(), what
        out[what + "_resid"] = _residual_ok(side, L64, X, 1.0, B64, q, dtype)
        out[what + "_ferr"] = float(np.abs(X - X_true.astype(np.float64)).max() / scale)
    print({k: f"{v:.3g}" for k, v in out.items()}, f"max|X| {scale:.3g}")
    assert out["oracle_resid"] <= 1.0
    assert out["gpu_resid"] <= 1.0
    assert out["gpu_ferr"] <= 16 * max(out["oracle_ferr"], q * u)
