"""symm (real) and symm / hemm (complex) through the C ABI.

CPU part: the oracle's restatement (full_symmetric walked literally, then
its gemm / 4M gemm) pinned against numpy's triu/tril mirror; where rblas
panics (oracle.symm_panics / hesy_panics walk the reference loops) against
the library's closed-form checks.

GPU part, tolerance written here (the gemm bound on the dense A):
  |dC| <= 2 (q + 2) u (|alpha| |A||B| + |beta| |C0|)   (q the order of A,
  u = 2^-24 for f32 / complex64, 2^-53 for f64 / complex128; complex
  magnitudes elementwise, with an extra factor 4 for the 4M split);
exact-integer cases are bit-exact; device and host entries agree bit for bit.
"""
import itertools

import numpy as np
import pytest

from rblas_amd import complex3, level3
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic, DeviceError

LEFT, RIGHT = 0, 1
UPPER, LOWER = 0, 1


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


def _dense(a, q, lda):
    return np.array([[a[i + j * lda] for j in range(q)] for i in range(q)])


def _mirror(S, uplo, herm=False):
    """numpy's mirror of the stored triangle (independent of the oracle's
    index walk)."""
    T = np.triu(S) if uplo == UPPER else np.tril(S)
    off = T - np.diag(np.diag(T))
    if herm:
        return off + off.conj().T + np.diag(np.diag(T).real)
    return off + off.T + np.diag(np.diag(T))


def _mat(x, rows, cols, ld):
    return np.array([[x[i + j * ld] for j in range(cols)] for i in range(rows)])


# ------------------------------------------------------------------ CPU side
@pytest.mark.parametrize("herm", [None, False, True])
def test_oracle_pinned_against_numpy(oracle, herm):
    rng = np.random.default_rng(3)
    for side, uplo in itertools.product((LEFT, RIGHT), (UPPER, LOWER)):
        m, n = 6, 5
        q = m if side == LEFT else n
        lda, ldb, ldc = q + 2, m + 1, m + 3
        cplx = herm is not None

        def draw(size):
            x = rng.standard_normal(size)
            return x + 1j * rng.standard_normal(size) if cplx else x
        a, b, c = draw((q - 1) * lda + q), draw((n - 1) * ldb + m), draw((n - 1) * ldc + m)
        c0 = c.copy()
        alpha, beta = (1.5 - 0.5j, 0.25 + 1j) if cplx else (1.5, -0.5)
        if cplx:
            oracle.hesy(herm, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc)
        else:
            oracle.symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc)
        F = _mirror(_dense(a, q, lda), uplo, bool(herm))
        B, C0 = _mat(b, m, n, ldb), _mat(c0, m, n, ldc)
        want = alpha * (F @ B if side == LEFT else B @ F) + beta * C0
        assert np.abs(_mat(c, m, n, ldc) - want).max() < 1e-12
        # The pin is sensitive to the triangle: the other triangle's mirror
        # (a planted mistake in the oracle's input) is far off.
        wrong = _mirror(_dense(a, q, lda), 1 - uplo, bool(herm))
        off = alpha * (wrong @ B if side == LEFT else B @ wrong) + beta * C0
        assert np.abs(_mat(c, m, n, ldc) - off).max() > 1e-3


def test_symm_panics_match_oracle(oracle):
    checked = panics = 0
    for side, uplo, m, n, lda, ldb, ldc in itertools.product((0, 1), (0, 1), (0, 1, 3), (0, 2, 3),
                                                             (1, 3, 4), (3, 4), (3, 4)):
        q = m if side == LEFT else n
        a_hi = (q - 1) * (lda + 1) if q else 0
        for a_len in sorted({0, a_hi, a_hi + 1}):
            for b_len in sorted({0, max(0, (n - 1) * ldb + m - 1), n * ldb}):
                for c_len in sorted({0, max(0, (n - 1) * ldc + m - 1), n * ldc}):
                    want = oracle.symm_panics(side, uplo, m, n, 1.0, a_len, lda, b_len, ldb, 0.5,
                                              c_len, ldc)
                    a, b, c = np.ones(a_len), np.ones(b_len), np.ones(c_len)
                    got = _status(lambda: level3.symm(side, uplo, m, n, 1.0, a, lda, b, ldb, 0.5,
                                                      c, ldc))
                    if got == "bad":  # a device-path limit, not a rblas outcome
                        continue
                    assert (got == "panic") == want, (side, uplo, m, n, lda, ldb, ldc, a_len,
                                                      b_len, c_len, got)
                    checked += 1
                    panics += want
    assert checked > 1000 and panics > 400


def test_hesy_panics_match_oracle(oracle):
    checked = panics = 0
    for herm, side, uplo, m, n, lda, ldb in itertools.product((False, True), (0, 1), (0, 1),
                                                              (0, 1, 3), (0, 2, 3), (1, 3, 4),
                                                              (3, 4)):
        q = m if side == LEFT else n
        a_hi = (q - 1) * (lda + 1) if q else 0
        for a_len in sorted({0, a_hi, a_hi + 1}):
            for b_len in sorted({0, max(0, (n - 1) * ldb + m - 1), n * ldb}):
                for c_len in sorted({0, max(0, (n - 1) * 3 + m - 1), n * 3}):
                    want = oracle.hesy_panics(side, uplo, m, n, a_len, lda, b_len, ldb, c_len, 3)
                    a, b, c = (np.ones(x, np.complex128) for x in (a_len, b_len, c_len))
                    fn = complex3.hemm if herm else complex3.symm
                    got = _status(lambda: fn(side, uplo, m, n, 1.0, a, lda, b, ldb, 0.5, c, 3))
                    if got == "bad":
                        continue
                    assert (got == "panic") == want, (herm, side, uplo, m, n, lda, ldb, a_len,
                                                      b_len, c_len, got)
                    checked += 1
                    panics += want
    assert checked > 1000 and panics > 300


def test_symm_bad_enums_and_empty():
    a = np.zeros(16)
    for side, uplo in ((2, 0), (0, 2), (-1, 0)):
        with pytest.raises(ValueError):
            level3.symm(side, uplo, 2, 2, 1.0, a, 2, a, 2, 0.0, a, 2)
        with pytest.raises(ValueError):
            complex3.hemm(side, uplo, 2, 2, 1.0, a.astype(np.complex128), 2,
                          a.astype(np.complex128), 2, 0.0, a.astype(np.complex128), 2)
    # Complex: m or n == 0 returns before any check (complex3.rs:190-192).
    z = np.zeros(0, np.complex128)
    complex3.symm(0, 0, 0, 4, 1.0, z, 1, z, 1, 0.0, z, 1)
    complex3.hemm(1, 1, 3, 0, 1.0, z, 1, z, 1, 0.0, z, 1)


# ------------------------------------------------------------------ GPU side
@pytest.mark.gpu
def test_symm_reference_kats(gpu):
    # level3.rs:769-792: A = [[2,1],[1,3]] upper-stored, C = A B exactly.
    a = np.array([2.0, 0.0, 1.0, 3.0])
    b = np.array([1.0, 2.0, 3.0, 4.0])
    c = np.zeros(4)
    level3.symm(LEFT, UPPER, 2, 2, 1.0, a, 2, b, 2, 0.0, c, 2)
    assert c.tolist() == [4.0, 7.0, 10.0, 15.0]
    # level3.rs:826-862: Right, lower, 7 x 6, against the explicit mirror, 1e-9.
    m, n = 7, 6
    al = np.cos(np.arange(n * n) * 0.17)
    b = np.arange(m * n) * 0.05
    c = np.zeros(m * n)
    level3.symm(RIGHT, LOWER, m, n, 1.0, al, n, b, m, 0.0, c, m)
    sym = lambda i, j: al[i + j * n] if i >= j else al[j + i * n]
    for j in range(n):
        for i in range(m):
            acc = sum(b[i + p * m] * sym(p, j) for p in range(n))
            assert abs(c[i + j * m] - acc) < 1e-9


@pytest.mark.gpu
def test_hemm_reference_kats(gpu):
    # complex3.rs:479-509: A = [[2, 1+i],[1-i, 3]] upper-stored (the lower
    # entry 9+9i is not read), B = (1, i): C = (1+i, 1+2i) to 1e-12.
    a = np.array([2, 9 + 9j, 1 + 1j, 3], np.complex128)
    b = np.array([1, 1j], np.complex128)
    c = np.zeros(2, np.complex128)
    complex3.hemm(LEFT, UPPER, 2, 1, 1.0, a, 2, b, 2, 0.0, c, 2)
    assert abs(c[0] - (1 + 1j)) < 1e-12 and abs(c[1] - (1 + 2j)) < 1e-12
    # complex3.rs:511-557: lower-stored 5 x 5, B 5 x 3, alpha 1.5 - 0.5i, 1e-9.
    q, n = 5, 3
    i = np.arange(q * q)
    a = np.sin(i * 0.3) + 1j * np.cos(i * 0.4)
    j = np.arange(q * n)
    b = j * 0.1 + 1j * j * 0.05
    c = np.zeros(q * n, np.complex128)
    alpha = 1.5 - 0.5j
    complex3.hemm(LEFT, LOWER, q, n, alpha, a, q, b, q, 0.0, c, q)

    def her(r, s):
        if r > s:
            return a[r + s * q]
        if r < s:
            return np.conj(a[s + r * q])
        return complex(a[r + s * q].real, 0.0)
    for jj in range(n):
        for ii in range(q):
            acc = sum(her(ii, p) * b[p + jj * q] for p in range(q))
            assert abs(c[ii + jj * q] - alpha * acc) < 1e-9


def _bound_ok(got, want, mag, q, u, extra=1.0):
    return np.all(np.abs(got - want) <= extra * 2 * (q + 2) * u * mag + 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("side,uplo", list(itertools.product((LEFT, RIGHT), (UPPER, LOWER))))
def test_symm_vs_oracle(gpu, oracle, dtype, side, uplo):
    rng = np.random.default_rng(10 + 2 * side + uplo)
    m, n = (150, 70) if side == LEFT else (70, 150)
    q = m if side == LEFT else n
    lda, ldb, ldc = q + 3, m + 5, m + 2
    a = rng.standard_normal((q - 1) * lda + q).astype(dtype)
    b = rng.standard_normal((n - 1) * ldb + m).astype(dtype)
    c = rng.standard_normal((n - 1) * ldc + m).astype(dtype)
    c_ref, c0 = c.copy(), c.copy()
    level3.symm(side, uplo, m, n, 0.75, a, lda, b, ldb, -1.25, c, ldc)
    oracle.symm(side, uplo, m, n, 0.75, a, lda, b, ldb, -1.25, c_ref, ldc)
    F = _mirror(_dense(a.astype(np.float64), q, lda), uplo)
    B, C0 = _mat(b.astype(np.float64), m, n, ldb), _mat(c0.astype(np.float64), m, n, ldc)
    want = 0.75 * (F @ B if side == LEFT else B @ F) - 1.25 * C0
    mag = 0.75 * (np.abs(F) @ np.abs(B) if side == LEFT else np.abs(B) @ np.abs(F)) + 1.25 * np.abs(C0)
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
    for got in (c, c_ref):
        assert _bound_ok(_mat(got.astype(np.float64), m, n, ldc), want, mag, q, u)
    # Entries between columns (rows m..ldc-1) are untouched.
    keep = np.ones(c.size, bool)
    for jj in range(n):
        keep[jj * ldc:jj * ldc + m] = False
    assert np.array_equal(c[keep], c0[keep])


@pytest.mark.gpu
@pytest.mark.parametrize("ctype", [np.complex64, np.complex128])
@pytest.mark.parametrize("herm,side,uplo",
                         list(itertools.product((False, True), (LEFT, RIGHT), (UPPER, LOWER))))
def test_hesy_vs_oracle(gpu, oracle, ctype, herm, side, uplo):
    rng = np.random.default_rng(40 + 4 * herm + 2 * side + uplo)
    m, n = (90, 40) if side == LEFT else (40, 90)
    q = m if side == LEFT else n
    lda, ldb, ldc = q + 1, m + 2, m + 3

    def draw(size):
        return (rng.standard_normal(size) + 1j * rng.standard_normal(size)).astype(ctype)
    a, b, c = draw((q - 1) * lda + q), draw((n - 1) * ldb + m), draw((n - 1) * ldc + m)
    c_ref, c0 = c.copy(), c.copy()
    alpha, beta = 0.5 + 1.25j, -0.75 + 0.5j
    fn = complex3.hemm if herm else complex3.symm
    fn(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc)
    oracle.hesy(herm, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c_ref, ldc)
    F = _mirror(_dense(a.astype(np.complex128), q, lda), uplo, herm)
    B, C0 = _mat(b.astype(np.complex128), m, n, ldb), _mat(c0.astype(np.complex128), m, n, ldc)
    want = alpha * (F @ B if side == LEFT else B @ F) + beta * C0
    mag = abs(alpha) * (np.abs(F) @ np.abs(B) if side == LEFT else np.abs(B) @ np.abs(F)) \
        + abs(beta) * np.abs(C0)
    u = 2.0 ** -24 if ctype == np.complex64 else 2.0 ** -53
    for got in (c, c_ref):
        assert _bound_ok(_mat(got.astype(np.complex128), m, n, ldc), want, mag, q, u, extra=4.0)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo", list(itertools.product((LEFT, RIGHT), (UPPER, LOWER))))
def test_symm_exact_integers(gpu, side, uplo):
    """Small integers: every product and sum is exact in f64, so C equals
    the integer result bit for bit; the unstored triangle holds values that
    would change it if read."""
    rng = np.random.default_rng(7 + 2 * side + uplo)
    m, n = (33, 20) if side == LEFT else (20, 33)
    q = m if side == LEFT else n
    S = rng.integers(-3, 4, (q, q)).astype(np.float64)
    F = _mirror(S, uplo)
    junk = np.where(np.triu(np.ones((q, q), bool)) if uplo == UPPER
                    else np.tril(np.ones((q, q), bool)), S, 1e6)
    a = junk.T.reshape(-1).copy()
    B = rng.integers(-3, 4, (m, n)).astype(np.float64)
    C0 = rng.integers(-3, 4, (m, n)).astype(np.float64)
    b, c = B.T.reshape(-1).copy(), C0.T.reshape(-1).copy()
    level3.symm(side, uplo, m, n, 2.0, a, q, b, m, -1.0, c, m)
    want = 2.0 * (F @ B if side == LEFT else B @ F) - C0
    assert np.array_equal(c.reshape(n, m).T, want)


@pytest.mark.gpu
def test_hemm_diagonal_imaginary_ignored(gpu):
    """The stored diagonal's imaginary part is not used (taken as zero)."""
    rng = np.random.default_rng(5)
    q, n = 37, 9
    a = (rng.standard_normal(q * q) + 1j * rng.standard_normal(q * q))
    b = rng.standard_normal(q * n) + 1j * rng.standard_normal(q * n)
    c1, c2 = np.zeros(q * n, np.complex128), np.zeros(q * n, np.complex128)
    a2 = a.copy()
    a2[np.arange(q) * (q + 1)] += 1000j
    complex3.hemm(LEFT, UPPER, q, n, 1.0, a, q, b, q, 0.0, c1, q)
    complex3.hemm(LEFT, UPPER, q, n, 1.0, a2, q, b, q, 0.0, c2, q)
    assert np.array_equal(c1, c2)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["d", "z_hemm"])
def test_device_entries_match_host(gpu, kind):
    rng = np.random.default_rng(8)
    m, n = 300, 130
    s = Stream()
    if kind == "d":
        a = rng.standard_normal(m * m)
        b = rng.standard_normal(m * n)
        c = rng.standard_normal(m * n)
        ch = c.copy()
        level3.symm(LEFT, LOWER, m, n, 1.5, a, m, b, m, 0.5, ch, m)
        da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
        level3.symm_dev(LEFT, LOWER, m, n, 1.5, da, m, db, m, 0.5, dc, m, s)
    else:
        a = rng.standard_normal(n * n) + 1j * rng.standard_normal(n * n)
        b = rng.standard_normal(m * n) + 1j * rng.standard_normal(m * n)
        c = rng.standard_normal(m * n) + 1j * rng.standard_normal(m * n)
        ch = c.copy()
        complex3.hemm(RIGHT, UPPER, m, n, 1 - 1j, a, n, b, m, 0.5j, ch, m)
        da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
        complex3.hemm_dev(RIGHT, UPPER, m, n, 1 - 1j, da, n, db, m, 0.5j, dc, m, s)
    s.synchronize()
    assert np.array_equal(dc.download(s), ch)
    s.close()
