"""Complex trsm through the C ABI.

CPU part: the oracle's restatement of complex3::trsm pinned against
rblas's round-trip test and numpy's solve; where rblas panics
(oracle.trsm_panics: complex3::trsm walks the same loops as level3::trsm)
against the library's closed-form checks.

GPU part, tolerance written here: componentwise residual
  |op(A) X - alpha B| <= 32 q u (|op(A)| |X| + |alpha B|)
(q the order of A, u = 2^-24 for complex64 / 2^-53 for complex128; 32 = 16
as for the real trsm, x2 for the complex product), for the GPU and the
oracle alike, on matrices of the batching family (diagonal 2 + |N(0,1)|,
off-diagonal N(0,1)/sqrt(q), complex); exact-integer systems bit-exact;
non-finite entries reach what the oracle's substitution reaches.
"""
import itertools

import numpy as np
import pytest

from rblas_amd import complex3
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic, DeviceError

COMBOS = list(itertools.product((0, 1), (0, 1), (0, 1, 2), (0, 1)))  # side, uplo, trans, diag


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


def _mat(x, rows, cols, ld):
    return np.array([[x[i + j * ld] for j in range(cols)] for i in range(rows)])


def _system(q, uplo, trans, diag, lda, ctype, rng):
    """Stored q x q A (ld lda), its op(A) as the solver sees it, and junk in
    the unused triangle that must not be read."""
    full = (rng.standard_normal((q, q)) + 1j * rng.standard_normal((q, q))) / np.sqrt(q)
    np.fill_diagonal(full, (2.0 + np.abs(rng.standard_normal(q))) * np.exp(1j * rng.uniform(0, 6.28, q)))
    tri = np.triu(np.ones((q, q), bool)) if uplo == 0 else np.tril(np.ones((q, q), bool))
    stored = np.where(tri, full, 1e3 + 1e3j)
    if diag == 1:
        np.fill_diagonal(stored, 7e3 - 7e3j)
    a = np.full((q - 1) * lda + q, -5e2 + 5e2j, np.complex128)
    for j in range(q):
        a[j * lda:j * lda + q] = stored[:, j]
    a = a.astype(ctype)
    S = np.where(tri, _mat(a.astype(np.complex128), q, q, lda), 0)
    if diag == 1:
        np.fill_diagonal(S, 1.0)
    op = S if trans == 0 else (S.T if trans == 1 else S.conj().T)
    return a, op


def _ratio(side, op, X, alpha, B0, q, u):
    if side == 0:
        r = np.abs(op @ X - alpha * B0)
        bound = 32 * q * u * (np.abs(op) @ np.abs(X) + np.abs(alpha * B0))
    else:
        r = np.abs(X @ op - alpha * B0)
        bound = 32 * q * u * (np.abs(X) @ np.abs(op) + np.abs(alpha * B0))
    return float((r / np.maximum(bound, 1e-300)).max())


# ------------------------------------------------------------------ CPU side
def test_oracle_reference_roundtrip(oracle):
    # complex3.rs:439-478: L X = B, then L X == B to 1e-9.
    l = np.array([2 + 1j, 1, 0.5 - 1j, 0, 3 - 1j, 1 + 1j, 0, 0, 4 + 0.5j])
    b0 = np.array([1 + 1j, 2, 3j])
    b = b0.copy()
    oracle.ctrsm(0, 1, 0, 0, 3, 1, 1.0, l, 3, b, 3)
    L = l.reshape(3, 3).T
    for i in range(3):
        s = sum(L[i, j] * b[j] for j in range(i + 1))
        assert abs(s - b0[i]) < 1e-9


@pytest.mark.parametrize("ctype", [np.complex64, np.complex128])
def test_oracle_pinned_against_numpy(oracle, ctype):
    rng = np.random.default_rng(2)
    u = 2.0 ** -24 if ctype == np.complex64 else 2.0 ** -53
    for side, uplo, trans, diag in COMBOS:
        m, n = (9, 4) if side == 0 else (4, 9)
        q = m if side == 0 else n
        lda, ldb = q + 2, m + 1
        a, op = _system(q, uplo, trans, diag, lda, ctype, rng)
        b = (rng.standard_normal((n - 1) * ldb + m) + 1j * rng.standard_normal((n - 1) * ldb + m)).astype(ctype)
        b0 = b.copy()
        alpha = 0.5 - 1.5j
        oracle.ctrsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb)
        B0 = _mat(b0.astype(np.complex128), m, n, ldb)
        X = _mat(b.astype(np.complex128), m, n, ldb)
        want = (np.linalg.solve(op, alpha * B0) if side == 0
                else np.linalg.solve(op.T, (alpha * B0).T).T)
        assert np.abs(X - want).max() <= 1e3 * q * u * np.abs(want).max()
        assert _ratio(side, op, X, alpha, B0, q, u) <= 1.0


def test_ctrsm_panics_match_oracle(oracle):
    checked = panics = 0
    for side, uplo, trans, diag in COMBOS:
        for m, n, lda, ldb in itertools.product((0, 1, 3), (0, 2, 3), (1, 3, 4), (3, 5)):
            q = m if side == 0 else n
            for a_len in sorted({0, max(0, (q - 1) * (lda + 1)), q * lda + 1}):
                for b_len in sorted({0, max(0, (m - 1) + (n - 1) * ldb), n * ldb}):
                    want = oracle.trsm_panics(side, uplo, min(trans, 1), diag, m, n, 2.0, a_len,
                                              lda, b_len, ldb)
                    a = np.ones(a_len, np.complex128)
                    b = np.ones(b_len, np.complex128)
                    got = _status(lambda: complex3.trsm(side, uplo, trans, diag, m, n, 2.0, a,
                                                        lda, b, ldb))
                    if got == "bad":
                        continue
                    assert (got == "panic") == want, (side, uplo, trans, diag, m, n, lda, ldb,
                                                      a_len, b_len, got)
                    checked += 1
                    panics += want
    assert checked > 1000 and panics > 300


def test_ctrsm_bad_enums_and_empty():
    z = np.ones(16, np.complex128)
    for args in ((2, 0, 0, 0), (0, 2, 0, 0), (0, 0, 3, 0), (0, 0, 0, 2)):
        with pytest.raises(ValueError):
            complex3.trsm(*args, 2, 2, 1.0, z, 2, z, 2)
    e = np.zeros(0, np.complex128)
    b = np.full(4, 3 + 1j)
    complex3.trsm(0, 1, 0, 0, 0, 4, 2.0, e, 1, b, 1)
    complex3.trsm(1, 1, 2, 0, 4, 0, 2.0, e, 1, b, 4)
    assert (b == 3 + 1j).all()


# ------------------------------------------------------------------ GPU side
@pytest.mark.gpu
def test_ctrsm_reference_roundtrip(gpu):
    l = np.array([2 + 1j, 1, 0.5 - 1j, 0, 3 - 1j, 1 + 1j, 0, 0, 4 + 0.5j])
    b0 = np.array([1 + 1j, 2, 3j])
    b = b0.copy()
    complex3.trsm(0, 1, 0, 0, 3, 1, 1.0, l, 3, b, 3)
    L = l.reshape(3, 3).T
    for i in range(3):
        assert abs(sum(L[i, j] * b[j] for j in range(i + 1)) - b0[i]) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("ctype", [np.complex64, np.complex128])
@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_ctrsm_vs_oracle(gpu, oracle, ctype, side, uplo, trans, diag):
    m, n = (150, 37) if side == 0 else (37, 150)
    q = m if side == 0 else n
    rng = np.random.default_rng(200 + 24 * side + 12 * uplo + 4 * trans + diag)
    lda, ldb = q + 3, m + 4
    a, op = _system(q, uplo, trans, diag, lda, ctype, rng)
    b = (rng.standard_normal((n - 1) * ldb + m) + 1j * rng.standard_normal((n - 1) * ldb + m)).astype(ctype)
    b0, b_ref = b.copy(), b.copy()
    alpha = 0.75 + 0.25j
    complex3.trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb)
    oracle.ctrsm(side, uplo, trans, diag, m, n, alpha, a, lda, b_ref, ldb)
    u = 2.0 ** -24 if ctype == np.complex64 else 2.0 ** -53
    B0 = _mat(b0.astype(np.complex128), m, n, ldb)
    for what, got in (("gpu", b), ("oracle", b_ref)):
        X = _mat(got.astype(np.complex128), m, n, ldb)
        ratio = _ratio(side, op, X, alpha, B0, q, u)
        assert ratio <= 1.0, f"{what} residual over bound ({ratio:.3g})"
    keep = np.ones(b.size, bool)
    for j in range(n):
        keep[j * ldb:j * ldb + m] = False
    assert np.array_equal(b[keep], b0[keep])


@pytest.mark.gpu
@pytest.mark.parametrize("ctype", [np.complex64, np.complex128])
@pytest.mark.parametrize("side,uplo,trans", list(itertools.product((0, 1), (0, 1), (0, 1, 2))))
def test_ctrsm_exact_integers(gpu, ctype, side, uplo, trans):
    """Bidiagonal op(A) with unit-modulus diagonal (+-1, +-i) and off-diagonal
    in {0, +-1, +-i}: every substitution step and every 4M update is exact, so
    X equals the Gaussian-integer solution bit for bit."""
    m, n = (140, 21) if side == 0 else (21, 140)
    q = m if side == 0 else n
    rng = np.random.default_rng(11 + 6 * side + 3 * uplo + trans)
    units = np.array([1, -1, 1j, -1j])
    D = np.diag(units[rng.integers(0, 4, q)])
    off = np.array([0, 1, -1, 1j, -1j])[rng.integers(0, 5, q - 1)]
    eff_upper = (uplo == 0) != (trans != 0)
    opA = D + (np.diag(off, 1) if eff_upper else np.diag(off, -1))
    stored = opA if trans == 0 else (opA.T if trans == 1 else opA.conj().T)
    X = rng.integers(-3, 4, (m, n)) + 1j * rng.integers(-3, 4, (m, n))
    B = opA @ X if side == 0 else X @ opA
    a = stored.T.reshape(-1).astype(ctype)
    b = B.T.reshape(-1).astype(ctype)
    complex3.trsm(side, uplo, trans, 0, m, n, 1.0, a, q, b, m)
    assert np.array_equal(b.reshape(n, m).T, X.astype(ctype))


@pytest.mark.gpu
def test_ctrsm_orders_and_special_cases(gpu, oracle):
    rng = np.random.default_rng(9)
    for q in (1, 63, 64, 65, 129):
        a, op = _system(q, 1, 0, 0, q, np.complex128, rng)
        b = rng.standard_normal(q * 3) + 1j * rng.standard_normal(q * 3)
        b_ref = b.copy()
        complex3.trsm(0, 1, 0, 0, q, 3, -2.0 + 1j, a, q, b, q)
        oracle.ctrsm(0, 1, 0, 0, q, 3, -2.0 + 1j, a, q, b_ref, q)
        assert np.allclose(b, b_ref, rtol=1e-12, atol=1e-14)
    # A NaN in B reaches exactly the entries the oracle's substitution reaches.
    for side, uplo in itertools.product((0, 1), (0, 1)):
        m, n = (70, 5) if side == 0 else (5, 70)
        q = m if side == 0 else n
        a, _ = _system(q, uplo, 0, 0, q, np.complex128, rng)
        b = rng.standard_normal(m * n) + 1j * rng.standard_normal(m * n)
        b[(m * n) // 2] = complex(np.nan, 0.0)
        b_ref = b.copy()
        complex3.trsm(side, uplo, 0, 0, m, n, 1.0, a, q, b, m)
        oracle.ctrsm(side, uplo, 0, 0, m, n, 1.0, a, q, b_ref, m)
        assert np.array_equal(np.isnan(b), np.isnan(b_ref)), (side, uplo)
    # alpha == 0: B becomes 0 * B (NaN stays NaN), then X = 0.
    a, _ = _system(40, 0, 0, 0, 40, np.complex128, rng)
    b = rng.standard_normal(80) + 1j * rng.standard_normal(80)
    b[7] = complex(np.nan, np.nan)
    b_ref = b.copy()
    complex3.trsm(0, 0, 0, 0, 40, 2, 0.0, a, 40, b, 40)
    oracle.ctrsm(0, 0, 0, 0, 40, 2, 0.0, a, 40, b_ref, 40)
    assert np.array_equal(np.isnan(b), np.isnan(b_ref))
    assert (b[~np.isnan(b)] == 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("side", [0, 1])
def test_ctrsm_device_entry_matches_host(gpu, side):
    rng = np.random.default_rng(31 + side)
    m, n = (300, 90) if side == 0 else (90, 300)
    q = m if side == 0 else n
    a, _ = _system(q, 0, 2, 0, q, np.complex128, rng)
    b = rng.standard_normal(m * n) + 1j * rng.standard_normal(m * n)
    bh = b.copy()
    complex3.trsm(side, 0, 2, 0, m, n, 1.5 - 0.5j, a, q, bh, m)
    s = Stream()
    da, db = DeviceArray.from_host(a, s), DeviceArray.from_host(b, s)
    complex3.trsm_dev(side, 0, 2, 0, m, n, 1.5 - 0.5j, da, q, db, m, s)
    s.synchronize()
    assert np.array_equal(db.download(s), bh)
    s.close()
