"""Helpers shared by the tests: views of rblas-style slices as matrices,
naive float64 references and the elementwise gemm error bound."""
from __future__ import annotations

import numpy as np

ROW, COL = 0, 1


def view(x: np.ndarray, layout: int, rows: int, cols: int, ld: int) -> np.ndarray:
    """The stored rows x cols matrix of slice x (no copy)."""
    if rows == 0 or cols == 0:
        return np.zeros((rows, cols), x.dtype)
    it = x.itemsize
    if layout == ROW:
        return np.lib.stride_tricks.as_strided(x, (rows, cols), (ld * it, it))
    return np.lib.stride_tricks.as_strided(x, (rows, cols), (it, ld * it))


def op(x, layout, trans, rows, cols, ld):
    """op(X) (rows x cols) of a stored slice; trans 1 = T, 2 = conj T."""
    if trans:
        s = view(x, layout, cols, rows, ld).T
        return np.conj(s) if trans == 2 else s
    return view(x, layout, rows, cols, ld)


def slice_len(layout, rows, cols, ld):
    if rows == 0 or cols == 0:
        return 0
    return (rows - 1) * ld + cols if layout == ROW else (cols - 1) * ld + rows


def gemm_expect(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """float64 reference of C := alpha op(A) op(B) + beta C, with rblas's
    beta == 0 (exact zeros) and alpha == 0 / k == 0 (no product) rules.
    Returns (expected m x n float64, |alpha||A||B| + |beta||C0| bound base)."""
    A = op(a, layout, ta, m, k, lda).astype(np.float64)
    B = op(b, layout, tb, k, n, ldb).astype(np.float64)
    C0 = view(c, layout, m, n, ldc).astype(np.float64)
    scaled = np.zeros_like(C0) if beta == 0 else beta * C0
    if k == 0 or alpha == 0:
        prod = np.zeros_like(C0)
        mag = np.zeros_like(C0)
    else:
        prod = alpha * (A @ B)
        mag = abs(alpha) * (np.abs(A) @ np.abs(B))
    if beta != 0:
        mag = mag + abs(beta) * np.abs(C0)
    return prod + scaled, mag


def unit_roundoff(dtype) -> float:
    return 2.0 ** -24 if np.dtype(dtype) == np.float32 else 2.0 ** -53


def assert_gemm_close(got, want, mag, k, dtype, c_const=2.0, what=""):
    """|got - want| <= c * (k + 2) * u * mag elementwise (the classical
    forward bound for a length-k dot product plus the alpha/beta update)."""
    u = unit_roundoff(dtype)
    tol = c_const * (k + 2) * u * mag + np.finfo(np.float64).tiny
    wide = np.complex128 if np.iscomplexobj(got) or np.iscomplexobj(want) else np.float64
    err = np.abs(got.astype(wide) - want)
    bad = err > tol
    if bad.any():
        idx = np.argwhere(bad)[0]
        raise AssertionError(
            f"{what}: {bad.sum()} entries over the bound; first {tuple(idx)} "
            f"got {got[tuple(idx)]!r} want {want[tuple(idx)]!r} tol {tol[tuple(idx)]!r}")
    return float((err / np.maximum(tol, 1e-300)).max()) if err.size else 0.0


def lcg_f32(state: int, count: int, shift: int, scale_bits: int, offset: float):
    """The reference tests' LCG (tests/correctness.rs:21-26, tests/parallel.rs:9-14):
    s = s*6364136223846793005 + 1442695040888963407 (wrapping);
    value = (s >> shift) as f32 / 2^scale_bits - offset."""
    out = np.empty(count, np.float32)
    s = state
    mask = (1 << 64) - 1
    denom = np.float32(2.0 ** scale_bits)
    for i in range(count):
        s = (s * 6364136223846793005 + 1442695040888963407) & mask
        out[i] = np.float32(s >> shift) / denom - np.float32(offset)
    return out, s
