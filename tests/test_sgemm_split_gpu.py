"""Split-precision SGEMM (BF16x3 / BF16x6 on bf16 MFMA) through the C ABI.

Tolerances (elementwise, written here):
  exact:   small-integer operands are exact in one bf16 term, so every mode
           must reproduce float64 bit for bit;
  BF16x6:  |dC| <= 2 (k+2) u |alpha||A||B| + 4u |beta C0| (u = 2^-24), the f32
           bound: the 3-term split represents each f32 to ~2^-24;
  BF16x3:  |dC| <= (3*2^-16 + 2 (k+2) u) |alpha||A||B| + 4u |beta C0|: each
           2-term product misses up to ~3*2^-16 of |a b|.
"""
import itertools

import numpy as np
import pytest

from rblas_amd import level3, workloads
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import Compute
from tests.util import COL, ROW, op, slice_len, view

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


def problem(layout, ta, tb, m, n, k, pad, seed):
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda = (ac if layout == ROW else ar) + pad
    ldb = (bc if layout == ROW else br) + pad
    ldc = (n if layout == ROW else m) + pad
    la, lb, lc = (slice_len(layout, ar, ac, lda), slice_len(layout, br, bc, ldb),
                  slice_len(layout, m, n, ldc))
    a = workloads.normal(seed, 0, la, np.float32)
    b = workloads.normal(seed, la, lb, np.float32)
    c = workloads.normal(seed, la + lb, lc, np.float32)
    return a, lda, b, ldb, c, ldc


@pytest.mark.parametrize("compute", [Compute.BF16X3, Compute.BF16X6])
@pytest.mark.parametrize("layout,ta,tb", list(itertools.product([ROW, COL], [0, 1], [0, 1])))
def test_split_exact_integers(gpu, compute, layout, ta, tb):
    rng = np.random.default_rng(31)
    m, n, k = 203, 141, 77
    a, lda, b, ldb, c, ldc = problem(layout, ta, tb, m, n, k, 7, 2)
    for x in (a, b, c):
        x[:] = rng.integers(-4, 5, x.size)
    c0 = c.copy()
    level3.sgemm_ex(layout, ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c, ldc, compute)
    A = op(a, layout, ta, m, k, lda).astype(np.float64)
    B = op(b, layout, tb, k, n, ldb).astype(np.float64)
    want = 2.0 * (A @ B) - view(c0, layout, m, n, ldc)
    assert np.array_equal(view(c, layout, m, n, ldc).astype(np.float64), want)
    mask = np.ones(c.size, bool)
    view(mask, layout, m, n, ldc)[:] = False
    assert np.array_equal(c[mask], c0[mask])


@pytest.mark.parametrize("compute", [Compute.BF16X3, Compute.BF16X6])
@pytest.mark.parametrize("layout,ta,tb", [(COL, 0, 0), (COL, 1, 1), (ROW, 0, 1)])
def test_split_accuracy_bound(gpu, compute, layout, ta, tb):
    m, n, k = 300, 260, 1000
    a, lda, b, ldb, c, ldc = problem(layout, ta, tb, m, n, k, 7, 3)
    c0 = c.copy()
    level3.sgemm_ex(layout, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c, ldc, compute)
    A = op(a, layout, ta, m, k, lda).astype(np.float64)
    B = op(b, layout, tb, k, n, ldb).astype(np.float64)
    C0 = view(c0, layout, m, n, ldc).astype(np.float64)
    want = 1.5 * (A @ B) - 0.25 * C0
    mag = 1.5 * (np.abs(A) @ np.abs(B))
    rep = 3 * 2.0 ** -16 if compute == Compute.BF16X3 else 0.0
    tol = (rep + 2 * (k + 2) * U) * mag + 4 * U * 0.25 * np.abs(C0)
    err = np.abs(view(c, layout, m, n, ldc) - want)
    assert (err <= tol).all(), float((err / tol).max())


def test_split_bf16x6_close_to_f32_error(gpu):
    """BF16x6's max error against float64 stays within 2x the plain f32
    kernel's on N(0,1) operands (the split SGEMM precision target)."""
    m = n = 512
    k = 4096
    a, lda, b, ldb, c, ldc = problem(COL, 0, 0, m, n, k, 0, 4)
    want = (a.reshape(k, m).T.astype(np.float64) @ b.reshape(n, k).T.astype(np.float64))
    errs = {}
    for compute in (Compute.F32, Compute.BF16X6, Compute.BF16X3):
        cc = np.zeros(m * n, np.float32)
        level3.sgemm_ex(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, cc, m, compute)
        errs[compute] = float(np.abs(cc.reshape(n, m).T - want).max())
    print({c.name: e for c, e in errs.items()})
    assert errs[Compute.BF16X6] <= 2.0 * errs[Compute.F32]
    assert errs[Compute.BF16X3] > errs[Compute.BF16X6]


@pytest.mark.parametrize("compute", [Compute.BF16X3, Compute.BF16X6])
def test_split_device_entry(gpu, compute):
    m, n, k = 517, 389, 1025
    a, lda, b, ldb, c, ldc = problem(COL, 0, 1, m, n, k, 7, 5)
    ch = c.copy()
    level3.sgemm_ex(COL, 0, 1, m, n, k, 1.5, a, lda, b, ldb, -0.25, ch, ldc, compute)
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    level3.sgemm_ex_dev(COL, 0, 1, m, n, k, 1.5, da, lda, db, ldb, -0.25, dc, ldc, compute, s)
    s.synchronize()
    assert np.array_equal(dc.download(s), ch)
    s.close()


@pytest.mark.parametrize("compute", [Compute.BF16X3, Compute.BF16X6])
@pytest.mark.parametrize("m,n,k,ta,tb", [(300, 260, 5000, 0, 0), (64, 40, 20000, 1, 1),
                                         (1024, 256, 9000, 0, 1)])
def test_split_k_exact_deterministic_and_bounded(gpu, compute, m, n, k, ta, tb):
    """Shapes whose 256 x 256 tile grid fills under half the CUs split k over
    several workgroups (ordered reduction of the partial tiles): exact on
    small integers, the same bits on a second call, within the mode's bound."""
    rng = np.random.default_rng(m + k)
    a, lda, b, ldb, c, ldc = problem(COL, ta, tb, m, n, k, 3, 8)
    ai, bi, ci = (rng.integers(-3, 4, x.size).astype(np.float32) for x in (a, b, c))
    cc = ci.copy()
    level3.sgemm_ex(COL, ta, tb, m, n, k, 1.0, ai, lda, bi, ldb, 0.5, cc, ldc, compute)
    A = op(ai, COL, ta, m, k, lda).astype(np.float64)
    B = op(bi, COL, tb, k, n, ldb).astype(np.float64)
    want = A @ B + 0.5 * view(ci, COL, m, n, ldc)
    assert np.array_equal(view(cc, COL, m, n, ldc).astype(np.float64), want)
    c1, c2 = c.copy(), c.copy()
    level3.sgemm_ex(COL, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c1, ldc, compute)
    level3.sgemm_ex(COL, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c2, ldc, compute)
    assert np.array_equal(c1, c2)
    A = op(a, COL, ta, m, k, lda).astype(np.float64)
    B = op(b, COL, tb, k, n, ldb).astype(np.float64)
    C0 = view(c, COL, m, n, ldc).astype(np.float64)
    want = 1.5 * (A @ B) - 0.25 * C0
    mag = 1.5 * (np.abs(A) @ np.abs(B))
    rep = 3 * 2.0 ** -16 if compute == Compute.BF16X3 else 0.0
    tol = (rep + 2 * (k + 2) * U) * mag + 4 * U * 0.25 * np.abs(C0) + 1e-30
    assert (np.abs(view(c1, COL, m, n, ldc) - want) <= tol).all()
