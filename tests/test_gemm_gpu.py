"""GPU gemm parity: librblas_hip.so against the oracle (rblas restated),
numpy float64 and exact integer arithmetic, through the C ABI.

Tolerance (written here, used by every float case): elementwise
|C_gpu - C_exact| <= 2 (k + 2) u (|alpha| |A||B| + |beta| |C0|), with u the
unit roundoff of the type (2^-24 f32, 2^-53 f64).
"""
import itertools
import math

import numpy as np
import pytest

from rblas_amd import level3, workloads
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic
from tests.util import COL, ROW, assert_gemm_close, gemm_expect, lcg_f32, slice_len, view

pytestmark = pytest.mark.gpu

LAYOUTS = list(itertools.product([ROW, COL], [0, 1, 2], [0, 1, 2]))


def make_problem(layout, ta, tb, m, n, k, pad, dtype, seed, dist="normal"):
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda = (ac if layout == ROW else ar) + pad
    ldb = (bc if layout == ROW else br) + pad
    ldc = (n if layout == ROW else m) + pad
    la, lb, lc = (slice_len(layout, ar, ac, lda), slice_len(layout, br, bc, ldb),
                  slice_len(layout, m, n, ldc))
    gen = workloads.normal if dist == "normal" else workloads.uniform
    a = gen(seed, 0, la, dtype)
    b = gen(seed, la, lb, dtype)
    c = gen(seed, la + lb, lc, dtype)
    return a, lda, b, ldb, c, ldc


def test_config1_sgemm_256(gpu, oracle):
    """Plain SGEMM 256^3 col-major NN, U[-1,1] seed 0, alpha 1, beta 0.5."""
    m = n = k = 256
    a, lda, b, ldb, c, ldc = make_problem(COL, 0, 0, m, n, k, 0, np.float32, 0, "uniform")
    c0 = c.copy()
    c_ref = c.copy()
    level3.gemm(COL, 0, 0, m, n, k, 1.0, a, lda, b, ldb, 0.5, c, ldc)
    oracle.gemm(COL, 0, 0, m, n, k, 1.0, a, lda, b, ldb, 0.5, c_ref, ldc)
    want, mag = gemm_expect(COL, 0, 0, m, n, k, 1.0, a, lda, b, ldb, 0.5, c0, ldc)
    r_gpu = assert_gemm_close(view(c, COL, m, n, ldc), want, mag, k, np.float32, what="gpu")
    r_cpu = assert_gemm_close(view(c_ref, COL, m, n, ldc), want, mag, k, np.float32, what="oracle")
    print(f"config1 max err/bound: gpu {r_gpu:.3g} oracle {r_cpu:.3g}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout,ta,tb", LAYOUTS)
def test_gemm_exact_integers(gpu, dtype, layout, ta, tb):
    """Small-integer operands make every product and partial sum exact, so
    the GPU result must equal the float64 result bit for bit."""
    rng = np.random.default_rng(100 + 9 * layout + 3 * ta + tb)
    m, n, k = 203, 141, 77
    a, lda, b, ldb, c, ldc = make_problem(layout, ta, tb, m, n, k, 7, dtype, 1)
    for x in (a, b, c):
        x[:] = rng.integers(-4, 5, x.size)
    c0 = c.copy()
    level3.gemm(layout, ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c, ldc)
    want, _ = gemm_expect(layout, ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c0, ldc)
    got = view(c, layout, m, n, ldc).astype(np.float64)
    assert np.array_equal(got, want)
    # padding between rows/columns of C is never written
    mask = np.ones(c.size, bool)
    view(mask, layout, m, n, ldc)[:] = False
    assert np.array_equal(c[mask], c0[mask])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout,ta,tb", LAYOUTS)
def test_gemm_random_vs_oracle(gpu, oracle, dtype, layout, ta, tb):
    """Config-2 form (ld +7, alpha 1.5, beta -0.25, N(0,1) seed 1) at a size
    with ragged tile edges on every side."""
    m, n, k = 301, 259, 517
    a, lda, b, ldb, c, ldc = make_problem(layout, ta, tb, m, n, k, 7, dtype, 1)
    c0, c_ref = c.copy(), c.copy()
    level3.gemm(layout, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c, ldc)
    oracle.gemm(layout, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c_ref, ldc)
    want, mag = gemm_expect(layout, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.25, c0, ldc)
    assert_gemm_close(view(c, layout, m, n, ldc), want, mag, k, dtype, what="gpu")
    assert_gemm_close(view(c_ref, layout, m, n, ldc), want, mag, k, dtype, what="oracle")
    # GPU and oracle each within the bound -> within twice it of each other
    diff = np.abs(view(c, layout, m, n, ldc).astype(np.float64) - view(c_ref, layout, m, n, ldc))
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
    assert (diff <= 4 * (k + 2) * u * mag + 1e-300).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gemm_special_cases(gpu, dtype):
    # kernel/gemm.rs:198-205, 322-338
    m, n, k = 70, 33, 19
    a = np.full(m * k, np.nan, dtype)
    b = np.ones(k * n, dtype)
    c = np.full(m * n, np.nan, dtype)
    level3.gemm(COL, 0, 0, m, n, k, 0.0, a, m, b, k, 0.0, c, m)  # alpha 0: A never read
    assert (c == 0).all()
    c[:] = 2
    level3.gemm(COL, 0, 0, m, n, 0, 1.0, a, m, b, k, 0.5, c, m)  # k 0: beta only
    assert (c == 1).all()
    c[:] = 3
    level3.gemm(COL, 0, 0, m, n, 0, 1.0, a, m, b, k, 1.0, c, m)  # beta 1: untouched
    assert (c == 3).all()
    a[:] = 1
    c[:] = np.nan
    level3.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)  # beta 0 clears NaN
    assert (c == k).all()
    c[:] = 1
    level3.gemm(COL, 1, 1, m, n, k, 1.0, a, k, b, n, 1.0, c, m)  # beta 1 accumulates
    assert (c == k + 1).all()


def test_reference_fixture_17x13x19(gpu):
    # kernel/gemm.rs:690-701 through the GPU, tolerance 1e-9 as in rblas
    m, n, k = 17, 13, 19
    a = np.sin(np.arange(m * k) * 0.5)
    b = np.cos(np.arange(k * n) * 0.3)
    c = np.zeros(m * n)
    level3.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = (a.reshape(k, m).T @ b.reshape(n, k).T).T.reshape(-1)
    assert np.abs(c - want).max() < 1e-9


def test_reference_fixture_8cubed_alpha_beta(gpu):
    # kernel/gemm.rs:704-717 through the GPU: f32, alpha 2, beta 3, C = 1.
    # Every product and partial sum is an integer below 2^24, so the whole C
    # is exact; rblas checks C[0] to 1e-3.
    m = n = k = 8
    a = np.arange(m * k, dtype=np.float32)
    b = (np.arange(k * n) % 5).astype(np.float32)
    c = np.ones(m * n, np.float32)
    level3.gemm(COL, 0, 0, m, n, k, 2.0, a, m, b, k, 3.0, c, m)
    acc = sum(float(a[p * m]) * float(b[p]) for p in range(k))
    assert abs(c[0] - (2 * acc + 3)) < 1e-3
    A = a.reshape(k, m).T.astype(np.float64)
    B = b.reshape(n, k).T.astype(np.float64)
    assert np.array_equal(c.reshape(n, m).T.astype(np.float64), 2 * A @ B + 3)


@pytest.mark.parametrize("case", ["nn_colmajor", "tn_colmajor", "rowmajor_nn"])
def test_reference_layout_kats(gpu, case):
    # level3.rs:642-736 through the GPU, against the naive product, 1e-9.
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
    level3.gemm(*args, 0.0, c, ldc)
    want, _ = gemm_expect(*args, 0.0, np.zeros(m * n), ldc)
    np.testing.assert_allclose(view(c, layout, m, n, ldc), want, rtol=0, atol=1e-9)


def test_reference_fixture_lcg_200x180x150(gpu):
    # tests/correctness.rs:29-58 through the GPU, max rel < 1e-3 as in rblas
    m, n, k = 200, 180, 150
    a, s = lcg_f32(0x1234_5678, m * k, 33, 31, 1.0)
    b, _ = lcg_f32(s, k * n, 33, 31, 1.0)
    c = np.zeros(m * n, np.float32)
    level3.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = (a.reshape(k, m).T.astype(np.float64) @ b.reshape(n, k).T).T.reshape(-1)
    rel = np.abs(c - want) / np.maximum(np.abs(want), 1e-4)
    assert rel.max() < 1e-3


def test_reference_fixture_lcg_64x5000x48(gpu):
    # tests/parallel.rs:17-63 through the GPU, max rel < 1e-2 as in rblas
    m, n, k = 64, 5000, 48
    a, s = lcg_f32(0xC0FFEE, m * k, 40, 24, 0.5)
    b, _ = lcg_f32(s, k * n, 40, 24, 0.5)
    c = np.zeros(m * n, np.float32)
    level3.gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    want = (a.reshape(k, m).T.astype(np.float64) @ b.reshape(n, k).T).T.reshape(-1)
    rel = np.abs(c - want) / np.maximum(np.abs(want), 1e-3)
    assert rel.max() < 1e-2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_overlapping_columns_nn(gpu, oracle, dtype):
    """rblas's NN path accepts lda < m (columns overlap) and computes with
    those strides; the GPU must read the same elements."""
    m, n, k = 40, 30, 20
    lda = 25
    rng = np.random.default_rng(4)
    a = rng.integers(-3, 4, (k - 1) * lda + m).astype(dtype)
    b = rng.integers(-3, 4, k * n).astype(dtype)
    c = np.zeros(m * n, dtype)
    c_ref = c.copy()
    level3.gemm(COL, 0, 0, m, n, k, 1.0, a, lda, b, k, 0.0, c, m)
    oracle.gemm(COL, 0, 0, m, n, k, 1.0, a, lda, b, k, 0.0, c_ref, m)
    assert np.array_equal(c, c_ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_device_entry_matches_host_entry(gpu, dtype):
    m, n, k = 513, 385, 1031
    a, lda, b, ldb, c, ldc = make_problem(COL, 1, 0, m, n, k, 7, dtype, 3)
    c_host = c.copy()
    level3.gemm(COL, 1, 0, m, n, k, 1.5, a, lda, b, ldb, -0.25, c_host, ldc)
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    level3.gemm_dev(COL, 1, 0, m, n, k, 1.5, da, lda, db, ldb, -0.25, dc, ldc, s)
    s.synchronize()
    assert np.array_equal(dc.download(s), c_host)
    s.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout,ta,tb", [(COL, 0, 0), (ROW, 1, 0), (COL, 1, 1)])
def test_gemm_4096_sampled(gpu, dtype, layout, ta, tb):
    """Config-2 size 4096^3: every entry of 48 sampled rows and 48 sampled
    columns of C against float64."""
    m = n = k = 4096
    a, lda, b, ldb, c, ldc = make_problem(layout, ta, tb, m, n, k, 7, dtype, 1)
    c0 = c.copy()
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    level3.gemm_dev(layout, ta, tb, m, n, k, 1.5, da, lda, db, ldb, -0.25, dc, ldc, s)
    s.synchronize()
    got = view(dc.download(s), layout, m, n, ldc)
    s.close()
    from tests.util import op
    A = op(a, layout, ta, m, k, lda).astype(np.float64)
    B = op(b, layout, tb, k, n, ldb).astype(np.float64)
    C0 = view(c0, layout, m, n, ldc).astype(np.float64)
    rng = np.random.default_rng(0)
    rows = rng.choice(m, 48, replace=False)
    cols = rng.choice(n, 48, replace=False)
    for sel_rows, sel_cols in ((rows, slice(None)), (slice(None), cols)):
        Ar = A[sel_rows]
        Bc = B[:, sel_cols]
        want = 1.5 * (Ar @ Bc) - 0.25 * C0[sel_rows][:, sel_cols] if isinstance(sel_rows, np.ndarray) \
            else 1.5 * (Ar @ Bc) - 0.25 * C0[:, sel_cols]
        mag = 1.5 * (np.abs(Ar) @ np.abs(Bc)) + 0.25 * np.abs(
            C0[sel_rows][:, sel_cols] if isinstance(sel_rows, np.ndarray) else C0[:, sel_cols])
        g = got[sel_rows][:, sel_cols] if isinstance(sel_rows, np.ndarray) else got[:, sel_cols]
        assert_gemm_close(g, want, mag, k, dtype, what="4096 sampled")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n,k", [(200, 190, 8192), (64, 48, 20000), (1024, 96, 4100)])
def test_gemm_split_k_exact_and_deterministic(gpu, dtype, m, n, k):
    """Shapes whose tile grid leaves most CUs idle run with k split over
    several workgroups and an ordered reduction of the partial tiles: exact
    on integers (every partial sum is exact), and repeatable bit for bit on
    random data."""
    rng = np.random.default_rng(m + n + k)
    for ta, tb in ((0, 0), (1, 1), (0, 1)):
        a, lda, b, ldb, c, ldc = make_problem(COL, ta, tb, m, n, k, 3, dtype, 2)
        for x in (a, b, c):
            x[:] = rng.integers(-3, 4, x.size)
        c0 = c.copy()
        level3.gemm(COL, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 2.0, c, ldc)
        want, _ = gemm_expect(COL, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 2.0, c0, ldc)
        assert np.array_equal(view(c, COL, m, n, ldc).astype(np.float64), want)
    a, lda, b, ldb, c, ldc = make_problem(COL, 0, 1, m, n, k, 3, dtype, 4)
    c1, c2 = c.copy(), c.copy()
    level3.gemm(COL, 0, 1, m, n, k, 1.5, a, lda, b, ldb, -0.25, c1, ldc)
    level3.gemm(COL, 0, 1, m, n, k, 1.5, a, lda, b, ldb, -0.25, c2, ldc)
    assert np.array_equal(c1, c2)
    want, mag = gemm_expect(COL, 0, 1, m, n, k, 1.5, a, lda, b, ldb, -0.25, c, ldc)
    assert_gemm_close(view(c1, COL, m, n, ldc), want, mag, k, dtype, what="split-k")


def test_try_gemm_valid_call_succeeds_reference(gpu):
    # level3.rs:574-591: the valid 4 x 3 x 5 call of ones returns Ok; every
    # entry of C is then k = 5.
    m, n, k = 4, 3, 5
    a, b, c = np.ones(m * k), np.ones(k * n), np.zeros(m * n)
    level3.try_gemm(COL, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert (c == k).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_try_gemm_ok_equals_gemm_bitwise(gpu, dtype):
    """Where validation passes, try_gemm runs the same gemm: C is equal bit
    for bit to level3.gemm's on the same inputs, over layouts, transposes,
    padded leading dimensions, exact-length slices and alpha / beta cases."""
    rng = np.random.default_rng(21)
    checked = 0
    for (layout, ta, tb), (m, n, k), pad, (alpha, beta) in itertools.product(
            LAYOUTS[::2], [(8, 8, 8), (37, 5, 130), (129, 66, 1)], (0, 3),
            [(1.5, -0.25), (1.0, 0.0), (0.0, 2.0)]):
        a, lda, b, ldb, c, ldc = make_problem(layout, ta, tb, m, n, k, pad, dtype,
                                              int(rng.integers(1 << 30)))
        c_try, c_gemm = c.copy(), c.copy()
        level3.try_gemm(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c_try, ldc)
        level3.gemm(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c_gemm, ldc)
        assert np.array_equal(c_try, c_gemm, equal_nan=True)
        checked += 1
    assert checked == 9 * 3 * 2 * 3


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout,ta,tb", list(itertools.product([COL, ROW], [0, 1], [0, 1])))
@pytest.mark.parametrize("k", [333, 5])
def test_wide_tile_edges_match_grouped(gpu, dtype, layout, ta, tb, k):
    """Shapes large enough for the 128 x 256 tile (after the RowMajor swap:
    f64 unless op(A) and op(B) are both stored K-contiguous, f32 with op(A)
    stored M-contiguous; the other cases run 128 x 128 tiles) with ragged
    edges in m, n and k: that tile
    computes every block of an edge tile instead of skipping those past the
    edge, and the k order of each C entry is the 128 x 128 / 64 x 64 tiles'
    one, so the single call agrees bit for bit with a grouped plan of the
    same problem (which runs 128 / 64 tiles); both meet the float64 bound on
    sampled rows."""
    from rblas_amd import grouped
    from tests.util import op
    m, n = 4000 + 3, 4100 + 29  # k = 5: one partial k-tile, staged through registers
    a, lda, b, ldb, c, ldc = make_problem(layout, ta, tb, m, n, k, 5, dtype, 9)
    c0 = c.copy()
    s = Stream()
    da, db, dc, dg = (DeviceArray.from_host(x, s) for x in (a, b, c, c))
    level3.gemm_dev(layout, ta, tb, m, n, k, 1.5, da, lda, db, ldb, -0.25, dc, ldc, s)
    plan = grouped.GroupedGemm([grouped.GemmProblem(layout, ta, tb, m, n, k, 1.5, da, lda, db,
                                                    ldb, -0.25, dg, ldc)], dtype)
    plan.run(s)
    s.synchronize()
    single, group = dc.download(s), dg.download(s)
    plan.close()
    s.close()
    assert np.array_equal(single, group)
    got = view(single, layout, m, n, ldc)
    A = op(a, layout, ta, m, k, lda).astype(np.float64)
    B = op(b, layout, tb, k, n, ldb).astype(np.float64)
    C0 = view(c0, layout, m, n, ldc).astype(np.float64)
    rows = np.concatenate([np.random.default_rng(1).choice(m, 40, replace=False),
                           np.arange(m - 8, m)])  # the last row tile too
    want = 1.5 * (A[rows] @ B) - 0.25 * C0[rows]
    mag = 1.5 * (np.abs(A[rows]) @ np.abs(B)) + 0.25 * np.abs(C0[rows])
    assert_gemm_close(got[rows], want, mag, k, dtype, what="wide tile")
