"""The benchmark workloads at their full sizes, checked through the C ABI.

Each case runs the exact configuration the benchmark times and checks the
result against float64 (complex128) on sampled rows and columns:

  plain gemm    m = n = k in {4096, 8192}, f32 / f64, every Layout x
                {NoTrans, Trans}^2 combination, stored ld = rows + 7,
                alpha 1.5, beta -0.25, N(0,1) operands (splitmix64 seed 1,
                drawn A, B, C), HBM-resident. Bound, elementwise:
                |dC| <= 2 (k+2) u (|alpha||A||B| + |beta||C0|).
  C64 4M gemm   4096^3, re / im i.i.d. N(0,1), ld = rows + 7, alpha
                1.5 - 0.5i, beta -0.25 + 0.125i, (N, N) and (ConjTrans,
                Trans). Bound: 8 (k+2) u (|alpha||A||B| + |beta||C0|).
  precision     SGEMM (alpha 1, beta 0) in FP32, BF16x3 and BF16x6 on
                (a) 8192^3 N(0,1), (b) the same with row i of A scaled by
                2^U[-20,20], (c) m = n = 1024, k = 65536. For each mode
                e = max over 48 sampled rows of |C - C64| / (|A||B|); the
                split-precision target is e(BF16x3) <= 2 e(FP32), and
                BF16x6 and FP32 stay within the f32 bound 2 (k+2) u.
  serving       the serving mix as drawn (gemm dims log-uniform to 4096,
                trsm order to 2048 with up to 512 right-hand sides, C64
                dims to 2048), plus one request at each maximum size,
                under Poisson arrivals so batches overlap; captured results
                against float64 (gemm: 2 (k+2) u bound on sampled rows and
                columns; trsm: residual 16 m u (|A||X| + |B|); complex:
                8 (k+2) u), and each captured request's hash equal alone.

u is the unit roundoff (2^-24 f32, 2^-53 f64).
"""
import numpy as np
import pytest

from rblas_amd import _ffi, complex3, level3, serve, workloads
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import Compute
from tests.util import COL, ROW, assert_gemm_close, op, view

pytestmark = pytest.mark.gpu

VARIANTS = [(layout, ta, tb) for layout in (COL, ROW) for ta in (0, 1) for tb in (0, 1)]
ALPHA, BETA = 1.5, -0.25
NSAMPLE = 48


class _Square:
    """One set of resident square operands (A, B, C0) of order n, ld n + 7:
    for a square problem every layout / transpose variant stores each
    operand as n x n with that ld, so one draw serves all eight."""

    def __init__(self, dtype, n):
        self.dtype, self.n, self.ld = np.dtype(dtype), n, n + 7
        self.len = (n - 1) * self.ld + n
        self.host = [workloads.normal(1, i * self.len, self.len, dtype) for i in range(3)]
        self.stream = Stream()
        self.dev = [DeviceArray.from_host(x, self.stream) for x in self.host]
        self.c = DeviceArray(self.len, dtype)
        self.stream.synchronize()

    def reset_c(self):
        st = _ffi.lib().rblas_memcpy_dtod(self.c.ptr, self.dev[2].ptr, self.c.nbytes,
                                          self.stream.handle)
        assert st == 0

    def close(self):
        for x in (*self.dev, self.c):
            x.free()
        self.stream.close()


@pytest.fixture(scope="module")
def square():
    held = {}

    def get(dtype, n):
        key = (np.dtype(dtype).name, n)
        if key not in held:
            for v in held.values():
                v.close()
            held.clear()
            held[key] = _Square(dtype, n)
        return held[key]

    yield get
    for v in held.values():
        v.close()


def sampled_check(got, A, B, C0, alpha, beta, k, dtype, c_const, seed=0, what=""):
    """Rows and columns of got (m x n) against float64 / complex128 on
    NSAMPLE sampled rows and NSAMPLE sampled columns, within
    c_const (k+2) u (|alpha||A||B| + |beta||C0|). A, B are op(A), op(B)."""
    wide = np.complex128 if np.iscomplexobj(got) else np.float64
    m, n = got.shape
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(m, min(NSAMPLE, m), replace=False))
    cols = np.sort(rng.choice(n, min(NSAMPLE, n), replace=False))
    Bw = np.asarray(B, dtype=wide)
    Ar = np.asarray(A[rows], dtype=wide)
    want = alpha * (Ar @ Bw) + beta * np.asarray(C0[rows], dtype=wide)
    mag = abs(alpha) * (np.abs(Ar) @ np.abs(Bw)) + abs(beta) * np.abs(np.asarray(C0[rows], wide))
    worst = assert_gemm_close(np.asarray(got[rows]), want, mag, k, dtype, c_const, what + " rows")
    del Bw
    Aw = np.asarray(A, dtype=wide)
    Bc = np.asarray(B[:, cols], dtype=wide)
    C0c = np.asarray(C0[:, cols], dtype=wide)
    want = alpha * (Aw @ Bc) + beta * C0c
    mag = abs(alpha) * (np.abs(Aw) @ np.abs(Bc)) + abs(beta) * np.abs(C0c)
    worst = max(worst, assert_gemm_close(np.asarray(got[:, cols]), want, mag, k, dtype, c_const,
                                         what + " cols"))
    return worst


@pytest.mark.parametrize("layout,ta,tb", VARIANTS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [4096, 8192])
def test_plain_gemm_full_size(gpu, square, dtype, n, layout, ta, tb):
    sq = square(dtype, n)
    print(f"gemm {np.dtype(dtype).name} n={n} layout={layout} ta={ta} tb={tb}", flush=True)
    sq.reset_c()
    ld = sq.ld
    level3.gemm_dev(layout, ta, tb, n, n, n, ALPHA, sq.dev[0], ld, sq.dev[1], ld, BETA, sq.c, ld,
                    sq.stream)
    sq.stream.synchronize()
    got = view(sq.c.download(sq.stream), layout, n, n, ld)
    a, b, c0 = sq.host
    A = op(a, layout, ta, n, n, ld)
    B = op(b, layout, tb, n, n, ld)
    C0 = view(c0, layout, n, n, ld)
    worst = sampled_check(got, A, B, C0, ALPHA, BETA, n, dtype, 2.0,
                          seed=layout * 4 + ta * 2 + tb, what=f"{n} {layout}{ta}{tb}")
    print(f"  max err / bound {worst:.3g}")


@pytest.mark.parametrize("ta,tb", [(0, 0), (2, 1)])
def test_c64_4m_gemm_4096(gpu, ta, tb):
    n = 4096
    ld = n + 7
    ln = (n - 1) * ld + n
    a, b, c = (workloads.normal(1, 2 * i * ln, 2 * ln, np.float64).view(np.complex128)
               for i in range(3))
    alpha, beta = 1.5 - 0.5j, -0.25 + 0.125j
    print(f"zgemm 4096 ta={ta} tb={tb}", flush=True)
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    complex3.gemm_dev(ta, tb, n, n, n, alpha, da, ld, db, ld, beta, dc, ld, s)
    s.synchronize()
    got = view(dc.download(s), COL, n, n, ld)
    for x in (da, db, dc):
        x.free()
    s.close()
    A = op(a, COL, ta, n, n, ld)
    B = op(b, COL, tb, n, n, ld)
    worst = sampled_check(got, A, B, view(c, COL, n, n, ld), alpha, beta, n, np.float64, 8.0,
                          seed=ta + tb, what="zgemm 4096")
    print(f"  max err / bound {worst:.3g}")


def _precision_operands(case):
    if case == "c":
        m = n = 1024
        k = 65536
    else:
        m = n = k = 8192
    A = workloads.normal(11, 0, m * k, np.float32)
    B = workloads.normal(11, m * k, k * n, np.float32)
    if case == "b":
        scale = np.exp2(np.random.default_rng(3).uniform(-20, 20, m)).astype(np.float32)
        A = (A.reshape(k, m) * scale[None, :]).reshape(-1)  # row i of A scaled
    return m, n, k, A, B


@pytest.mark.parametrize("case", ["a", "b", "c"])
def test_split_sgemm_precision_target(gpu, case):
    m, n, k, A, B = _precision_operands(case)
    print(f"precision case {case}: {m} x {n} x {k}", flush=True)
    rows = np.sort(np.random.default_rng(5).choice(m, NSAMPLE, replace=False))
    Am = A.reshape(k, m).T[rows].astype(np.float64)
    Bm = B.reshape(n, k).T.astype(np.float64)
    ref = Am @ Bm
    mag = np.abs(Am) @ np.abs(Bm)
    s = Stream()
    da, db = DeviceArray.from_host(A, s), DeviceArray.from_host(B, s)
    dc = DeviceArray(m * n, np.float32)
    err = {}
    for comp in (Compute.F32, Compute.BF16X3, Compute.BF16X6):
        level3.sgemm_ex_dev(COL, 0, 0, m, n, k, 1.0, da, m, db, k, 0.0, dc, m, comp, s)
        s.synchronize()
        C = dc.download(s).reshape(n, m).T[rows].astype(np.float64)
        e = np.abs(C - ref)
        err[comp] = float((e / mag).max())
        if comp != Compute.BF16X3:
            assert_gemm_close(C, ref, mag, k, np.float32, 2.0, f"{comp.name} case {case}")
    for x in (da, db, dc):
        x.free()
    s.close()
    print({c.name: f"{e:.3e}" for c, e in err.items()},
          f"BF16X3/FP32 {err[Compute.BF16X3] / err[Compute.F32]:.3f}")
    assert err[Compute.BF16X3] <= 2.0 * err[Compute.F32]
    assert err[Compute.BF16X6] <= 2.0 * err[Compute.F32]


def _serving_trace():
    """The serving mix drawn as the benchmark draws it (seed 7 stream, rate
    400 requests/s), plus one request at each maximum size."""
    req = serve.workload(36, 400.0, seed=7)
    big = [(serve.GEMM, 0, 4096, 4096, 4096), (serve.GEMM, 1, 4096, 4096, 4096),
           (serve.TRSM, 1, 2048, 512, 0), (serve.ZGEMM, 1, 2048, 2048, 2048)]
    for j, (kind, dt, m, n, k) in enumerate(big):
        r = req[3 + 8 * j]
        r["kind"], r["dtype"], r["m"], r["n"], r["k"] = kind, dt, m, n, k
    return req


def _check_request(req, got):
    kind, m, n, k = (int(req[f]) for f in ("kind", "m", "n", "k"))
    if kind == serve.TRSM:
        A, B = serve.operands(req)
        A = A.reshape(m, m).T.astype(np.float64)
        X = got.reshape(n, m).T.astype(np.float64)
        B = B.reshape(n, m).T.astype(np.float64)
        u = 2.0 ** -24 if got.dtype == np.float32 else 2.0 ** -53
        r = np.abs(A @ X - B)
        assert (r <= 16 * m * u * (np.abs(A) @ np.abs(X) + np.abs(B)) + 1e-300).all()
        return
    A, B, C0 = serve.operands(req)
    cplx = kind == serve.ZGEMM
    dtype = np.float64 if cplx else got.dtype
    sampled_check(got.reshape(n, m).T, A.reshape(k, m).T, B.reshape(n, k).T, C0.reshape(n, m).T,
                  1.0, 0.5, k, dtype, 8.0 if cplx else 2.0, what=f"serve kind {kind}")


def test_serving_mix_full_sizes(gpu):
    req = _serving_trace()
    assert req["m"].max() == 4096 and (req["m"][req["kind"] == serve.TRSM] == 2048).any()
    assert (req["n"][req["kind"] == serve.TRSM] == 512).any()
    assert (req["m"][req["kind"] == serve.ZGEMM] == 2048).any()
    picks = sorted({3, 11, 19, 27} | set(range(0, len(req), 5)))
    cap = {i: serve.result_buffer(req[i]) for i in picks}
    print(f"serving {len(req)} requests, capturing {len(picks)}", flush=True)
    res, stats = serve.run(req, serve.BATCHED, capture=cap)
    assert (res["done_s"] >= res["arrival_s"]).all()
    print(f"  batches {stats['batches']} max in flight {stats['max_inflight_seen']}", flush=True)
    for i in picks:
        _check_request(req[i], cap[i])
    for i in (3, 11, 19, 27):  # the maximum-size requests alone
        one = req[i:i + 1].copy()
        one["arrival_s"] = 0.0
        r1, _ = serve.run(one, serve.BATCHED)
        assert r1["hash"][0] == res["hash"][i], i
