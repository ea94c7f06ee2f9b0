"""Level-1 dot / nrm2 / axpy through the C ABI.

CPU part: rblas's panic conditions (restated in the oracle) on a sweep of
argument combinations; no device is touched on those paths.
GPU part: rblas's own KATs (level1.rs:169-228), the oracle and float64 on
random data, strides, and special values. Tolerances (written here):
  dot:  |d - exact| <= (n+2) u sum|x_i y_i|
  nrm2: relative 4 (n+2) u (f64); f32 is summed in f64: 2u relative
  axpy: bit-exact (each element is one fma, like rblas).
"""
import itertools
import math

import numpy as np
import pytest

from rblas_amd import level1
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic


def _sweep():
    rng = np.random.default_rng(99)
    for _ in range(300):
        n = int(rng.integers(0, 6))
        incx, incy = (int(v) for v in rng.integers(0, 4, 2))
        xl, yl = (int(v) for v in rng.integers(0, 14, 2))
        alpha = float(rng.choice([0.0, 1.5]))
        yield n, incx, incy, xl, yl, alpha


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_level1_panics_match_oracle(oracle, dtype):
    checked = 0
    for n, incx, incy, xl, yl, alpha in _sweep():
        x, y = np.zeros(xl, dtype), np.zeros(yl, dtype)
        if oracle.axpy_panics(n, alpha, xl, incx, yl, incy):
            with pytest.raises(BlasPanic):
                level1.axpy(n, alpha, x, incx, y, incy)
            checked += 1
        if oracle.dot_panics(n, xl, incx, yl, incy):
            with pytest.raises(BlasPanic):
                level1.dot(n, x, incx, y, incy)
            checked += 1
        if oracle.nrm2_panics(n, xl, incx):
            with pytest.raises(BlasPanic):
                level1.nrm2(n, x, incx)
            checked += 1
    assert checked > 200


def test_level1_noops_need_no_device():
    # axpy with alpha == 0 returns before any check (level1.rs:18-20)
    x = np.zeros(1)
    y = np.full(1, 3.0)
    level1.axpy(10, 0.0, x, 0, y, 0)
    assert y[0] == 3.0
    assert level1.dot(0, x, 0, x, 0) == 0.0
    assert level1.nrm2(0, x, 0) == 0.0


@pytest.mark.gpu
def test_level1_reference_kats(gpu):
    # level1.rs:169-228
    x = np.array([1.0, 2, 3, 4], np.float32)
    y = np.array([10.0, 20, 30, 40], np.float32)
    level1.axpy(4, 2.0, x, 1, y, 1)
    assert list(y) == [12, 24, 36, 48]
    x = np.array([1.0, 0, 2, 0, 3])
    y = np.array([10.0, 20, 30])
    level1.axpy(3, 1.0, x, 2, y, 1)
    assert list(y) == [11, 22, 33]
    assert level1.dot(3, np.array([1.0, 2, 3], np.float32), 1,
                      np.array([4.0, 5, 6], np.float32), 1) == 32.0
    assert abs(level1.nrm2(2, np.array([3.0, 4], np.float32), 1) - 5) < 1e-6
    big = 1e200
    got = level1.nrm2(3, np.array([big] * 3), 1)
    assert math.isfinite(got) and abs(got - big * math.sqrt(3)) / (big * math.sqrt(3)) < 1e-12
    v = np.array([1.0, -2, 3, -4, 5])
    assert abs(level1.nrm2(5, v, 1) - math.sqrt(55)) < 1e-12
    assert abs(level1.nrm2(2, np.array([3.0, 99, 4]), 2) - 5) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 1), (1, 3), (3, 2)])
def test_level1_vs_oracle(gpu, oracle, dtype, incx, incy):
    rng = np.random.default_rng(7 + incx * 3 + incy)
    u = np.finfo(dtype).eps / 2
    for n in (1, 3, 17, 1000, 65537, 1 << 20):
        x = rng.standard_normal((n - 1) * incx + 1).astype(dtype)
        y = rng.standard_normal((n - 1) * incy + 1).astype(dtype)
        xs, ys = x[::incx][:n].astype(np.float64), y[::incy][:n].astype(np.float64)
        d = level1.dot(n, x, incx, y, incy)
        exact = float(xs @ ys)
        bound = (n + 2) * u * float(np.abs(xs) @ np.abs(ys))
        assert abs(d - exact) <= bound
        assert abs(oracle.dot(n, x, incx, y, incy) - exact) <= bound
        nr = level1.nrm2(n, x, incx)
        want = float(np.sqrt((xs ** 2).sum()))
        assert abs(nr - want) <= 4 * (n + 2) * u * want
        y_gpu, y_ref = y.copy(), y.copy()
        level1.axpy(n, 0.75, x, incx, y_gpu, incy)
        oracle.axpy(n, 0.75, x, incx, y_ref, incy)
        assert np.array_equal(y_gpu, y_ref)  # both: one fma per element


@pytest.mark.gpu
def test_nrm2_extremes(gpu):
    tiny = np.full(1000, 1e-300)
    assert abs(level1.nrm2(1000, tiny, 1) - 1e-300 * math.sqrt(1000)) <= 1e-12 * 1e-300 * 32
    huge = np.full(1000, 1e300)
    assert abs(level1.nrm2(1000, huge, 1) / (1e300 * math.sqrt(1000)) - 1) < 1e-12
    mixed = np.array([1e300, 1.0, 1e-300, 3e300])
    assert abs(level1.nrm2(4, mixed, 1) / (math.sqrt(10) * 1e300) - 1) < 1e-12
    assert math.isnan(level1.nrm2(3, np.array([1.0, np.nan, 2.0]), 1))
    assert level1.nrm2(2, np.array([np.inf, 1.0]), 1) == math.inf
    f = np.full(4, 1e38, np.float32)  # squares overflow f32, not the f64 sum
    assert abs(level1.nrm2(4, f, 1) - 2e38) / 2e38 < 1e-6


NONFINITE_CASES = {
    "inf": ([np.inf], "inf"),
    "inf_finite": ([3.0, np.inf, 4.0], "inf"),
    "neg_inf": ([1.0, -np.inf], "inf"),
    "two_inf": ([np.inf, np.inf], "nan"),
    "inf_1_neginf": ([np.inf, 1.0, -np.inf], "nan"),
    "three_inf": ([-np.inf, 2.0, np.inf, 0.0, np.inf], "nan"),
    "nan": ([1.0, np.nan, 2.0], "nan"),
    "inf_nan": ([np.inf, np.nan], "nan"),
    "nan_inf": ([0.0, np.nan, 5.0, -np.inf], "nan"),
    "zeros": ([0.0, 0.0], "zero"),
}


def nonfinite_vector(vals, dtype, n, spread):
    """vals placed at spread positions of an n-vector of N(0,1) entries
    (far apart, so they land in different workgroups of the device
    reduction), or the bare values when n == len(vals)."""
    if n == len(vals):
        return np.array(vals, dtype)
    x = np.random.default_rng(n).standard_normal(n).astype(dtype)
    pos = np.linspace(0, n - 1, len(vals)).astype(np.int64) if spread else np.arange(len(vals))
    x[pos] = vals
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", sorted(NONFINITE_CASES))
def test_nrm2_nonfinite_matches_oracle(gpu, oracle, dtype, case):
    """level1.rs:96-123 on non-finite input: any NaN gives NaN, a second
    infinity gives inf / inf = NaN at level1.rs:116, one infinity gives +Inf.
    The GPU must return the oracle's class of result on short vectors,
    strided vectors and long vectors whose special entries sit in different
    workgroups."""
    vals, want = NONFINITE_CASES[case]
    for n, incx, spread in ((len(vals), 1, False), (4099, 1, True), (2_000_003, 1, True),
                            (1001, 3, True)):
        base = nonfinite_vector(vals, dtype, n, spread)
        if incx > 1:
            x = np.full((n - 1) * incx + 1, np.nan, dtype)  # gaps hold NaN, never read
            x[::incx] = base
        else:
            x = base
        got = level1.nrm2(n, x, incx)
        ref = oracle.nrm2(n, x, incx)
        if want == "nan":
            assert math.isnan(got) and math.isnan(ref), (n, incx, got, ref)
        elif want == "inf":
            assert got == math.inf and ref == math.inf, (n, incx, got, ref)
        elif n == len(vals):
            assert got == 0 and ref == 0
        else:
            u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
            assert abs(got - ref) <= 4 * (n + 2) * u * ref, (n, incx, got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_level1_device_entries(gpu, dtype):
    rng = np.random.default_rng(5)
    n = 300001
    x = rng.standard_normal(n).astype(dtype)
    y = rng.standard_normal(n).astype(dtype)
    s = Stream()
    dx, dy = DeviceArray.from_host(x, s), DeviceArray.from_host(y, s)
    r = DeviceArray(1, dtype)
    level1.dot_dev(n, dx, 1, dy, 1, r, s)
    s.synchronize()
    assert r.download(s)[0] == np.array(level1.dot(n, x, 1, y, 1), dtype)
    level1.nrm2_dev(n, dx, 1, r, s)
    s.synchronize()
    assert r.download(s)[0] == np.array(level1.nrm2(n, x, 1), dtype)
    level1.axpy_dev(n, 1.25, dx, 1, dy, 1, s)
    s.synchronize()
    yh = y.copy()
    level1.axpy(n, 1.25, x, 1, yh, 1)
    assert np.array_equal(dy.download(s), yh)
    # n == 0: rblas returns 0; the device result is stored as 0 (over a NaN).
    r = DeviceArray.from_host(np.array([np.nan], dtype), s)
    level1.dot_dev(0, dx, 1, dy, 1, r, s)
    s.synchronize()
    assert r.download(s)[0] == 0
    r = DeviceArray.from_host(np.array([np.nan], dtype), s)
    level1.nrm2_dev(0, dx, 1, r, s)
    s.synchronize()
    assert r.download(s)[0] == 0
    s.close()
