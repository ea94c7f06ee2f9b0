"""rblas_dgemm_dev_noalloc: the DGEMM device entry point whose k-split
workspace is a fixed device array instead of a pool allocation. Through the
C ABI its result must equal rblas_dgemm_dev's bit for bit (same launch
shape, same summation order), on consecutive calls that share the array,
the second one smaller (so stale partials or counters of the first would
show), and on shapes with and without a k split."""
import numpy as np
import pytest

from rblas_amd import level3, workloads
from rblas_amd.device import DeviceArray, Stream
from tests.util import COL, ROW, assert_gemm_close, gemm_expect, slice_len, view

pytestmark = pytest.mark.gpu


def _operands(layout, ta, tb, m, n, k, pad, seed, s):
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    lda = (ac if layout == ROW else ar) + pad
    ldb = (bc if layout == ROW else br) + pad
    ldc = (n if layout == ROW else m) + pad
    la, lb, lc = (slice_len(layout, ar, ac, lda), slice_len(layout, br, bc, ldb),
                  slice_len(layout, m, n, ldc))
    a = workloads.normal(seed, 0, la, np.float64)
    b = workloads.normal(seed, la, lb, np.float64)
    c = workloads.normal(seed, la + lb, lc, np.float64)
    dev = [DeviceArray.from_host(x, s) for x in (a, b, c)]
    return (a, b, c), dev, (lda, ldb, ldc)


# (layout, ta, tb, m, n, k), in call order: the largest workspace the f64
# tile kernel uses (255 tiles of 128^2, 2 splits: 63.75 MiB of partials),
# then smaller k-split shapes at 128 and 64 tiles, one without a split, and
# the first shapes again.
SEQUENCE = [
    (COL, 0, 0, 1920, 2176, 4096),    # 255 tiles of 128^2, 2 splits
    (COL, 0, 0, 1024, 1024, 65536),   # 64 tiles of 128^2, 4 splits (32 MiB of partials)
    (ROW, 1, 0, 200, 190, 8192),      # 64-tiles, k split
    (COL, 0, 1, 300, 260, 517),       # no split
    (COL, 1, 1, 640, 520, 20000),     # 128-tiles, k split, ragged edges
]


def test_noalloc_matches_dgemm_dev_on_consecutive_calls(gpu):
    s = Stream()
    for i, (layout, ta, tb, m, n, k) in enumerate(SEQUENCE + SEQUENCE[:2]):
        print(f"call {i}: layout {layout} ta {ta} tb {tb} {m} x {n} x {k}", flush=True)
        host, (da, db, dc), (lda, ldb, ldc) = _operands(layout, ta, tb, m, n, k, 3, 40 + i, s)
        dc_ref = DeviceArray.from_host(host[2], s)
        level3.gemm_dev(layout, ta, tb, m, n, k, 1.5, da, lda, db, ldb, -0.25, dc_ref, ldc, s)
        level3.dgemm_dev_noalloc(layout, ta, tb, m, n, k, 1.5, da, lda, db, ldb, -0.25, dc, ldc, s)
        s.synchronize()
        got, ref = dc.download(s), dc_ref.download(s)
        assert np.array_equal(got, ref), i
        want, mag = gemm_expect(layout, ta, tb, m, n, k, 1.5, host[0], lda, host[1], ldb, -0.25,
                                host[2], ldc)
        assert_gemm_close(view(got, layout, m, n, ldc), want, mag, k, np.float64,
                          what=f"noalloc call {i}")
        for x in (da, db, dc, dc_ref):
            x.free()
    s.close()


def test_noalloc_validation_and_no_ops(gpu):
    s = Stream()
    a = DeviceArray(64, np.float64)
    c = DeviceArray.from_host(np.full(64, 3.0), s)
    with pytest.raises(ValueError):  # overlapping C columns
        level3.dgemm_dev_noalloc(COL, 0, 0, 8, 4, 2, 1.0, a, 8, a, 2, 0.0, c, 7, s)
    level3.dgemm_dev_noalloc(COL, 0, 0, 8, 4, 0, 1.0, a, 8, a, 1, 1.0, c, 8, s)  # k 0, beta 1
    level3.dgemm_dev_noalloc(COL, 0, 0, 8, 4, 0, 1.0, a, 8, a, 1, 2.0, c, 8, s)  # beta only
    s.synchronize()
    assert (c.download(s)[:32] == 6.0).all() and (c.download(s)[32:] == 3.0).all()
    with pytest.raises(TypeError):
        level3.dgemm_dev_noalloc(COL, 0, 0, 8, 4, 2, 1.0, DeviceArray(64, np.float32), 8,
                                 a, 2, 0.0, c, 8, s)
    s.close()
