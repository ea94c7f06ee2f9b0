"""The seeded operand generator: a large draw computed in chunks on host
threads equals the same draw computed in one piece (the stream is
counter-based), for odd offsets and lengths that end mid-chunk."""
import numpy as np
import pytest

from rblas_amd import workloads


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("gen", ["normal", "uniform"])
def test_chunked_draw_equals_one_piece(gen, dtype):
    fn = getattr(workloads, gen)
    count = 2 * workloads._CHUNK + 12345
    for offset in (0, 7):
        whole = fn(3, offset, count, dtype, _whole=True)
        chunked = fn(3, offset, count, dtype)
        assert np.array_equal(whole, chunked)
    # small draws take the one-piece path
    assert np.array_equal(fn(3, 5, 1001, dtype), fn(3, 5, 1001, dtype, _whole=True))
