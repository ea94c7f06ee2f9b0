"""Seeded synthetic operands for the benchmark and test workloads.

One documented generator shared by every consumer (tests, bench, serving):
counter-based splitmix64. Element i of a draw with seed s and stream offset o
is mix(s + (o + i + 1) * 0x9E3779B97F4A7C15), where mix is splitmix64's
finaliser. Operands of one problem are drawn in the order A, B, C from
consecutive offsets.

  uniform(-1, 1): f32 from the top 24 bits, f64 from the top 53 bits,
                  value = 2*u - 1 (exact in the target type).
  normal(0, 1):   Box-Muller in float64 on consecutive pairs
                  (u1 = (hi53 + 1) * 2^-53, u2 = hi53 * 2^-53), cast to T.
"""
from __future__ import annotations

import concurrent.futures as cf
import os

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, offset: int, count: int) -> np.ndarray:
    """`count` outputs of the counter-based splitmix64 stream."""
    with np.errstate(over="ignore"):
        ctr = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + ctr * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


_CHUNK = 1 << 22  # elements per worker task (even, so normal pairs stay aligned)


def _parallel(fn, seed: int, offset: int, count: int, dtype) -> np.ndarray | None:
    """fn's draw computed in chunks on host threads (numpy's loops release
    the GIL). The stream is counter-based, so chunk j is the draw at offset
    + j * _CHUNK and the result equals the one-piece draw. None for small
    draws."""
    if count <= 2 * _CHUNK:
        return None
    out = np.empty(count, dtype)
    starts = range(0, count, _CHUNK)

    def one(c0):
        c1 = min(count, c0 + _CHUNK)
        out[c0:c1] = fn(seed, offset + c0, c1 - c0, dtype, _whole=True)

    workers = min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(one, starts))
    return out


def uniform(seed: int, offset: int, count: int, dtype, _whole: bool = False) -> np.ndarray:
    if not _whole:
        big = _parallel(uniform, seed, offset, count, dtype)
        if big is not None:
            return big
    z = splitmix64(seed, offset, count)
    if np.dtype(dtype) == np.float32:
        u = (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
        return (np.float32(2) * u - np.float32(1)).astype(np.float32)
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (2.0 * u - 1.0).astype(dtype)


def normal(seed: int, offset: int, count: int, dtype, _whole: bool = False) -> np.ndarray:
    if not _whole:
        big = _parallel(normal, seed, offset, count, dtype)
        if big is not None:
            return big
    pairs = (count + 1) // 2
    z = splitmix64(seed, offset, 2 * pairs)
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * pairs, np.float64)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:count].astype(dtype)


def stored_shape(trans: bool, rows: int, cols: int) -> tuple[int, int]:
    """Stored (rows, cols) of an operand whose op() is rows x cols."""
    return (cols, rows) if trans else (rows, cols)


def slice_len(layout_row_major: bool, rows: int, cols: int, ld: int) -> int:
    """Length rblas requires for a stored rows x cols matrix (level3.rs:121-124)."""
    if rows == 0 or cols == 0:
        return 0
    return (rows - 1) * ld + cols if layout_row_major else (cols - 1) * ld + rows
