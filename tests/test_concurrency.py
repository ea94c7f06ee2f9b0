"""Concurrent callers.

rblas's functions are reentrant and callable from any thread (`Float: Send +
Sync`, lib.rs:70; concurrent calls are allowed), so the C ABI must be too.
Eight threads call the host-slice entry points (gemm without and with a k
split, trsm, dot, complex gemm) and a `_dev` gemm on a stream of their own,
three rounds each, all at once. Every result must equal, bit for bit, the
same call made alone beforehand: the kernels are deterministic and their
launch configuration depends on the shape only, so any difference means two
calls shared a workspace, a counter or a stream.
"""
import threading

import numpy as np
import pytest

from rblas_amd import complex3, level1, level3
from rblas_amd.device import DeviceArray, Stream

THREADS = 8
ROUNDS = 3


def _inputs(seed):
    rng = np.random.default_rng(100 + seed)
    m, n, k = 150 + seed, 97 + 2 * seed, 260 + 3 * seed
    g = dict(m=m, n=n, k=k, a=rng.standard_normal((m + 3) * k), b=rng.standard_normal((k + 1) * n),
             c=rng.standard_normal((m + 5) * n))
    # 256 x 256 x 16384: four 128-tiles, so the k range is split.
    s = dict(a=rng.standard_normal(256 * 16384), b=rng.standard_normal(16384 * 256),
             c=rng.standard_normal(256 * 256))
    q, nr = 180, 40
    full = rng.standard_normal((q, q)) / np.sqrt(q)
    np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
    t = dict(q=q, nr=nr, a=np.tril(full).T.reshape(-1).copy(), b=rng.standard_normal(q * nr))
    d = dict(x=rng.standard_normal(100_000), y=rng.standard_normal(100_000))
    z = dict(a=(rng.standard_normal(40 * 50) + 1j * rng.standard_normal(40 * 50)),
             b=(rng.standard_normal(50 * 30) + 1j * rng.standard_normal(50 * 30)),
             c=(rng.standard_normal(40 * 30) + 1j * rng.standard_normal(40 * 30)))
    f = dict(a=rng.standard_normal(200 * 300).astype(np.float32),
             b=rng.standard_normal(300 * 150).astype(np.float32),
             c=rng.standard_normal(200 * 150).astype(np.float32))
    return g, s, t, d, z, f


def _run(inputs):
    g, s, t, d, z, f = inputs
    out = {}
    m, n, k = g["m"], g["n"], g["k"]
    c = g["c"].copy()
    level3.gemm(1, 0, 0, m, n, k, 1.5, g["a"], m + 3, g["b"], k + 1, -0.25, c, m + 5)
    out["gemm"] = c
    c = s["c"].copy()
    level3.gemm(1, 0, 0, 256, 256, 16384, 1.0, s["a"], 256, s["b"], 16384, 0.5, c, 256)
    out["gemm_split"] = c
    b = t["b"].copy()
    level3.trsm(0, 1, 0, 0, t["q"], t["nr"], 2.0, t["a"], t["q"], b, t["q"])
    out["trsm"] = b
    out["dot"] = np.float64(level1.dot(d["x"].size, d["x"], 1, d["y"], 1))
    c = z["c"].copy()
    complex3.gemm(0, 2, 40, 30, 50, 1 - 0.5j, z["a"], 40, z["b"], 30, 0.25j, c, 40)
    out["zgemm"] = c
    st = Stream()
    da, db, dc = (DeviceArray.from_host(x, st) for x in (f["a"], f["b"], f["c"]))
    level3.gemm_dev(1, 0, 0, 200, 150, 300, 1.0, da, 200, db, 300, 0.5, dc, 200, st)
    st.synchronize()
    out["sgemm_dev"] = dc.download(st)
    for x in (da, db, dc):
        x.free()
    st.close()
    return out


@pytest.mark.gpu
def test_concurrent_callers_match_lone_calls(gpu):
    inputs = [_inputs(i) for i in range(THREADS)]
    alone = [_run(x) for x in inputs]
    got = [[None] * ROUNDS for _ in range(THREADS)]
    errors = []

    def worker(i):
        try:
            for r in range(ROUNDS):
                got[i][r] = _run(inputs[i])
        except Exception as e:  # reported by the main thread
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(THREADS)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for i in range(THREADS):
        for r in range(ROUNDS):
            for key, want in alone[i].items():
                assert np.array_equal(got[i][r][key], want), (i, r, key)
