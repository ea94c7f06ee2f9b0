"""Serving loop (continuous batching with HBM admission).

Checks, with the tolerances written here:
  - results: captured results of gemm / trsm / complex-gemm requests against
    float64 references built from the same operands regenerated on the host
    (gemm: 2 (k+2) u |A||B| + |beta||C0| elementwise, complex: 8 (k+2) u, trsm
    residual 16 m u (|A||X| + |B|));
  - each request's result is the same, bit for bit (64-bit hash of its bits),
    alone and in the mix, for continuous batching;
  - the admission cap holds and every request completes.
"""
import numpy as np
import pytest

from rblas_amd import serve, workloads


def test_device_generator_is_the_documented_one():
    """The host regeneration used by these tests follows workloads.py; the
    trsm matrix is lower triangular with the documented entries."""
    req = np.zeros(1, serve.REQUEST_DTYPE)[0]
    req["kind"], req["dtype"], req["m"], req["n"], req["seed"] = serve.TRSM, 1, 5, 2, 99
    A, B = serve.operands(req)
    M = A.reshape(5, 5).T
    assert np.allclose(np.triu(M, 1), 0) and (np.diag(M) >= 2).all()
    u = workloads.uniform(99, 0, 25, np.float64).reshape(5, 5).T
    assert M[3, 1] == u[3, 1] / np.sqrt(5)
    assert np.array_equal(B, workloads.uniform(99, 25, 10, np.float64))


def test_serve_rejects_malformed_requests():
    bad = serve.workload(4, 0.0)
    bad["m"][2] = 0
    with pytest.raises(ValueError):
        serve.run(bad)
    unordered = serve.workload(4, 10.0)
    unordered["arrival_s"][3] = 0.0
    with pytest.raises(ValueError):
        serve.run(unordered)
    big = serve.workload(1, 0.0)
    big["kind"], big["m"], big["n"], big["k"] = serve.GEMM, 4096, 4096, 4096
    with pytest.raises(ValueError):
        serve.run(big, hbm_cap=1e6)
    with pytest.raises(ValueError):
        serve.run(serve.workload(2, 0.0), hash_col_step=-1)


def _small_trace(count, seed, rate):
    rng = np.random.default_rng(seed)
    req = serve.workload(count, rate, seed)
    # smaller sizes so the host references stay quick
    g = req["kind"] != serve.TRSM
    for f in ("m", "n", "k"):
        req[f][g] = np.floor(np.exp(rng.uniform(np.log(8), np.log(400), g.sum())))
    t = req["kind"] == serve.TRSM
    req["m"][t] = rng.integers(1, 700, t.sum())
    req["n"][t] = rng.integers(1, 60, t.sum())
    return req


def _check(req, got):
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
    wide = np.complex128 if cplx else np.float64
    Am = A.reshape(k, m).T.astype(wide)
    Bm = B.reshape(n, k).T.astype(wide)
    C0m = C0.reshape(n, m).T.astype(wide)
    want = Am @ Bm + 0.5 * C0m
    mag = np.abs(Am) @ np.abs(Bm) + 0.5 * np.abs(C0m)
    u = 2.0 ** -24 if got.dtype == np.float32 else 2.0 ** -53
    c = 8 if cplx else 2
    err = np.abs(got.reshape(n, m).T.astype(wide) - want)
    assert (err <= c * (k + 2) * u * mag + 1e-300).all()


@pytest.mark.gpu
def test_serve_results_correct_and_same_alone_and_in_mix(gpu):
    req = _small_trace(120, 3, rate=2000.0)
    picks = [i for i in range(120) if i % 9 == 0]
    cap = {i: serve.result_buffer(req[i]) for i in picks}
    res, stats = serve.run(req, serve.BATCHED, max_inflight=3, max_batch=16, capture=cap)
    assert (res["done_s"] >= res["arrival_s"]).all() and stats["batches"] >= 2
    for i in picks:
        _check(req[i], cap[i])
    for i in picks:  # alone: a trace of one request, same seed and shape
        one = req[i:i + 1].copy()
        one["arrival_s"] = 0.0
        r1, _ = serve.run(one, serve.BATCHED)
        assert r1["hash"][0] == res["hash"][i], i


@pytest.mark.gpu
def test_serve_plain_mode_correct(gpu):
    req = _small_trace(40, 4, rate=0.0)
    picks = list(range(0, 40, 5))
    cap = {i: serve.result_buffer(req[i]) for i in picks}
    res, stats = serve.run(req, serve.PLAIN, capture=cap)
    assert stats["batches"] == 40 and stats["max_inflight_seen"] == 1
    for i in picks:
        _check(req[i], cap[i])


@pytest.mark.gpu
def test_serve_admission_cap_holds(gpu):
    req = _small_trace(60, 5, rate=0.0)
    res, stats = serve.run(req, serve.BATCHED, hbm_cap=1e18)
    cap = float(res["footprint_bytes"].max()) * 1.5  # the largest request fits, barely
    res2, stats2 = serve.run(req, serve.BATCHED, max_inflight=3, hbm_cap=cap)
    assert stats2["peak_footprint_bytes"] <= cap
    assert stats2["batches"] > stats["batches"]  # the cap split the work
    assert np.array_equal(res2["hash"], res["hash"])  # same bits under any admission


@pytest.mark.gpu
def test_serve_sampled_hash_same_alone_and_in_mix(gpu):
    """hash_col_step > 1 fingerprints every hash_col_step-th column: the same
    bits alone and in the mix, and a different fingerprint from the full one
    for a result with more than one column."""
    req = _small_trace(40, 6, rate=3000.0)
    res, _ = serve.run(req, serve.BATCHED, max_inflight=3, max_batch=8, hash_col_step=3)
    full, _ = serve.run(req, serve.BATCHED, max_inflight=3, max_batch=8)
    for i in range(0, 40, 7):
        one = req[i:i + 1].copy()
        one["arrival_s"] = 0.0
        r1, _ = serve.run(one, serve.BATCHED, hash_col_step=3)
        assert r1["hash"][0] == res["hash"][i], i
        if int(req["n"][i]) > 1:
            assert res["hash"][i] != full["hash"][i], i
