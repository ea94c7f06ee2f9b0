"""Grouped GEMM plans (one persistent launch over many gemm problems).

Every problem is rblas's gemm (level3.rs:32-62) with its own arguments, so
each result is checked like a single gemm: exact on small integers, and on
random data within |dC| <= 2 (k+2) u (|alpha||A||B| + |beta||C0|)
elementwise (u the unit roundoff). A problem's launch shape depends on its
own shape only, so its bits are the same alone, in any group, replayed
from a graph, and run again.
"""
import numpy as np
import pytest

from rblas_amd import grouped, level3
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic
from tests.util import COL, ROW, assert_gemm_close, gemm_expect, slice_len, view


class _Host:
    """A stand-in operand for validation-only tests (never dereferenced)."""

    def __init__(self, size, dtype):
        self.ptr, self.size, self.dtype = 0, size, np.dtype(dtype)


def test_grouped_validation_reports_first_bad_problem():
    f = np.float64
    ok = grouped.GemmProblem(COL, 0, 0, 4, 4, 4, 1.0, _Host(16, f), 4, _Host(16, f), 4, 0.0,
                             _Host(16, f), 4)
    short_a = grouped.GemmProblem(COL, 0, 0, 4, 4, 4, 1.0, _Host(15, f), 4, _Host(16, f), 4, 0.0,
                                  _Host(16, f), 4)
    bad_enum = grouped.GemmProblem(COL, 5, 0, 4, 4, 4, 1.0, _Host(16, f), 4, _Host(16, f), 4, 0.0,
                                   _Host(16, f), 4)
    with pytest.raises(BlasPanic) as e:
        grouped.GroupedGemm([ok, ok, short_a, bad_enum], f)
    assert "problem 2" in str(e.value)
    with pytest.raises(ValueError) as e:
        grouped.GroupedGemm([ok, bad_enum, short_a], f)
    assert "problem 1" in str(e.value)


def _random_group(count, dtype, seed, lo=1, hi=300, integers=False, specials=True):
    """count problems, dims log-uniform in [lo, hi], all layouts/transposes,
    ld padded by 0..3; a few zero-sized, k == 0, alpha == 0, beta 0 / 1."""
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        m, n, k = (int(np.exp(rng.uniform(np.log(lo), np.log(hi + 1)))) for _ in range(3))
        layout, ta, tb = int(rng.integers(0, 2)), int(rng.integers(0, 3)), int(rng.integers(0, 3))
        alpha, beta = float(rng.choice([1.5, -0.5, 1.0])), float(rng.choice([-0.25, 0.0, 1.0, 2.0]))
        if specials:
            if i % 17 == 3:
                k = 0
            if i % 19 == 5:
                alpha = 0.0
            if i % 23 == 7:
                m = 0
        pad = int(rng.integers(0, 4))
        ar, ac = (k, m) if ta else (m, k)
        br, bc = (n, k) if tb else (k, n)
        lda = max(1, (ac if layout == ROW else ar) + pad)
        ldb = max(1, (bc if layout == ROW else br) + pad)
        ldc = max(1, (n if layout == ROW else m) + pad)
        sizes = [max(1, slice_len(layout, ar, ac, lda)), max(1, slice_len(layout, br, bc, ldb)),
                 max(1, slice_len(layout, m, n, ldc))]
        if integers:
            a, b, c = (rng.integers(-3, 4, s).astype(dtype) for s in sizes)
        else:
            a, b, c = (rng.standard_normal(s).astype(dtype) for s in sizes)
        probs.append(dict(layout=layout, ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta,
                          a=a, lda=lda, b=b, ldb=ldb, c=c, ldc=ldc))
    return probs


def _upload(probs, s):
    out = []
    for p in probs:
        da, db, dc = (DeviceArray.from_host(p[x], s) for x in ("a", "b", "c"))
        out.append(grouped.GemmProblem(p["layout"], p["ta"], p["tb"], p["m"], p["n"], p["k"],
                                       p["alpha"], da, p["lda"], db, p["ldb"], p["beta"], dc,
                                       p["ldc"]))
    return out


def _expect(p):
    return gemm_expect(p["layout"], p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["a"],
                       p["lda"], p["b"], p["ldb"], p["beta"], p["c"], p["ldc"])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grouped_exact_integers(gpu, dtype):
    s = Stream()
    probs = _random_group(96, dtype, 1, integers=True)
    dev = _upload(probs, s)
    plan = grouped.GroupedGemm(dev, dtype)
    plan.run(s)
    s.synchronize()
    for p, d in zip(probs, dev):
        got = d.c.download(s)
        want, _ = _expect(p)
        assert np.array_equal(view(got, p["layout"], p["m"], p["n"], p["ldc"]).astype(np.float64),
                              want)
        keep = np.ones(got.size, bool)
        view(keep, p["layout"], p["m"], p["n"], p["ldc"])[:] = False
        assert np.array_equal(got[keep], p["c"][keep])
    info = plan.info()
    assert info["problems"] == 96 and info["units"] > 0
    plan.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grouped_random_vs_bound_and_single_calls(gpu, dtype):
    """Random data, shapes up to 600, plus six problems with k = 20000 (long
    enough for the planner to split their k): each problem within the gemm
    bound, and each equal to the same problem run alone as a group of one,
    bit for bit."""
    s = Stream()
    probs = _random_group(48, dtype, 2, lo=8, hi=600)
    for i, p in enumerate(probs[:6]):  # a few long-k problems
        rng = np.random.default_rng(50 + i)
        m, n, k = 70 + 10 * i, 90, 20000
        p.update(layout=COL, ta=0, tb=i % 2, m=m, n=n, k=k, lda=m, ldb=(n if i % 2 else k), ldc=m)
        p["a"] = rng.standard_normal(m * k).astype(dtype)
        p["b"] = rng.standard_normal(k * n).astype(dtype)
        p["c"] = rng.standard_normal(m * n).astype(dtype)
    dev = _upload(probs, s)
    plan = grouped.GroupedGemm(dev, dtype)
    plan.run(s)
    s.synchronize()
    assert plan.info()["split_problems"] >= 1
    results = [d.c.download(s) for d in dev]
    for p, got in zip(probs, results):
        want, mag = _expect(p)
        assert_gemm_close(view(got, p["layout"], p["m"], p["n"], p["ldc"]), want, mag, p["k"],
                          dtype, what="grouped")
    for idx in (0, 1, 5, 9, 20, 33):
        p = probs[idx]
        alone = _upload([p], s)
        g1 = grouped.GroupedGemm(alone, dtype)
        g1.run(s)
        s.synchronize()
        assert np.array_equal(alone[0].c.download(s), results[idx]), f"problem {idx}"
        g1.close()
    plan.close()
    s.close()


@pytest.mark.gpu
def test_grouped_rerun_and_graph_replay(gpu):
    """The schedule counters reset themselves: plain reruns and graph
    replays (with eager work in between) reproduce the first run exactly."""
    dtype = np.float64
    s = Stream()
    probs = _random_group(40, dtype, 3, lo=16, hi=400)
    dev = _upload(probs, s)
    plan = grouped.GroupedGemm(dev, dtype)
    plan.run(s)
    s.synchronize()
    first = [d.c.download(s) for d in dev]

    def reset():
        for p, d in zip(probs, dev):
            d.c.upload(p["c"], s)

    for replay in range(4):
        reset()
        if replay % 2:
            plan.run(s)
        else:
            plan.run_graph(s)
        x = DeviceArray.from_host(np.ones(1000), s)  # unrelated eager work between replays
        level3.gemm_dev(COL, 0, 0, 10, 10, 10, 1.0, x, 10, x, 10, 0.0, x, 10, s)
        s.synchronize()
        for i, d in enumerate(dev):
            assert np.array_equal(d.c.download(s), first[i]), (replay, i)
        x.free()
    plan.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grouped_batching_workload_sample(gpu, dtype):
    """The batching workload's shape family (512 problems, m, n, k
    log-uniform in [16, 2048], random transposes, U[-1, 1]); a sample of the
    problems checked entry by entry against float64."""
    rng = np.random.default_rng(4)
    s = Stream()
    probs, dev = [], []
    for i in range(512):
        m, n, k = (int(np.exp(rng.uniform(np.log(16), np.log(2049)))) for _ in range(3))
        ta, tb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        ar, ac = (k, m) if ta else (m, k)
        br, bc = (n, k) if tb else (k, n)
        a = rng.uniform(-1, 1, ar * ac).astype(dtype)
        b = rng.uniform(-1, 1, br * bc).astype(dtype)
        c = rng.uniform(-1, 1, m * n).astype(dtype)
        probs.append(dict(layout=COL, ta=ta, tb=tb, m=m, n=n, k=k, alpha=1.0, beta=0.5, a=a,
                          lda=ar, b=b, ldb=br, c=c, ldc=m))
    dev = _upload(probs, s)
    plan = grouped.GroupedGemm(dev, dtype)
    plan.run(s)
    s.synchronize()
    for idx in rng.choice(512, 24, replace=False):
        p = probs[idx]
        got = dev[idx].c.download(s)
        want, mag = _expect(p)
        assert_gemm_close(view(got, COL, p["m"], p["n"], p["ldc"]), want, mag, p["k"], dtype,
                          what=f"problem {idx}")
    plan.close()
    s.close()
