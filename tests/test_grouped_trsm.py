"""Grouped trsm plans and non-finite propagation of both trsm paths.

Each problem is rblas's trsm (level3.rs:390-520). Tolerance (written here):
componentwise residual |op(A) X - alpha B| <= 16 q u (|op(A)||X| + |alpha B|)
(q the order of A, u the unit roundoff of the operand type; the grouped path
computes in f64, so f32 problems meet it with room). Exact-integer systems
are bit-exact. A NaN planted in B reaches exactly the entries rblas's
substitution reaches (the oracle's NaN pattern).
"""
import itertools

import numpy as np
import pytest

from rblas_amd import grouped, level3
from rblas_amd.device import DeviceArray, Stream
from rblas_amd.types import BlasPanic
from tests.util import COL, view


class _Host:
    def __init__(self, size, dtype):
        self.ptr, self.size, self.dtype = 0, size, np.dtype(dtype)


def test_grouped_trsm_validation():
    f = np.float64
    ok = grouped.TrsmProblem(0, 1, 0, 0, 4, 3, 1.0, _Host(16, f), 4, _Host(12, f), 4)
    short_b = grouped.TrsmProblem(0, 1, 0, 0, 4, 3, 1.0, _Host(16, f), 4, _Host(11, f), 4)
    too_big = grouped.TrsmProblem(1, 1, 0, 0, 2, 3000, 1.0, _Host(3000 * 3000, f), 3000,
                                  _Host(6000, f), 2)
    with pytest.raises(BlasPanic) as e:
        grouped.GroupedTrsm([ok, short_b], f)
    assert "problem 1" in str(e.value)
    with pytest.raises(ValueError) as e:
        grouped.GroupedTrsm([ok, ok, too_big], f)
    assert "problem 2" in str(e.value)


def _system(q, uplo, trans, diag, lda, dtype, rng):
    """Stored A whose op(A) is triangular: diag 2+|N(0,1)|, off-diagonal
    N(0,1)/sqrt(q); the other triangle holds large values that must not be
    read, and so does the diagonal for Unit."""
    full = rng.standard_normal((q, q)) / np.sqrt(q)
    np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
    tri = np.triu(np.ones((q, q), bool)) if uplo == 0 else np.tril(np.ones((q, q), bool))
    stored = np.where(tri, full, 1e3)
    if diag == 1:
        np.fill_diagonal(stored, 7e3)
    a = np.full((q - 1) * lda + q, -5e2)
    view(a, COL, q, q, lda)[:] = stored
    a = a.astype(dtype)
    S = np.where(tri, view(a, COL, q, q, lda).astype(np.float64), 0.0)
    if diag == 1:
        np.fill_diagonal(S, 1.0)
    return a, (S.T if trans else S)


def _residual_ratio(side, opA, X, alpha, B0, q, dtype):
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
    if side == 0:
        r = np.abs(opA @ X - alpha * B0)
        bound = 16 * q * u * (np.abs(opA) @ np.abs(X) + np.abs(alpha * B0))
    else:
        r = np.abs(X @ opA - alpha * B0)
        bound = 16 * q * u * (np.abs(X) @ np.abs(opA) + np.abs(alpha * B0))
    return float((r / np.maximum(bound, 1e-300)).max())


def _make_group(count, dtype, seed, qmax=300, rmax=40):
    rng = np.random.default_rng(seed)
    probs = []
    combos = list(itertools.product((0, 1), (0, 1), (0, 1, 2), (0, 1)))
    for i in range(count):
        side, uplo, trans, diag = combos[i % len(combos)]
        q = int(np.exp(rng.uniform(0, np.log(qmax))))
        nr = int(rng.integers(1, rmax + 1))
        m, n = (q, nr) if side == 0 else (nr, q)
        lda, ldb = q + int(rng.integers(0, 3)), m + int(rng.integers(0, 3))
        a, opA = _system(q, uplo, trans, diag, lda, dtype, rng)
        b = rng.standard_normal((n - 1) * ldb + m).astype(dtype)
        alpha = float(rng.choice([1.0, 0.75, -2.0]))
        probs.append(dict(side=side, uplo=uplo, trans=trans, diag=diag, m=m, n=n, alpha=alpha,
                          a=a, lda=lda, b=b, ldb=ldb, opA=opA, q=q))
    return probs


def _upload(probs, s):
    out = []
    for p in probs:
        da, db = DeviceArray.from_host(p["a"], s), DeviceArray.from_host(p["b"], s)
        out.append(grouped.TrsmProblem(p["side"], p["uplo"], p["trans"], p["diag"], p["m"], p["n"],
                                       p["alpha"], da, p["lda"], db, p["ldb"]))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grouped_trsm_all_combos(gpu, dtype):
    s = Stream()
    probs = _make_group(96, dtype, 11)
    dev = _upload(probs, s)
    plan = grouped.GroupedTrsm(dev, dtype)
    plan.run(s)
    s.synchronize()
    for p, d in zip(probs, dev):
        got = d.b.download(s)
        X = view(got, COL, p["m"], p["n"], p["ldb"]).astype(np.float64)
        B0 = view(p["b"], COL, p["m"], p["n"], p["ldb"]).astype(np.float64)
        ratio = _residual_ratio(p["side"], p["opA"], X, p["alpha"], B0, p["q"], dtype)
        assert ratio <= 1.0, (p["side"], p["uplo"], p["trans"], p["diag"], p["q"], ratio)
        keep = np.ones(got.size, bool)
        view(keep, COL, p["m"], p["n"], p["ldb"])[:] = False
        assert np.array_equal(got[keep], p["b"][keep])
    plan.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans", list(itertools.product((0, 1), (0, 1), (0, 1))))
def test_grouped_trsm_exact_integers(gpu, side, uplo, trans):
    """Bidiagonal op(A), entries in {-1, 0, 1}, diagonal +-1, q from 1 to
    1500: the exact integer solution, bit for bit (both widths of unit)."""
    rng = np.random.default_rng(3 + 4 * side + 2 * uplo + trans)
    s = Stream()
    probs, wants = [], []
    for q in (1, 17, 64, 65, 200, 1024, 1500):
        nr = int(rng.integers(1, 40))
        m, n = (q, nr) if side == 0 else (nr, q)
        S = np.diag(rng.choice([-1.0, 1.0], q))
        off = rng.integers(-1, 2, max(q - 1, 0)).astype(np.float64)
        eff_upper = (uplo == 0) != bool(trans)
        opA = S + (np.diag(off, 1) if eff_upper else np.diag(off, -1))
        stored = opA.T if trans else opA
        X = rng.integers(-4, 5, (m, n)).astype(np.float64)
        B = opA @ X if side == 0 else X @ opA
        probs.append(dict(a=stored.T.reshape(-1).copy(), b=B.T.reshape(-1).copy(), q=q, m=m, n=n))
        wants.append(X)
    dev = []
    for p in probs:
        da, db = DeviceArray.from_host(p["a"], s), DeviceArray.from_host(p["b"], s)
        dev.append(grouped.TrsmProblem(side, uplo, trans, 0, p["m"], p["n"], 1.0, da, p["q"], db,
                                       p["m"]))
    plan = grouped.GroupedTrsm(dev, np.float64)
    plan.run(s)
    s.synchronize()
    for p, d, X in zip(probs, dev, wants):
        assert np.array_equal(d.b.download(s).reshape(p["n"], p["m"]).T, X), p["q"]
    plan.close()
    s.close()


@pytest.mark.gpu
def test_grouped_trsm_matches_single_calls_and_replays(gpu):
    dtype = np.float64
    s = Stream()
    probs = _make_group(40, dtype, 12, qmax=700, rmax=70)
    dev = _upload(probs, s)
    plan = grouped.GroupedTrsm(dev, dtype)
    plan.run(s)
    s.synchronize()
    first = [d.b.download(s) for d in dev]
    for p, got in zip(probs, first):  # single-call path: same solution within the bound
        b1 = p["b"].copy()
        level3.trsm(p["side"], p["uplo"], p["trans"], p["diag"], p["m"], p["n"], p["alpha"],
                    p["a"], p["lda"], b1, p["ldb"])
        X1 = view(b1, COL, p["m"], p["n"], p["ldb"])
        X2 = view(got, COL, p["m"], p["n"], p["ldb"])
        scale = np.abs(X1).max() + 1e-300
        assert np.abs(X1 - X2).max() <= 1e-10 * scale
    for replay in range(3):
        for p, d in zip(probs, dev):
            d.b.upload(p["b"], s)
        plan.run_graph(s) if replay % 2 == 0 else plan.run(s)
        s.synchronize()
        for i, d in enumerate(dev):
            assert np.array_equal(d.b.download(s), first[i]), (replay, i)
    plan.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grouped_trsm_wide_units_match_single_calls(gpu, dtype):
    """48 systems of order up to 1024 with 100..250 right-hand sides: the
    group runs them as 32-wide units (its planner widens them while the units
    still fill the CUs), a single call as 16-wide ones. Each right-hand side is computed by the same
    sequence of f64 operations at either width, so the results agree bit for
    bit; they also meet the residual bound."""
    rng = np.random.default_rng(31)
    s = Stream()
    probs = []
    combos = list(itertools.product((0, 1), (0, 1), (0, 1, 2), (0, 1)))
    for i in range(48):
        side, uplo, trans, diag = combos[i % len(combos)]
        q = int(np.exp(rng.uniform(np.log(20), np.log(1024))))
        nr = int(rng.integers(100, 251))
        m, n = (q, nr) if side == 0 else (nr, q)
        lda, ldb = q + 1, m + 2
        a, opA = _system(q, uplo, trans, diag, lda, dtype, rng)
        b = rng.standard_normal((n - 1) * ldb + m).astype(dtype)
        probs.append(dict(side=side, uplo=uplo, trans=trans, diag=diag, m=m, n=n, alpha=0.75,
                          a=a, lda=lda, b=b, ldb=ldb, opA=opA, q=q))
    dev = _upload(probs, s)
    plan = grouped.GroupedTrsm(dev, dtype)
    nrhs = [p["n"] if p["side"] == 0 else p["m"] for p in probs]
    units16 = sum((r + 15) // 16 for r in nrhs)
    assert sum((r + 31) // 32 for r in nrhs) <= plan.info()["units"] < units16  # some widened
    plan.run(s)
    s.synchronize()
    for i, (p, d) in enumerate(zip(probs, dev)):
        got = d.b.download(s)
        b1 = p["b"].copy()
        level3.trsm(p["side"], p["uplo"], p["trans"], p["diag"], p["m"], p["n"], p["alpha"],
                    p["a"], p["lda"], b1, p["ldb"])
        assert np.array_equal(got, b1), (i, p["side"], p["uplo"], p["trans"], p["q"])
        if i % 6 == 0:
            X = view(got, COL, p["m"], p["n"], p["ldb"]).astype(np.float64)
            B0 = view(p["b"], COL, p["m"], p["n"], p["ldb"]).astype(np.float64)
            assert _residual_ratio(p["side"], p["opA"], X, p["alpha"], B0, p["q"], dtype) <= 1.0
    plan.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans", list(itertools.product((0, 1), (0, 1), (0, 1))))
def test_trsm_single_call_wide_units_exact(gpu, side, uplo, trans):
    """A single call with 6200 right-hand sides takes 32-wide units (16-wide
    ones would be more than the CUs need): bidiagonal op(A) with entries in
    {-1, 0, 1}, diagonal +-1, order 300, exact integer solution bit for bit."""
    q, nr = 300, 6200
    m, n = (q, nr) if side == 0 else (nr, q)
    rng = np.random.default_rng(40 + 4 * side + 2 * uplo + trans)
    S = np.diag(rng.choice([-1.0, 1.0], q))
    off = rng.integers(-1, 2, q - 1).astype(np.float64)
    eff_upper = (uplo == 0) != bool(trans)
    opA = S + (np.diag(off, 1) if eff_upper else np.diag(off, -1))
    stored = opA.T if trans else opA
    X = rng.integers(-4, 5, (m, n)).astype(np.float64)
    B = opA @ X if side == 0 else X @ opA
    b = B.T.reshape(-1).copy()
    level3.trsm(side, uplo, trans, 0, m, n, 1.0, stored.T.reshape(-1).copy(), q, b, m)
    assert np.array_equal(b.reshape(n, m).T, X)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans", list(itertools.product((0, 1), (0, 1), (0, 1))))
def test_trsm_nan_reaches_what_rblas_reaches(gpu, oracle, side, uplo, trans):
    """One NaN in B: the set of NaN results equals the oracle's (rblas's
    substitution order), for the single-call and the grouped path."""
    rng = np.random.default_rng(40 + 4 * side + 2 * uplo + trans)
    q, nr = 150, 5
    m, n = (q, nr) if side == 0 else (nr, q)
    a, _ = _system(q, uplo, trans, 0, q, np.float64, rng)
    b = rng.standard_normal(m * n)
    b[(m // 2) + (n // 2) * m] = np.nan
    want = b.copy()
    oracle.trsm(side, uplo, trans, 0, m, n, 1.0, a, q, want, m)
    single = b.copy()
    level3.trsm(side, uplo, trans, 0, m, n, 1.0, a, q, single, m)
    s = Stream()
    da, db = DeviceArray.from_host(a, s), DeviceArray.from_host(b, s)
    plan = grouped.GroupedTrsm([grouped.TrsmProblem(side, uplo, trans, 0, m, n, 1.0, da, q, db,
                                                    m)], np.float64)
    plan.run(s)
    s.synchronize()
    grp = db.download(s)
    assert 0 < np.isnan(want).sum() < want.size
    assert np.array_equal(np.isnan(single), np.isnan(want))
    assert np.array_equal(np.isnan(grp), np.isnan(want))
    plan.close()
    s.close()


@pytest.mark.gpu
def test_grouped_trsm_batching_workload_sample(gpu):
    """256 lower systems, n log-uniform in [32, 1024], nrhs in [1, 256],
    diag 2+|N(0,1)|, off-diagonal N(0,1)/sqrt(n), f64; residuals of a sample."""
    rng = np.random.default_rng(5)
    s = Stream()
    probs = []
    for _ in range(256):
        q = int(np.exp(rng.uniform(np.log(32), np.log(1025))))
        nr = int(rng.integers(1, 257))
        full = rng.standard_normal((q, q)) / np.sqrt(q)
        np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
        L = np.tril(full)
        probs.append(dict(q=q, nr=nr, L=L, a=L.T.reshape(-1).copy(),
                          b=rng.standard_normal(q * nr)))
    dev = []
    for p in probs:
        da, db = DeviceArray.from_host(p["a"], s), DeviceArray.from_host(p["b"], s)
        dev.append(grouped.TrsmProblem(0, 1, 0, 0, p["q"], p["nr"], 1.0, da, p["q"], db, p["q"]))
    plan = grouped.GroupedTrsm(dev, np.float64)
    plan.run(s)
    s.synchronize()
    for idx in rng.choice(256, 20, replace=False):
        p = probs[idx]
        X = dev[idx].b.download(s).reshape(p["nr"], p["q"]).T
        B0 = p["b"].reshape(p["nr"], p["q"]).T
        assert _residual_ratio(0, p["L"], X, 1.0, B0, p["q"], np.float64) <= 1.0, idx
    plan.close()
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("q,nr", [(705, 40), (769, 17), (1000, 300), (1024, 1024)])
@pytest.mark.parametrize("side,uplo,trans", [(0, 1, 0), (0, 0, 0), (1, 0, 1), (1, 1, 0)])
def test_trsm_split_single_call_matches_group(gpu, q, nr, side, uplo, trans):
    """Single calls whose units fill at most half the CUs run each unit as
    two workgroups (the second half of the diagonal blocks waits for the X of
    the first): bit for bit the group's one-workgroup solve of the same
    problem, forward and backward, Left and Right, within the residual
    bound."""
    dtype = np.float64
    rng = np.random.default_rng(q + nr + 8 * side + 4 * uplo + 2 * trans)
    m, n = (q, nr) if side == 0 else (nr, q)
    lda, ldb = q + 3, m + 1
    a, opA = _system(q, uplo, trans, 0, lda, dtype, rng)
    b = rng.standard_normal((n - 1) * ldb + m).astype(dtype)
    single = b.copy()
    level3.trsm(side, uplo, trans, 0, m, n, 0.5, a, lda, single, ldb)
    s = Stream()
    da, db = DeviceArray.from_host(a, s), DeviceArray.from_host(b, s)
    plan = grouped.GroupedTrsm([grouped.TrsmProblem(side, uplo, trans, 0, m, n, 0.5, da, lda, db,
                                                    ldb)], dtype)
    plan.run(s)
    s.synchronize()
    assert np.array_equal(single, db.download(s))
    X = view(single, COL, m, n, ldb)
    B0 = view(b, COL, m, n, ldb)
    assert _residual_ratio(side, opA, X, 0.5, B0, q, dtype) <= 1.0
    plan.close()
    s.close()
