"""Serving loop front end (rblas_serve_run, csrc/serve.cpp).

No rblas counterpart. Builds request traces and runs them through the
native serving loop in continuous-batching or plain per-call mode.

The serving workload: Poisson arrivals at rate `rate` (requests/s), `count`
requests, 60% gemm (m, n, k log-uniform in [32, 4096], f32 or f64 with
probability 1/2), 25% trsm (order U[64, 2048], right-hand sides U[1, 512],
f64), 15% complex f64 gemm (m, n, k log-uniform in [64, 2048]). The A and B
of gemm and complex requests are slices of a read-only operand pool per
element type, generated in HBM before the first arrival; the written
operands (C, a trsm's A and B) are generated in HBM per request (see
csrc/serve.hip). The admission cap is on the summed footprint (written
operands and workspace) of admitted, unfinished requests (200 GB by
default).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi, workloads
from .level3 import raise_status

GEMM, TRSM, ZGEMM = 0, 1, 2
BATCHED, PLAIN = 0, 1
HBM_CAP = 200e9
POOL_SEED = 0x5EEDB10C5EEDB10C  # RBLAS_SERVE_POOL_SEED
POOL_MIN = 1 << 25              # pool elements per element type, at least
_M64 = (1 << 64) - 1

REQUEST_DTYPE = np.dtype([("kind", np.int32), ("dtype", np.int32), ("m", np.int32),
                          ("n", np.int32), ("k", np.int32), ("pad", np.int32),
                          ("seed", np.uint64), ("arrival_s", np.float64)])
RESULT_DTYPE = np.dtype([("arrival_s", np.float64), ("admit_s", np.float64),
                         ("enqueued_s", np.float64), ("done_s", np.float64),
                         ("flops", np.float64),
                         ("footprint_bytes", np.float64), ("hash", np.uint64),
                         ("batch", np.int32), ("pad", np.int32)])


def _loguni(rng, lo, hi, size):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size))).astype(np.int32)


def workload(count: int = 10_000, rate: float = 0.0, seed: int = 7) -> np.ndarray:
    """The serving workload's request trace; rate 0 puts every arrival at 0
    (a closed-loop run that measures capacity)."""
    rng = np.random.default_rng(seed)
    req = np.zeros(count, REQUEST_DTYPE)
    kind = rng.choice([GEMM, TRSM, ZGEMM], size=count, p=[0.60, 0.25, 0.15])
    req["kind"] = kind
    g, t, z = kind == GEMM, kind == TRSM, kind == ZGEMM
    req["dtype"] = 1
    req["dtype"][g] = rng.integers(0, 2, g.sum())
    for f in ("m", "n", "k"):
        req[f][g] = _loguni(rng, 32, 4096, g.sum())
        req[f][z] = _loguni(rng, 64, 2048, z.sum())
    req["m"][t] = rng.integers(64, 2049, t.sum())
    req["n"][t] = rng.integers(1, 513, t.sum())
    req["k"][t] = 0
    req["seed"] = rng.integers(1, 2**63, count, dtype=np.uint64)
    if rate > 0:
        req["arrival_s"] = np.cumsum(rng.exponential(1.0 / rate, count))
    return req


def run(requests: np.ndarray, mode: int = BATCHED, max_inflight: int = 4, max_batch: int = 32,
        hbm_cap: float = HBM_CAP, capture: dict | None = None, hash_col_step: int = 1):
    """Serve `requests` (REQUEST_DTYPE). `capture` maps request index -> a
    host array that receives its dense result; hash_col_step > 1 hashes
    every hash_col_step-th column of each result only. Returns (results,
    stats)."""
    requests = np.ascontiguousarray(requests, REQUEST_DTYPE)
    results = np.zeros(len(requests), RESULT_DTYPE)
    capture = capture or {}
    ids = (C.c_int32 * max(1, len(capture)))(*capture.keys())
    outs = (C.c_void_p * max(1, len(capture)))(*[a.ctypes.data for a in capture.values()])
    cfg = _ffi.ServeConfig(mode, max_inflight, max_batch, len(capture), hash_col_step, 0, hbm_cap,
                           C.cast(ids, C.POINTER(C.c_int32)), C.cast(outs, C.POINTER(C.c_void_p)))
    stats = _ffi.ServeStats()
    st = _ffi.lib().rblas_serve_run(requests.ctypes.data, len(requests), C.byref(cfg),
                                    results.ctypes.data, C.byref(stats))
    raise_status(st)
    return results, {f: getattr(stats, f) for f, _ in _ffi.ServeStats._fields_}


def summary(results: np.ndarray, stats: dict) -> dict:
    lat = results["done_s"] - results["arrival_s"]
    return {"requests": int(len(results)), "wall_s": stats["wall_s"],
            "tflops": stats["flops"] / stats["wall_s"] / 1e12 if stats["wall_s"] > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "batches": stats["batches"],
            "max_inflight_seen": stats["max_inflight_seen"],
            "peak_footprint_gb": stats["peak_footprint_bytes"] / 1e9}


# ------------------------------------------------------- host regeneration
def _mix64(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def pool_elems(requests: np.ndarray) -> dict:
    """Pool size (elements) per element type (0 f32, 1 f64) of a run over
    `requests`: 2^25, or the largest pooled operand."""
    need = {0: 0, 1: 0}
    for r in np.atleast_1d(requests):
        kind, m, n, k = (int(r[f]) for f in ("kind", "m", "n", "k"))
        if kind == TRSM:
            continue
        c = 2 if kind == ZGEMM else 1
        dt = 1 if kind == ZGEMM else int(r["dtype"])
        need[dt] = max(need[dt], c * m * k, c * k * n)
    return {dt: max(POOL_MIN, v) for dt, v in need.items()}


def _pool_slice(seed: int, j: int, length: int, elems: int, dt) -> np.ndarray:
    off = _mix64(seed ^ j) % (elems - length + 1)
    return workloads.uniform(POOL_SEED, off, length, dt)


def operands(req, pools: dict | None = None) -> tuple:
    """The request's operands exactly as the device generates them (numpy),
    column-major flattened: gemm (A, B, C0); trsm (A, B0); zgemm complex
    (A, B, C0). `pools`: pool_elems() of the run's trace (default: the
    minimum pool, which any trace whose pooled operands stay within 2^25
    elements uses)."""
    kind, m, n, k, seed = int(req["kind"]), int(req["m"]), int(req["n"]), int(req["k"]), \
        int(req["seed"])
    pools = pools or {0: POOL_MIN, 1: POOL_MIN}
    if kind == GEMM:
        dt = np.float32 if int(req["dtype"]) == 0 else np.float64
        el = pools[int(req["dtype"])]
        la, lb, lc = m * k, k * n, m * n
        return (_pool_slice(seed, 1, la, el, dt), _pool_slice(seed, 2, lb, el, dt),
                workloads.uniform(seed, la + lb, lc, dt))
    if kind == TRSM:
        dt = np.float32 if int(req["dtype"]) == 0 else np.float64
        u = workloads.uniform(seed, 0, m * m, dt).reshape(m, m).T  # (row, col)
        ud = u.astype(np.float64)
        A = np.where(np.tri(m, dtype=bool), (ud / np.sqrt(m)).astype(dt), dt(0))
        np.fill_diagonal(A, (2.0 + np.abs(np.diag(ud))).astype(dt))
        return A.T.reshape(-1).copy(), workloads.uniform(seed, m * m, m * n, dt)
    la, lb, lc = 2 * m * k, 2 * k * n, 2 * m * n
    def cplx(x):
        return x.view(np.complex128)
    return (cplx(_pool_slice(seed, 1, la, pools[1], np.float64)),
            cplx(_pool_slice(seed, 2, lb, pools[1], np.float64)),
            cplx(workloads.uniform(seed, la + lb, lc, np.float64)))


def result_buffer(req) -> np.ndarray:
    kind, m, n = int(req["kind"]), int(req["m"]), int(req["n"])
    if kind == ZGEMM:
        return np.zeros(m * n, np.complex128)
    return np.zeros(m * n, np.float32 if int(req["dtype"]) == 0 else np.float64)
