#!/usr/bin/env python3
"""Benchmark of the MI355X Level-3 path through librblas_hip.so (one GPU).

Headline (`value`): DGEMM 8192^3 TFLOP/s. One step = one gemm call on
HBM-resident operands: C := 1.5 A B - 0.25 C, f64, column-major NN,
m = n = k = 8192, lda = ldb = ldc = 8199 (rows + 7), entries N(0,1) from the
seeded splitmix64 generator (rblas_amd.workloads), seed 1, drawn A, B, C.
Timed with HIP events on the stream the kernel runs on.

`components` measures the rest of the metric (and of the workloads), each on
synthetic data of the stated shape:
  - SGEMM 8192^3 in plain FP32, BF16x3 and BF16x6 (TFLOP/s, % of peak), and
    the precision study: inputs (a) N(0,1), (b) rows of A scaled by
    2^U[-20,20], (c) m = n = 1024, k = 65536; max |C - C64| / (|A||B|)
    over sampled rows against a float64 product;
  - grouped GEMM (512 problems, m, n, k log-uniform in [16, 2048], random
    transposes, U[-1,1], f32 and f64) in one persistent launch, against the
    same problems as one launch each;
  - single trsm calls (Left / Lower / NoTrans, q = nrhs in {1024, 2048,
    8192}, f64 and f32; Right / Lower / NoTrans at 8192, f64);
  - DSYMM 8192^3 and ZHEMM 4096^3 (dense A from the stored triangle, then
    the gemm);
  - grouped trsm (256 lower systems, n log-uniform [32, 1024], nrhs in
    [1, 256], diag 2+|N(0,1)|, off-diagonal N(0,1)/sqrt(n), f64) against
    per-call trsm;
  - serving: 10k requests (60% gemm, 25% trsm, 15% complex gemm), Poisson
    at 70% of the measured closed-loop capacity, 200 GB admission cap:
    p50/p99 latency and TFLOP/s, continuous batching vs plain per-call on
    the same trace, and plain per-call at 70% of its own capacity;
  - DGEMM and SGEMM 8192^3 in all 8 Layout x Transpose variants;
  - the small SGEMM (col-major NN 256^3, U[-1,1] seed 0, beta 0.5): device
    time per call and host-slice call time;
  - the headline DGEMM through the host-slice entry point (PCIe-inclusive);
  - complex trsm (ZTRSM / CTRSM 4096 x 4096, Left and Right);
  - C64 gemm (4M) 4096^3, Level-1 bandwidth (f64 and f32) and Level-2
    bandwidth (gemv NoTrans / Trans, ger; f64, f32, complex f64).
`roofline` prices the DGEMM kernel against MI355X's FP64 MFMA peak, with HBM
traffic per launch from the committed rocprofv3 PMC summary; `cpu_baseline`
times the oracle (rblas's CPU algorithm restated in C++) on the headline
DGEMM problem itself on the host cores, and `cpu_baseline.rows` the oracle
on the grouped GEMM and batched trsm workloads, a prefix of the serving
trace and the f64 Level-1 calls (full runs only).
"""
from __future__ import annotations

import argparse
import functools
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6    # MI355X dense FP64 matrix TFLOP/s
FP32_PEAK = 157.3   # MI355X dense FP32 matrix TFLOP/s
HBM_PEAK_GBS = 8000.0
METRIC = ("TFLOP/s & %peak: DGEMM, 3xBF16 SGEMM 8192^3; grouped GEMM TFLOP/s; serve p99 ms")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--quick", action="store_true", help="headline DGEMM only")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--serve-requests", type=int, default=10000)
    p.add_argument("--serve-inflight", type=int, default=4,
                   help="batches in flight for --only serve_batched (profiling)")
    p.add_argument("--only", choices=sorted(ONLY),
                   help="run one component and print its JSON (for profiling)")
    p.add_argument("--copy-calibrate", action="store_true",
                   help="only run 5 device-to-device copies of 1 GiB (counter calibration)")
    return p.parse_args()


PROFILE = os.path.join(ROOT, "profiles", "dgemm_8192_nn.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per DGEMM launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration), or None."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1].strip()))
    except OSError:
        pass
    return None


def copy_calibrate():
    from rblas_amd import _ffi
    from rblas_amd.device import DeviceArray, Stream
    n = 1 << 30
    s = Stream()
    src, dst = DeviceArray(n, np.uint8), DeviceArray(n, np.uint8)
    for _ in range(5):
        _ffi.lib().rblas_memcpy_dtod(dst.ptr, src.ptr, n, s.handle)
    s.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "copies": 5}), flush=True)


def timed(fn, stream, steps, warmup, min_warm_s=0.05):
    """ms per call of fn over `steps` calls, after at least `warmup` calls and
    at least min_warm_s seconds of them (a few short launches leave the clocks
    below their steady state: the first grouped run measured 7% slow)."""
    from rblas_amd.device import Event
    t0 = time.perf_counter()
    done = 0
    while done < warmup or time.perf_counter() - t0 < min_warm_s:
        fn()
        done += 1
        if done >= warmup:
            stream.synchronize()
    stream.synchronize()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    stream.synchronize()
    ms = e1.elapsed_ms(e0) / steps
    e0.close()
    e1.close()
    return ms


def square_gemm(dtype, n, ld, steps, warmup, compute=None, operands=None, min_warm_s=0.05):
    """ms per call of the 8192-style gemm (optionally a split-SGEMM compute
    mode). min_warm_s: see timed (the headline line passes 0: exactly
    --warmup calls)."""
    from rblas_amd import level3, workloads
    from rblas_amd.device import DeviceArray, Stream
    la = (n - 1) * ld + n
    if operands is None:
        operands = (workloads.normal(1, 0, la, dtype), workloads.normal(1, la, la, dtype),
                    workloads.normal(1, 2 * la, la, dtype))
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in operands)
    if compute is None:
        fn = lambda: level3.gemm_dev(1, 0, 0, n, n, n, 1.5, da, ld, db, ld, -0.25, dc, ld, s)
    else:
        fn = lambda: level3.sgemm_ex_dev(1, 0, 0, n, n, n, 1.5, da, ld, db, ld, -0.25, dc, ld,
                                         compute, s)
    ms = timed(fn, s, steps, warmup, min_warm_s)
    for x in (da, db, dc):
        x.free()
    s.close()
    return ms


def precision_case(case, rng_seed=3):
    """Max |C - C64| / (|A||B|) over 48 sampled rows, for FP32, BF16x3 and
    BF16x6 (alpha 1, beta 0), with the float64 product of the same rows as
    the reference."""
    from rblas_amd import level3, workloads
    from rblas_amd.device import DeviceArray, Stream
    from rblas_amd.types import Compute
    if case == "c":
        m = n = 1024
        k = 65536
    else:
        m = n = k = 8192
    A = workloads.normal(11, 0, m * k, np.float32)
    B = workloads.normal(11, m * k, k * n, np.float32)
    if case == "b":
        rng = np.random.default_rng(rng_seed)
        scale = np.exp2(rng.uniform(-20, 20, m)).astype(np.float32)
        A = (A.reshape(k, m) * scale[None, :]).reshape(-1)  # row i of A scaled
    rows = np.random.default_rng(5).choice(m, 48, replace=False)
    Am = A.reshape(k, m).T[rows].astype(np.float64)
    Bm = B.reshape(n, k).T.astype(np.float64)
    ref = Am @ Bm
    mag = np.abs(Am) @ np.abs(Bm)
    s = Stream()
    da, db = DeviceArray.from_host(A, s), DeviceArray.from_host(B, s)
    dc = DeviceArray(m * n, np.float32)
    out = {}
    for name, comp in (("fp32", Compute.F32), ("bf16x3", Compute.BF16X3),
                       ("bf16x6", Compute.BF16X6)):
        level3.sgemm_ex_dev(1, 0, 0, m, n, k, 1.0, da, m, db, k, 0.0, dc, m, comp, s)
        s.synchronize()
        C = dc.download(s).reshape(n, m).T[rows].astype(np.float64)
        out[name] = float((np.abs(C - ref) / mag).max())
    for x in (da, db, dc):
        x.free()
    s.close()
    out["bf16x3_over_fp32"] = out["bf16x3"] / out["fp32"]
    out["bf16x6_over_fp32"] = out["bf16x6"] / out["fp32"]
    return {k: float(f"{v:.4g}") for k, v in out.items()}


def gemm_variants(dtype, n=8192, steps=5):
    """All 8 Layout x Transpose variants at n^3 on one set of resident
    operands (stored ld = n + 7 fits every variant of a square problem)."""
    from rblas_amd import level3, workloads
    from rblas_amd.device import DeviceArray, Stream
    ld = n + 7
    la = (n - 1) * ld + n
    s = Stream()
    da, db, dc = (DeviceArray.from_host(workloads.normal(1, i * la, la, dtype), s)
                  for i in range(3))
    out = {}
    for layout, lname in ((1, "col"), (0, "row")):
        for ta in (0, 1):
            for tb in (0, 1):
                ms = timed(lambda: level3.gemm_dev(layout, ta, tb, n, n, n, 1.5, da, ld, db, ld,
                                                   -0.25, dc, ld, s), s, steps, 1)
                out[f"{lname}_{'NT'[ta]}{'NT'[tb]}"] = round(2.0 * n ** 3 / ms / 1e9, 2)
    for x in (da, db, dc):
        x.free()
    s.close()
    return out


def sgemm_256(calls=200):
    """The small plain SGEMM: col-major NN 256^3, f32, U[-1,1] seed 0 drawn A,
    B, C, alpha 1, beta 0.5, ld 256. Latency-bound: reports the device time
    per call (resident operands, back-to-back calls) and the host-slice call
    time (copies in and out, synchronous)."""
    from rblas_amd import level3, workloads
    from rblas_amd.device import DeviceArray, Stream
    n = 256
    a = workloads.uniform(0, 0, n * n, np.float32)
    b = workloads.uniform(0, n * n, n * n, np.float32)
    c = workloads.uniform(0, 2 * n * n, n * n, np.float32)
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    ms = timed(lambda: level3.gemm_dev(1, 0, 0, n, n, n, 1.0, da, n, db, n, 0.5, dc, n, s), s,
               calls, 10)
    for x in (da, db, dc):
        x.free()
    s.close()
    ch = c.copy()
    for _ in range(5):
        level3.gemm(1, 0, 0, n, n, n, 1.0, a, n, b, n, 0.5, ch, n)
    t0 = time.perf_counter()
    for _ in range(50):
        level3.gemm(1, 0, 0, n, n, n, 1.0, a, n, b, n, 0.5, ch, n)
    host_us = (time.perf_counter() - t0) / 50 * 1e6
    return {"device_us_per_call": round(ms * 1e3, 2),
            "device_tflops": round(2.0 * n ** 3 / ms / 1e9, 3),
            "host_call_us": round(host_us, 1)}


def dgemm_host_call(n=8192, calls=2):
    """The headline DGEMM through the host-slice entry point (operands in
    host memory: copies in over PCIe, the kernel, C copied back,
    synchronous): ms per call and the PCIe-inclusive TFLOP/s."""
    from rblas_amd import level3, workloads
    ld = n + 7
    la = (n - 1) * ld + n
    a, b, c = (workloads.normal(1, i * la, la, np.float64) for i in range(3))
    level3.gemm(1, 0, 0, n, n, n, 1.5, a, ld, b, ld, -0.25, c, ld)
    t0 = time.perf_counter()
    for _ in range(calls):
        level3.gemm(1, 0, 0, n, n, n, 1.5, a, ld, b, ld, -0.25, c, ld)
    ms = (time.perf_counter() - t0) / calls * 1e3
    return {"ms_per_call": round(ms, 2), "tflops_pcie_inclusive": round(2.0 * n ** 3 / ms / 1e9, 2)}


def skinny_sgemm_tflops(steps=5):
    """m = n = 1024, k = 65536 SGEMM (the precision study's case c)."""
    from rblas_amd import level3, workloads
    from rblas_amd.device import DeviceArray, Stream
    from rblas_amd.types import Compute
    m = n = 1024
    k = 65536
    s = Stream()
    da = DeviceArray.from_host(workloads.uniform(2, 0, m * k, np.float32), s)
    db = DeviceArray.from_host(workloads.uniform(2, m * k, k * n, np.float32), s)
    dc = DeviceArray(m * n, np.float32)
    out = {}
    for name, comp in (("fp32", Compute.F32), ("bf16x3", Compute.BF16X3),
                       ("bf16x6", Compute.BF16X6)):
        ms = timed(lambda: level3.sgemm_ex_dev(1, 0, 0, m, n, k, 1.0, da, m, db, k, 0.0, dc, m,
                                               comp, s), s, steps, 1)
        out[name] = round(2.0 * m * n * k / ms / 1e9, 2)
    for x in (da, db, dc):
        x.free()
    s.close()
    return out


def grouped_gemm_set(dtype):
    """The grouped GEMM workload, one problem at a time: 512 column-major
    problems, m, n, k = floor(exp U[ln 16, ln 2049]), transposes uniform over
    the four combinations, alpha 1, beta 0.5, operands U[-1, 1] drawn from one
    stream (seed 21) in problem order. Yields (ta, tb, m, n, k, lda, ldb,
    [A, B, C0])."""
    from rblas_amd import workloads
    rng = np.random.default_rng(4)
    off = 0
    for _ in range(512):
        m, n, k = (int(np.exp(rng.uniform(np.log(16), np.log(2049)))) for _ in range(3))
        ta, tb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        ar, ac = (k, m) if ta else (m, k)
        br, bc = (n, k) if tb else (k, n)
        host = []
        for cnt in (ar * ac, br * bc, m * n):
            host.append(workloads.uniform(21, off, cnt, dtype))
            off += cnt
        yield ta, tb, m, n, k, ar, br, host


def grouped_trsm_set():
    """The batched trsm workload: 256 lower-triangular f64 systems (Left,
    NoTrans, NonUnit, alpha 1), order floor(exp U[ln 32, ln 1025]), right-hand
    sides U{1..256}, diagonal 2 + |N(0,1)|, off-diagonal N(0,1)/sqrt(order),
    seed 5. Yields (order, nrhs, A, B0), A column-major."""
    rng = np.random.default_rng(5)
    for _ in range(256):
        q = int(np.exp(rng.uniform(np.log(32), np.log(1025))))
        nr = int(rng.integers(1, 257))
        full = rng.standard_normal((q, q)) / np.sqrt(q)
        np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
        yield q, nr, np.tril(full).T.reshape(-1).copy(), rng.standard_normal(q * nr)


def grouped_gemm(dtype):
    from rblas_amd import grouped, level3
    from rblas_amd.device import DeviceArray, Stream
    s = Stream()
    probs, flops = [], 0.0
    for ta, tb, m, n, k, ar, br, host in grouped_gemm_set(dtype):
        arrs = [DeviceArray.from_host(h, s) for h in host]
        probs.append(grouped.GemmProblem(1, ta, tb, m, n, k, 1.0, arrs[0], ar, arrs[1], br, 0.5,
                                         arrs[2], m))
        flops += 2.0 * m * n * k
    plan = grouped.GroupedGemm(probs, dtype)
    ms = timed(lambda: plan.run(s), s, 10, 2)
    ms_graph = timed(lambda: plan.run_graph(s), s, 10, 2)

    def percall():
        for p in probs:
            level3.gemm_dev(1, p.transa, p.transb, p.m, p.n, p.k, 1.0, p.a, p.lda, p.b, p.ldb,
                            0.5, p.c, p.ldc, s)
    ms_pc = timed(percall, s, 3, 1)
    info = plan.info()
    peak = FP32_PEAK if dtype == np.float32 else FP64_PEAK
    # Per-problem roofline: max(flops / peak, bytes / HBM) summed over problems.
    sz = np.dtype(dtype).itemsize
    ideal_s = 0.0
    for p in probs:
        f = 2.0 * p.m * p.n * p.k
        b = (p.m * p.k + p.k * p.n + 2 * p.m * p.n) * sz
        ideal_s += max(f / (peak * 1e12), b / (HBM_PEAK_GBS * 1e9))
    plan.close()
    for p in probs:
        for x in (p.a, p.b, p.c):
            x.free()
    s.close()
    return {"tflops": round(flops / ms / 1e9, 2), "ms": round(ms, 3),
            "graph_ms": round(ms_graph, 3), "per_call_ms": round(ms_pc, 3),
            "speedup_vs_per_call": round(ms_pc / ms, 2),
            "roofline_frac": round(ideal_s * 1e3 / ms, 3), "units": info["units"],
            "tiles64": info["tiles64"], "tiles128": info["tiles128"],
            "split_problems": info["split_problems"]}


def grouped_trsm():
    from rblas_amd import grouped, level3
    from rblas_amd.device import DeviceArray, Stream
    s = Stream()
    probs, flops, ideal_s = [], 0.0, 0.0
    for q, nr, ha, hb in grouped_trsm_set():
        a = DeviceArray.from_host(ha, s)
        b = DeviceArray.from_host(hb, s)
        probs.append(grouped.TrsmProblem(0, 1, 0, 0, q, nr, 1.0, a, q, b, q))
        flops += float(q) * q * nr
        # Per-system roofline: max(q^2 nrhs / FP64 peak, (q^2/2 + 2 q nrhs) 8 B / HBM).
        ideal_s += max(float(q) * q * nr / (FP64_PEAK * 1e12),
                       (q * q / 2.0 + 2.0 * q * nr) * 8 / (HBM_PEAK_GBS * 1e9))
    plan = grouped.GroupedTrsm(probs, np.float64)
    ms = timed(lambda: plan.run(s), s, 10, 2)

    def percall():
        for p in probs:
            level3.trsm_dev(0, 1, 0, 0, p.m, p.n, 1.0, p.a, p.lda, p.b, p.ldb, s)
    ms_pc = timed(percall, s, 2, 1)
    plan.close()
    for p in probs:
        p.a.free()
        p.b.free()
    s.close()
    return {"tflops": round(flops / ms / 1e9, 3), "ms": round(ms, 3),
            "per_call_ms": round(ms_pc, 3), "speedup_vs_per_call": round(ms_pc / ms, 1),
            "roofline_frac": round(ideal_s * 1e3 / ms, 3)}


def symm_hemm():
    """DSYMM (Left, Upper, m = n = 8192, ld 8199; 2 m^2 n flops) and ZHEMM
    (Left, Lower, m = n = 4096; 8 m^2 n real flops), resident operands."""
    from rblas_amd import complex3, level3, workloads
    from rblas_amd.device import DeviceArray, Stream
    s = Stream()
    out = {}
    n = 8192
    ld = n + 7
    la = (n - 1) * ld + n
    da, db, dc = (DeviceArray.from_host(workloads.normal(5, i * la, la, np.float64), s)
                  for i in range(3))
    ms = timed(lambda: level3.symm_dev(0, 0, n, n, 1.5, da, ld, db, ld, -0.25, dc, ld, s), s, 5, 1)
    out["dsymm_8192"] = {"ms": round(ms, 3), "tflops": round(2.0 * n ** 3 / ms / 1e9, 2)}
    for x in (da, db, dc):
        x.free()
    n = 4096
    z = [workloads.normal(6, 2 * i * n * n, 2 * n * n, np.float64).view(np.complex128)
         for i in range(3)]
    da, db, dc = (DeviceArray.from_host(x, s) for x in z)
    ms = timed(lambda: complex3.hemm_dev(0, 1, n, n, 1.5 - 0.5j, da, n, db, n, 0.25, dc, n, s),
               s, 5, 1)
    out["zhemm_4096"] = {"ms": round(ms, 3), "tflops": round(8.0 * n ** 3 / ms / 1e9, 2)}
    for x in (da, db, dc):
        x.free()
    s.close()
    return out


def trsm_single():
    """Single trsm calls on resident operands: Left / Lower / NoTrans,
    q = nrhs in {1024, 2048, 8192}, f64 and f32, and Right / Lower / NoTrans
    at 8192 f64 (TFLOP/s in q^3)."""
    from rblas_amd import level3
    from rblas_amd.device import DeviceArray, Stream
    rng = np.random.default_rng(11)
    s = Stream()
    out = {}
    for q in (1024, 2048, 8192):
        full = rng.standard_normal((q, q)) / np.sqrt(q)
        np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
        a64 = np.tril(full).T.reshape(-1).copy()
        b64 = rng.standard_normal(q * q)
        del full
        for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
            da = DeviceArray.from_host(a64.astype(dt), s)
            db = DeviceArray.from_host(b64.astype(dt), s)
            steps = 5 if q == 8192 else 10
            ms = timed(lambda: level3.trsm_dev(0, 1, 0, 0, q, q, 1.0, da, q, db, q, s), s, steps, 2)
            out[f"{tag}_{q}"] = {"ms": round(ms, 3), "tflops": round(float(q) ** 3 / ms / 1e9, 2)}
            da.free()
            db.free()
        if q == 8192:
            da = DeviceArray.from_host(a64, s)
            db = DeviceArray.from_host(b64, s)
            ms = timed(lambda: level3.trsm_dev(1, 1, 0, 0, q, q, 1.0, da, q, db, q, s), s, 5, 2)
            out["f64_8192_right"] = {"ms": round(ms, 3),
                                     "tflops": round(float(q) ** 3 / ms / 1e9, 2)}
            da.free()
            db.free()
    s.close()
    return out


def ctrsm_single():
    """Complex trsm on resident operands, q = nrhs = 4096: ZTRSM and CTRSM
    Left / Lower / NoTrans, and ZTRSM Right / Upper / ConjTrans (TFLOP/s in
    4 q^2 nrhs real flops)."""
    from rblas_amd import complex3
    from rblas_amd.device import DeviceArray, Stream
    rng = np.random.default_rng(12)
    s = Stream()
    out = {}
    q = 4096
    full = (rng.standard_normal((q, q)) + 1j * rng.standard_normal((q, q))) / np.sqrt(q)
    np.fill_diagonal(full, 2.0 + np.abs(rng.standard_normal(q)))
    b = rng.standard_normal(q * q) + 1j * rng.standard_normal(q * q)
    for tag, side, uplo, trans, a, ct in (
            ("z_llnn", 0, 1, 0, np.tril(full), np.complex128),
            ("c_llnn", 0, 1, 0, np.tril(full), np.complex64),
            ("z_ruct", 1, 0, 2, np.triu(full), np.complex128)):
        da = DeviceArray.from_host(a.T.reshape(-1).astype(ct), s)
        db = DeviceArray.from_host(b.astype(ct), s)
        ms = timed(lambda: complex3.trsm_dev(side, uplo, trans, 0, q, q, 1.0, da, q, db, q, s),
                   s, 5, 2)
        out[f"{tag}_{q}"] = {"ms": round(ms, 3), "tflops": round(4.0 * q ** 3 / ms / 1e9, 2)}
        da.free()
        db.free()
    s.close()
    return out


def level2_bw(n=16384, steps=10):
    """gemv / ger on resident operands, column-major, ld q + 7: DGEMV and
    SGEMV NoTrans and Trans at q = n, ZGEMV NoTrans and ConjTrans at q = n/2,
    DGER at q = n (one N(0,1) buffer, seed 31, serves every case).
    Algorithmic bytes: gemv reads A once plus x and y and writes y; ger reads
    and writes A and reads x and y."""
    from rblas_amd import level2, workloads
    from rblas_amd.device import DeviceArray, Stream
    s = Stream()
    out = {}
    base = workloads.normal(31, 0, n * (n + 7) + 4 * n, np.float64)
    cases = (("dgemv_n", np.float64, n, 0), ("dgemv_t", np.float64, n, 1),
             ("dger", np.float64, n, None),
             ("sgemv_n", np.float32, n, 0), ("sgemv_t", np.float32, n, 1),
             ("zgemv_n", np.complex128, n // 2, 0), ("zgemv_c", np.complex128, n // 2, 2))
    for name, dt, q, trans in cases:
        ld = q + 7
        per = 2 if np.dtype(dt).kind == "c" else 1
        la = per * ld * q
        real = np.float32 if dt == np.float32 else np.float64
        a = base[:la].astype(real).view(dt)
        x = base[la:la + per * q].astype(real).view(dt)
        y = base[la + per * q:la + 2 * per * q].astype(real).view(dt)
        da, dx, dy = (DeviceArray.from_host(v, s) for v in (a, x, y))
        del a
        es = np.dtype(dt).itemsize
        if trans is None:
            ms = timed(lambda: level2.ger_dev(1, q, q, 0.5, dx, 1, dy, 1, da, ld, s), s, steps, 2)
            nbytes = 2 * q * q * es + 2 * q * es
        else:
            al, be = (1.5, -0.25) if per == 1 else (1.5 - 0.5j, 0.25j)
            ms = timed(lambda: level2.gemv_dev(1, trans, q, q, al, da, ld, dx, 1, be, dy, 1, s),
                       s, steps, 2)
            nbytes = q * q * es + 3 * q * es
        gbs = nbytes / ms / 1e6
        out[f"{name}_{q}"] = {"ms": round(ms, 4), "gb_s": round(gbs, 1),
                              "hbm_frac": round(gbs / HBM_PEAK_GBS, 3)}
        for v in (da, dx, dy):
            v.free()
    s.close()
    return out


# The serving runs fingerprint every 8th column of each result (the tests
# hash every column): the hash is the harness's record of a result, not
# serving work.
SERVE_HASH_COL_STEP = 8


def serving(count):
    """Returns (summary, the timed request trace)."""
    from rblas_amd import serve
    run = functools.partial(serve.run, hash_col_step=SERVE_HASH_COL_STEP)
    closed = serve.workload(count, 0.0)
    res, st = run(closed, serve.BATCHED)
    capacity = count / st["wall_s"]
    rate = 0.7 * capacity
    trace = serve.workload(count, rate)
    res, st = run(trace, serve.BATCHED)
    batched = serve.summary(res, st)
    res2, st2 = run(trace, serve.PLAIN)
    plain = serve.summary(res2, st2)
    # The plain per-call path cannot sustain the batched loop's rate (its
    # queue grows without bound there); it is also run at 70% of its own
    # closed-loop capacity, a rate it sustains, for a p99-vs-p99 comparison.
    _, stc = run(closed, serve.PLAIN)
    plain_capacity = count / stc["wall_s"]
    plain_rate = 0.7 * plain_capacity
    res3, st3 = run(serve.workload(count, plain_rate), serve.PLAIN)
    plain_own = serve.summary(res3, st3)
    keys = ("tflops", "p50_ms", "p99_ms", "wall_s", "batches", "peak_footprint_gb")
    return {"requests": count, "rate_rps": round(rate, 1),
            "capacity_rps": round(capacity, 1), "hash_col_step": SERVE_HASH_COL_STEP,
            "batched": {k: round(batched[k], 3) for k in keys},
            "plain": {k: round(plain[k], 3) for k in keys},
            "plain_capacity_rps": round(plain_capacity, 1),
            "plain_at_own_70pct": dict({"rate_rps": round(plain_rate, 1)},
                                       **{k: round(plain_own[k], 3) for k in keys})}, trace


def serving_batched(count, inflight=4):
    """The continuous-batching runs of serving() alone (closed loop, then
    Poisson at 70% of its capacity), for profiling the batched kernel mix;
    inflight 1 runs one batch at a time, so each kernel's trace duration is
    its own execution time rather than time shared with other batches."""
    from rblas_amd import serve
    run = functools.partial(serve.run, hash_col_step=SERVE_HASH_COL_STEP,
                            max_inflight=inflight)
    res, st = run(serve.workload(count, 0.0), serve.BATCHED)
    rate = 0.7 * count / st["wall_s"]
    res, st = run(serve.workload(count, rate), serve.BATCHED)
    out = serve.summary(res, st)
    out["rate_rps"] = rate
    return {k: round(v, 3) for k, v in out.items()}


def c64_gemm(n=4096):
    from rblas_amd import complex3, workloads
    from rblas_amd.device import DeviceArray, Stream
    s = Stream()
    arrs = []
    for i in range(3):
        x = workloads.normal(13, 2 * i * n * n, 2 * n * n, np.float64).view(np.complex128)
        arrs.append(DeviceArray.from_host(x, s))
    ms = timed(lambda: complex3.gemm_dev(0, 0, n, n, n, 1.5 - 0.5j, arrs[0], n, arrs[1], n,
                                         -0.25, arrs[2], n, s), s, 5, 1)
    for x in arrs:
        x.free()
    s.close()
    return {"tflops": round(8.0 * n ** 3 / ms / 1e9, 2), "ms": round(ms, 3)}


def level1_bw(n=1 << 26):
    """dot / axpy / nrm2 at n = 2^26, unit stride, f64 and f32: GB/s of
    algorithmic bytes (2ns, 3ns, ns)."""
    from rblas_amd import level1, workloads
    from rblas_amd.device import DeviceArray, Stream
    s = Stream()
    out = {}
    for dt, pre in ((np.float64, "d"), (np.float32, "s")):
        sz = np.dtype(dt).itemsize
        x = DeviceArray.from_host(workloads.uniform(3, 0, n, dt), s)
        y = DeviceArray.from_host(workloads.uniform(3, n, n, dt), s)
        r = DeviceArray(1, dt)
        for name, fn, nbytes in (("dot", lambda: level1.dot_dev(n, x, 1, y, 1, r, s), 2 * n * sz),
                                 ("axpy", lambda: level1.axpy_dev(n, 1e-3, x, 1, y, 1, s),
                                  3 * n * sz),
                                 ("nrm2", lambda: level1.nrm2_dev(n, x, 1, r, s), n * sz)):
            ms = timed(fn, s, 20, 3)
            out[pre + name + "_gbs"] = round(nbytes / ms / 1e6, 1)
        for a in (x, y, r):
            a.free()
    s.close()
    return out


def cpu_baseline(n, ld):
    """Oracle DGEMM (rblas's threaded CPU gemm restated: column bands of
    about 4 per worker, each packing its own A panels) on the headline
    problem itself, m = n = k = n, ld, alpha 1.5, beta -0.25, the same
    operands, on the host cores: median of 3 calls after a warm-up call on a
    64-column slice. The former 512-column slice is kept as an extra field:
    its 12-column bands re-pack A far more often per flop than the full
    problem's 132-column bands, so it understates rblas."""
    from oracle import oracle as orc
    from rblas_amd import workloads
    threads = min(os.cpu_count() or 1, 16)
    la = (n - 1) * ld + n
    a = workloads.normal(1, 0, la, np.float64)
    b = workloads.normal(1, la, la, np.float64)
    c = workloads.normal(1, 2 * la, la, np.float64)
    orc.gemm(1, 0, 0, n, 64, n, 1.5, a, ld, b, ld, -0.25, c.copy(), ld, threads=threads)  # warm

    def median_tflops(ncols, reps=3):
        times = []
        for _ in range(reps):
            cc = c.copy()
            t = time.perf_counter()
            orc.gemm(1, 0, 0, n, ncols, n, 1.5, a, ld, b, ld, -0.25, cc, ld, threads=threads)
            times.append(time.perf_counter() - t)
        return 2.0 * n * ncols * n / sorted(times)[len(times) // 2] / 1e12

    full = median_tflops(n)
    slice512 = median_tflops(512)
    return {"value": round(full, 5), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"DGEMM m=n=k={n} col-major NN, ld={ld}, alpha 1.5, beta -0.25 (the "
                      f"headline problem, same operands), median of 3, oracle restatement of "
                      f"rblas's threaded CPU gemm on {threads} threads",
            "slice_8192x512x8192_tflops": round(slice512, 5)}


def cpu_rows(serve_trace, budget_s=10.0):
    """The CPU side of the batching, serving and Level-1 rows, on the same
    synthetic inputs as their GPU figures, timed with the oracle (the
    restatement of rblas's CPU algorithms): gemm threaded as rblas's
    `threads` feature is, trsm and Level-1 serial as rblas's are.

    - grouped GEMM (f64, f32): the 512 problems one after another;
    - batched trsm: the 256 systems, one serial solve per system, spread over
      the worker threads (the systems are independent);
    - serving: the serving trace's requests in arrival order on one FIFO
      server, service times measured and operand generation not timed,
      until 100 requests or `budget_s` of service time;
    - Level-1: dot / axpy / nrm2 at n = 2^26, f64, one core."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc
    from rblas_amd import serve, workloads
    threads = min(os.cpu_count() or 1, 16)
    out = {"cores": threads, "kind": "port"}
    for dtype, name in ((np.float64, "grouped_gemm_f64"), (np.float32, "grouped_gemm_f32")):
        flops, secs = 0.0, 0.0
        for ta, tb, m, n, k, ar, br, (a, b, c) in grouped_gemm_set(dtype):
            t = time.perf_counter()
            orc.gemm(1, ta, tb, m, n, k, 1.0, a, ar, b, br, 0.5, c, m, threads=threads)
            secs += time.perf_counter() - t
            flops += 2.0 * m * n * k
        out[name] = {"tflops": round(flops / secs / 1e12, 4), "ms": round(secs * 1e3, 1)}
    systems = list(grouped_trsm_set())
    flops = sum(float(q) * q * nr for q, nr, _, _ in systems)
    t = time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(lambda p: orc.trsm(0, 1, 0, 0, p[0], p[1], 1.0, p[2], p[0], p[3], p[0]),
                      systems))
    secs = time.perf_counter() - t
    del systems
    out["grouped_trsm_f64"] = {"tflops": round(flops / secs / 1e12, 5), "ms": round(secs * 1e3, 1)}
    free_at, service, flops, lat = 0.0, 0.0, 0.0, []
    for req in serve_trace[:100]:
        if service >= budget_s:
            break
        kind, m, n, k = (int(req[f]) for f in ("kind", "m", "n", "k"))
        ops = serve.operands(req)
        t = time.perf_counter()
        if kind == serve.GEMM:
            orc.gemm(1, 0, 0, m, n, k, 1.0, ops[0], m, ops[1], k, 0.5, ops[2], m,
                     threads=threads)
            f = 2.0 * m * n * k
        elif kind == serve.TRSM:
            orc.trsm(0, 1, 0, 0, m, n, 1.0, ops[0], m, ops[1], m)
            f = float(m) * m * n
        else:
            orc.cgemm(0, 0, m, n, k, 1.0, ops[0], m, ops[1], k, 0.5, ops[2], m, threads=threads)
            f = 8.0 * m * n * k
        d = time.perf_counter() - t
        start = max(float(req["arrival_s"]), free_at)
        free_at = start + d
        lat.append(free_at - float(req["arrival_s"]))
        service += d
        flops += f
    out["serve_prefix"] = {"requests": len(lat), "tflops": round(flops / service / 1e12, 4),
                           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
                           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
                           "service_s": round(service, 2)}
    n = 1 << 26
    x = workloads.uniform(3, 0, n, np.float64)
    y = workloads.uniform(3, n, n, np.float64)
    for name, fn, nbytes in (("ddot", lambda: orc.dot(n, x, 1, y, 1), 16 * n),
                             ("daxpy", lambda: orc.axpy(n, 1e-3, x, 1, y, 1), 24 * n),
                             ("dnrm2", lambda: orc.nrm2(n, x, 1), 8 * n)):
        fn()
        t = time.perf_counter()
        fn()
        out[name + "_gbs_1core"] = round(nbytes / (time.perf_counter() - t) / 1e9, 2)
    return out


ONLY = {
    "sgemm_bf16x3": lambda a: {"ms": square_gemm(np.float32, a.size, a.size + 7, 5, 2,
                                                 compute=1)},
    "sgemm_bf16x6": lambda a: {"ms": square_gemm(np.float32, a.size, a.size + 7, 5, 2,
                                                 compute=2)},
    "sgemm_fp32": lambda a: {"ms": square_gemm(np.float32, a.size, a.size + 7, 5, 2)},
    "grouped_gemm_f64": lambda a: grouped_gemm(np.float64),
    "grouped_gemm_f32": lambda a: grouped_gemm(np.float32),
    "grouped_trsm_f64": lambda a: grouped_trsm(),
    "trsm": lambda a: trsm_single(),
    "symm_hemm": lambda a: symm_hemm(),
    "ctrsm": lambda a: ctrsm_single(),
    "level2": lambda a: level2_bw(),
    "serve": lambda a: serving(a.serve_requests)[0],
    "serve_batched": lambda a: serving_batched(a.serve_requests, a.serve_inflight),
    "c64_gemm": lambda a: c64_gemm(),
    "sgemm_skinny": lambda a: skinny_sgemm_tflops(),
    "level1": lambda a: level1_bw(),
    "dgemm_variants": lambda a: gemm_variants(np.float64),
    "sgemm_variants": lambda a: gemm_variants(np.float32),
    "sgemm_256": lambda a: sgemm_256(),
    "dgemm_host": lambda a: dgemm_host_call(),
}


def main():
    args = parse()
    if args.only:
        print(json.dumps({args.only: ONLY[args.only](args)}), flush=True)
        return
    if args.copy_calibrate:
        copy_calibrate()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is designed but not built)"}))
        sys.exit(2)
    from rblas_amd import workloads
    from rblas_amd.types import Compute
    n = args.size
    ld = n + 7
    flop = 2.0 * n * n * n
    log("dgemm ...")
    ms = square_gemm(np.float64, n, ld, args.steps, args.warmup, min_warm_s=0.0)
    tflops = flop / (ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic N(0,1), splitmix64 seed 1",
        "config": {"workload": f"dgemm col-major NN m=n=k={n} ld={ld} alpha=1.5 beta=-0.25",
                   "pct_fp64_peak": round(100 * tflops / FP64_PEAK, 2)},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": FP64_PEAK,
                     "unit": "TFLOP/s", "frac": round(tflops / FP64_PEAK, 4),
                     "traffic": measured_traffic()},
    }
    if not args.quick:
        comp = {"dgemm_8192_tflops": round(tflops, 3),
                "dgemm_8192_pct_peak": round(100 * tflops / FP64_PEAK, 2)}
        la = (n - 1) * ld + n
        ops32 = (workloads.normal(1, 0, la, np.float32), workloads.normal(1, la, la, np.float32),
                 workloads.normal(1, 2 * la, la, np.float32))
        for name, mode in (("sgemm_fp32", None), ("sgemm_bf16x3", Compute.BF16X3),
                           ("sgemm_bf16x6", Compute.BF16X6)):
            log(f"{name} ...")
            ms32 = square_gemm(np.float32, n, ld, 5, 2, compute=mode, operands=ops32)
            t = flop / (ms32 * 1e-3) / 1e12
            comp[f"{name}_8192_tflops"] = round(t, 2)
            comp[f"{name}_8192_pct_fp32_peak"] = round(100 * t / FP32_PEAK, 2)
        del ops32
        comp["bf16x3_over_fp32_speed"] = round(comp["sgemm_bf16x3_8192_tflops"] /
                                               comp["sgemm_fp32_8192_tflops"], 3)
        comp["bf16x6_over_fp32_speed"] = round(comp["sgemm_bf16x6_8192_tflops"] /
                                               comp["sgemm_fp32_8192_tflops"], 3)
        log("precision ...")
        comp["precision_err_over_bound"] = {c: precision_case(c) for c in ("a", "b", "c")}
        comp["sgemm_skinny_1024x1024x65536_tflops"] = skinny_sgemm_tflops()
        log("variants ...")
        comp["dgemm_8192_variants_tflops"] = gemm_variants(np.float64)
        comp["sgemm_8192_variants_tflops"] = gemm_variants(np.float32)
        comp["sgemm_256"] = sgemm_256()
        comp["dgemm_8192_host_call"] = dgemm_host_call()
        log("grouped gemm ...")
        comp["grouped_gemm_f64"] = grouped_gemm(np.float64)
        comp["grouped_gemm_f32"] = grouped_gemm(np.float32)
        log("trsm ...")
        comp["trsm"] = trsm_single()
        log("symm / hemm ...")
        comp["symm_hemm"] = symm_hemm()
        log("complex trsm ...")
        comp["ctrsm"] = ctrsm_single()
        log("grouped trsm ...")
        comp["grouped_trsm_f64"] = grouped_trsm()
        log("serving ...")
        comp["serve"], trace = serving(args.serve_requests)
        comp["serve_p99_ms"] = comp["serve"]["batched"]["p99_ms"]
        log("c64 / level-1 ...")
        comp["c64_gemm_4096"] = c64_gemm()
        comp["level1"] = level1_bw()
        log("level-2 ...")
        comp["level2"] = level2_bw()
        out["components"] = comp
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(n, ld)
        if not args.quick:
            log("cpu rows ...")
            out["cpu_baseline"]["rows"] = cpu_rows(trace)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
