#!/usr/bin/env python3
"""Benchmark: DGEMM 8192^3 on one MI355X through librblas_hip.so.

Workload (one "step" = one gemm call on HBM-resident operands):
  C := 1.5 * A * B - 0.25 * C, f64, column-major NoTrans/NoTrans,
  m = n = k = 8192, lda = ldb = ldc = 8199 (rows + 7), entries N(0,1) from
  the seeded splitmix64 generator (rblas_amd.workloads), seed 1, drawn in
  the order A, B, C.

Prints one JSON line. `value` is whole-job TFLOP/s (2mnk flop per call) over
the K timed calls, timed with HIP events on the stream the kernel runs on.
`roofline` prices the gemm kernel against MI355X's FP64 MFMA peak;
`cpu_baseline` times the oracle (rblas's CPU algorithm restated in C++) on a
column slice of the same problem on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--extra", action="store_true",
                   help="also time SGEMM (plain FP32) at the same size")
    p.add_argument("--copy-calibrate", action="store_true",
                   help="only run 5 device-to-device copies of 1 GiB (counter calibration)")
    return p.parse_args()


PROFILE = os.path.join(ROOT, "profiles", "dgemm_8192_nn.md")


def measured_traffic():
    """HBM bytes per gemm launch from the committed rocprofv3 PMC summary
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


def gen_operands(n, ld, dtype, seed=1):
    from rblas_amd import workloads
    la = (n - 1) * ld + n
    a = workloads.normal(seed, 0, la, dtype)
    b = workloads.normal(seed, la, la, dtype)
    c = workloads.normal(seed, 2 * la, la, dtype)
    return a, b, c


def time_gemm(dtype, n, ld, steps, warmup):
    from rblas_amd import level3
    from rblas_amd.device import DeviceArray, Event, Stream
    a, b, c = gen_operands(n, ld, dtype)
    s = Stream()
    da, db, dc = (DeviceArray.from_host(x, s) for x in (a, b, c))
    del a, b, c
    args = (1, 0, 0, n, n, n, 1.5, da, ld, db, ld, -0.25, dc, ld, s)
    for _ in range(warmup):
        level3.gemm_dev(*args)
    s.synchronize()
    e0, e1 = Event(), Event()
    e0.record(s)
    for _ in range(steps):
        level3.gemm_dev(*args)
    e1.record(s)
    s.synchronize()
    ms = e1.elapsed_ms(e0)
    for x in (da, db, dc):
        x.free()
    s.close()
    return ms / steps


def cpu_baseline(n, ld):
    """Oracle DGEMM on a 8192 x 512 x 8192 column slice of the same problem,
    threaded over the host cores (rblas's column-band driver restated)."""
    from oracle import oracle as orc
    from rblas_amd import workloads
    ncols = 512
    threads = min(os.cpu_count() or 1, 16)
    la = (n - 1) * ld + n
    a = workloads.normal(1, 0, la, np.float64)
    b = workloads.normal(1, la, (ncols - 1) * ld + n, np.float64)
    c = workloads.normal(1, 2 * la, (ncols - 1) * ld + n, np.float64)
    orc.gemm(1, 0, 0, n, 64, n, 1.5, a, ld, b, ld, -0.25, c.copy(), ld, threads=threads)  # warm
    times = []
    for _ in range(3):
        cc = c.copy()
        t = time.perf_counter()
        orc.gemm(1, 0, 0, n, ncols, n, 1.5, a, ld, b, ld, -0.25, cc, ld, threads=threads)
        times.append(time.perf_counter() - t)
    t = sorted(times)[1]
    tflops = 2.0 * n * ncols * n / t / 1e12
    return {"value": round(tflops, 5), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"DGEMM m=k={n}, n={ncols} column slice (same operands/ld/alpha/beta), "
                      f"median of 3, oracle restatement of rblas's threaded CPU gemm"}


def main():
    args = parse()
    if args.copy_calibrate:
        copy_calibrate()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)
    n = args.size
    ld = n + 7
    flop = 2.0 * n * n * n
    ms = time_gemm(np.float64, n, ld, args.steps, args.warmup)
    tflops = flop / (ms * 1e-3) / 1e12
    out = {
        "metric": "DGEMM 8192^3 TFLOP/s (f64, MFMA)",
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
                   "pct_fp64_peak": round(100 * tflops / FP64_MFMA_PEAK_TFLOPS, 2)},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3),
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / FP64_MFMA_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic()},
    }
    if args.extra:
        ms32 = time_gemm(np.float32, n, ld, args.steps, args.warmup)
        t32 = flop / (ms32 * 1e-3) / 1e12
        out["extra"] = {"sgemm_fp32_tflops": round(t32, 3),
                        "sgemm_fp32_pct_peak": round(100 * t32 / FP32_MFMA_PEAK_TFLOPS, 2)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, ld)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
