"""Autotuner for the grouped-gemm planner (runs on the GPU).

The planner (csrc/grouped_api.cpp) prices a work unit of tile side t over
kt k-tiles as c0[t] + c1[t] * kt microseconds, picks per problem the tile
side with the smaller total, and splits the k of units dearer than
unit_cap_us. This module measures those numbers on the device:

  1. for each dtype and tile side, groups of identical problems whose units
     all have kt k-tiles (kt over a geometric range) run in one persistent
     launch; the CU-time per unit (launch time * resident workgroups / units)
     is fitted by least squares to c0 + c1 * kt;
  2. with that model, unit_cap_us is chosen among candidates by timing the
     batching workload (512 problems, m, n, k log-uniform in [16, 2048]).

`python -m rblas_amd.autotune` prints the fitted table and writes it to
rblas_amd/tuning_gfx950.json; `load()` installs a saved table through
rblas_grouped_set_tuning. Forcing one tile side or no split during the
measurement uses the same cost-model interface (a prohibitive price on the
other choice).
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from . import grouped
from .device import DeviceArray, Event, Stream

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning_gfx950.json")
BIG = 1e9


def _time(fn, stream, reps, min_warm_s=0.05):
    """ms per call over `reps` calls, after warm-up calls filling at least
    min_warm_s seconds (short launches alone leave the clocks low)."""
    t0 = time.perf_counter()
    fn()
    stream.synchronize()
    while time.perf_counter() - t0 < min_warm_s:
        fn()
        stream.synchronize()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    stream.synchronize()
    ms = e1.elapsed_ms(e0) / reps
    e0.close()
    e1.close()
    return ms


def _uniform_group(dtype, tile, kt, units_wanted, stream):
    """Problems of 2 x 2 tiles each, k = kt k-tiles, enough of them for
    `units_wanted` units; operands shared (read-only A, B; C per problem)."""
    bk = 32 if np.dtype(dtype) == np.float32 else 16
    m = n = 2 * tile
    k = kt * bk
    count = max(1, units_wanted // 4)
    rng = np.random.default_rng(kt)
    a = DeviceArray.from_host(rng.uniform(-1, 1, m * k).astype(dtype), stream)
    b = DeviceArray.from_host(rng.uniform(-1, 1, k * n).astype(dtype), stream)
    cs = [DeviceArray(m * n, dtype) for _ in range(count)]
    probs = [grouped.GemmProblem(1, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m) for c in cs]
    return probs, [a, b] + cs


def measure_unit_costs(dtype, tile, kts=(1, 2, 4, 8, 16, 32, 64, 128), units=4096, reps=5):
    """[(kt, microseconds of CU time per unit)] for one dtype and tile side."""
    s = Stream()
    saved = grouped.get_tuning(dtype)
    # Force the tile side and no split through the cost model.
    forced = [1.0, 1.0, BIG, 1.0, BIG] if tile == 64 else [BIG, 1.0, 1.0, 1.0, BIG]
    out = []
    try:
        grouped.set_tuning(dtype, forced)
        for kt in kts:
            probs, bufs = _uniform_group(dtype, tile, kt, units, s)
            plan = grouped.GroupedGemm(probs, dtype)
            info = plan.info()
            assert info["split_problems"] == 0
            assert (info["tiles64"] if tile == 64 else info["tiles128"]) == info["units"]
            ms = _time(lambda: plan.run(s), s, reps)
            out.append((kt, ms * 1e3 * info["grid"] / info["units"]))
            plan.close()
            for x in bufs:
                x.free()
    finally:
        grouped.set_tuning(dtype, saved)
        s.close()
    return out


def fit(points):
    """Least-squares c0, c1 of us = c0 + c1 * kt (both kept positive)."""
    x = np.array([p[0] for p in points], float)
    y = np.array([p[1] for p in points], float)
    c1, c0 = np.polyfit(x, y, 1)
    return max(c0, 0.1), max(c1, 1e-3)


def batching_workload(dtype, stream, seed=4, count=512):
    """The batching workload: m, n, k log-uniform in [16, 2048], random
    NoTrans/Trans on each operand, U[-1, 1] entries, alpha 1, beta 0.5."""
    rng = np.random.default_rng(seed)
    probs, bufs, flops = [], [], 0.0
    for _ in range(count):
        m, n, k = (int(np.exp(rng.uniform(np.log(16), np.log(2049)))) for _ in range(3))
        ta, tb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        ar, ac = (k, m) if ta else (m, k)
        br, bc = (n, k) if tb else (k, n)
        a = DeviceArray.from_host(rng.uniform(-1, 1, ar * ac).astype(dtype), stream)
        b = DeviceArray.from_host(rng.uniform(-1, 1, br * bc).astype(dtype), stream)
        c = DeviceArray.from_host(rng.uniform(-1, 1, m * n).astype(dtype), stream)
        probs.append(grouped.GemmProblem(1, ta, tb, m, n, k, 1.0, a, ar, b, br, 0.5, c, m))
        bufs += [a, b, c]
        flops += 2.0 * m * n * k
    return probs, bufs, flops


def tune_cap(dtype, base, caps=(20, 40, 80, 160, 320, 1e6), reps=5):
    """Pick unit_cap_us on the batching workload given the fitted costs."""
    s = Stream()
    probs, bufs, _ = batching_workload(dtype, s)
    best = None
    results = []
    saved = grouped.get_tuning(dtype)
    try:
        for cap in caps:
            grouped.set_tuning(dtype, list(base[:4]) + [cap])
            plan = grouped.GroupedGemm(probs, dtype)
            ms = _time(lambda: plan.run(s), s, reps)
            plan.close()
            results.append((cap, ms))
            if best is None or ms < best[1]:
                best = (cap, ms)
    finally:
        grouped.set_tuning(dtype, saved)
        for x in bufs:
            x.free()
        s.close()
    return best[0], results


def tune(verbose=True) -> dict:
    table = {}
    for dtype, name in ((np.float32, "f32"), (np.float64, "f64")):
        coeffs = []
        for tile in (64, 128):
            pts = measure_unit_costs(dtype, tile)
            c0, c1 = fit(pts)
            coeffs += [c0, c1]
            if verbose:
                print(f"{name} tile {tile}: " + ", ".join(f"kt {k}: {u:.2f} us" for k, u in pts)
                      + f" -> c0 {c0:.2f} c1 {c1:.3f}", flush=True)
        cap, results = tune_cap(dtype, coeffs)
        if verbose:
            print(f"{name} unit cap: " + ", ".join(f"{c:g}: {ms:.3f} ms" for c, ms in results)
                  + f" -> {cap:g}", flush=True)
        table[name] = coeffs + [float(cap)]
    return table


def load(path: str = TABLE) -> bool:
    """Install a saved table; False if there is none."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        table = json.load(f)
    for name, dtype in (("f32", np.float32), ("f64", np.float64)):
        if name in table:
            grouped.set_tuning(dtype, table[name])
    return True


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else TABLE
    t = tune()
    with open(out, "w") as f:
        json.dump(t, f, indent=1)
    print(f"wrote {out}: {t}")
