"""Autotuner of the batched engine: the kernel of each LU pair class.

For every tunable class -- 3 (16 < n <= 32) and 4 (32 < n <= 64) -- and both
factor precisions (the FP32 factors of a mixed batch, the FP64 factors of a
plain batch and of the mixed batch's FP64 fallback), every candidate kernel
is timed on a sample batch drawn like the default workload (n uniform in
4..48, N log-uniform in [16, 4096]): one eager factorization per candidate
with HIP events around every pair-elimination launch (Batch.time_factorize),
the class's launch times summed, the median of `reps` factorizations kept.
The fastest candidate per (precision, class) goes to alglin_amd/tuning.json,
which Batch.plan applies through alglin_batch_set_class_kernel.

Candidates (alglin_batch_set_class_kernel codes):
  class 3, FP32 factors: the matrix-core pair (0), the two-wave split (1)
  class 4, FP32 factors: the matrix-core pair (0), column-split waves with
                         4 / 8 / 16 waves per pair (1), register tiles (2)
  class 3, FP64 factors: the matrix-core pair (0), the two-wave split (1)
  class 4, FP64 factors: the matrix-core pair (0, n <= 48), column-split
                         waves 4 / 8 / 16 (1), register tiles (2)

Usage: python -m alglin_amd.autotune [--nsys 1024] [--seed 3] [--reps 3] [--out alglin_amd/tuning.json]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import systems
from .bcr import TUNING_PATH, Batch, device_info

CANDIDATES = {
    (True, 3): [(0, 0), (1, 0)],
    (True, 4): [(0, 0), (1, 4), (1, 8), (1, 16), (2, 0)],
    (False, 3): [(0, 0), (1, 0)],
    (False, 4): [(0, 0), (1, 4), (1, 8), (1, 16), (2, 0)],
}


def class_ms(bt: Batch, cls: int, reps: int) -> float:
    """Median over reps of the summed launch times of class `cls` in one eager
    factorization (after one untimed factorization)."""
    bt.time_factorize()
    ts = []
    for _ in range(reps):
        rows = bt.time_factorize()
        ts.append(float(rows[rows[:, 1] == cls, 3].sum()))
    return float(np.median(ts))


def tune(nsys=1024, seed=3, reps=3, log=sys.stderr):
    dims = systems.uneven_batch_dims(nsys, seed)
    out = {"device": device_info(), "workload": f"{nsys} systems, n uniform in 4..48, N log-uniform [16,4096], "
                                                f"qr=qx=n/4, nr=nx=2, seed {seed}",
           "method": f"median of {reps} eager factorizations, HIP events per launch", "fp32_factors": {},
           "fp64_factors": {}}
    for fp32 in (True, False):
        bt = Batch(mixed=fp32)
        bt.plan(dims, tuning=False)
        bt.generate(seed=seed, shift=4.0)
        for (prec, cls), cands in CANDIDATES.items():
            if prec != fp32:
                continue
            if not ((dims[:, 1] > (16 if cls == 3 else 32)) & (dims[:, 1] <= (32 if cls == 3 else 64))).any():
                continue
            times = {}
            for kernel, waves in cands:
                bt.set_class_kernel(cls, kernel, waves)
                t0 = time.perf_counter()
                times[f"{kernel}:{waves}"] = class_ms(bt, cls, reps)
                print(f"[autotune] {'fp32' if fp32 else 'fp64'} class {cls} kernel {kernel} waves {waves}: "
                      f"{times[f'{kernel}:{waves}']:.2f} ms ({time.perf_counter() - t0:.1f} s)", file=log, flush=True)
            best = min(times, key=times.get)
            k, w = (int(v) for v in best.split(":"))
            out["fp32_factors" if fp32 else "fp64_factors"][str(cls)] = dict(kernel=k, waves=w, ms=times[best],
                                                                             candidates_ms=times)
            bt.set_class_kernel(cls, k, w)
        del bt
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nsys", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=TUNING_PATH)
    a = ap.parse_args()
    res = tune(a.nsys, a.seed, a.reps)
    with open(a.out + ".tmp", "w") as f:
        json.dump(res, f, indent=1)
    os.replace(a.out + ".tmp", a.out)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
