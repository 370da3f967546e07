"""Benchmark of FCWT::cwt on one MI355X: the 100k x 3000 headline workload.

Workload: n = 100,000 float32 samples (padded to N = 131,072), a seeded sum of
5 unit sinusoids at frequencies drawn from [1, 400] Hz at fs = 1000 plus
N(0, 0.2) noise; Morlet sigma = 2.0; 3000 LOGSCALES frequencies 1-450 Hz;
output 3000 x 100,000 complex64 (2.4 GB) written to HBM.

A step is one full cwt (forward FFT + every scale) on device pointers: the
input is already in HBM and the output buffer preallocated there.  The value
is output GB/s = 8 * fn * n bytes / step time.  The roofline entry reports the
same algorithmic bytes over the summed kernel time of the step, from HIP
events recorded on the launch stream, against MI355X's 8 TB/s HBM peak.
The CPU baseline is the oracle's float32 port (numpy float32 FFTs plus the
reference's window multiply in C) on one host core, timed on a bounded
sample of the same workload (the first `--cpu-rows` scales).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-rows R]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "CWT output GB/s (% HBM peak) at 100k x 3000; batched/adjoint vs plain GPU"


def headline_signal(n=100_000, fs=1000, seed=2024):
    rng = np.random.default_rng(seed)
    f = rng.uniform(1.0, 400.0, 5)
    ph = rng.uniform(0, 2 * np.pi, 5)
    t = np.arange(n, dtype=np.float64) / fs
    x = np.sin(2 * np.pi * f[:, None] * t[None, :] + ph[:, None]).sum(axis=0)
    return (x + 0.2 * rng.standard_normal(n)).astype(np.float32)


def cpu_baseline(x, scales, fb, rows):
    """Oracle float32 port on one core over the first `rows` scales."""
    from oracle import oracle as O

    out = np.empty((rows, x.size), np.complex64)
    O.cwt_fp32_port(x, scales[:rows], fb, rows=rows, out=out)  # warm (imports, tables)
    t0 = time.perf_counter()
    O.cwt_fp32_port(x, scales[:rows], fb, rows=rows, out=out)
    dt = time.perf_counter() - t0
    return 8.0 * rows * x.size / dt / 1e9, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-rows", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU: FCWT::cwt runs on a single MI355X here "
                 "(multi-GPU sharding is described in DESIGN.md, not run); use --gpus 1")

    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import fcwt_amd

    n, fs, f0, f1, fn, fb = 100_000, 1000, 1.0, 450.0, 3000, 2.0
    x = headline_signal(n, fs)
    morl = fcwt_amd.Morlet(fb)
    scs = fcwt_amd.Scales(morl, fcwt_amd.FCWT_LOGSCALES, fs, f0, f1, fn)
    plan = fcwt_amd.Plan(n, scs.scales, fb, threads=1, normalize=True)
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    out = torch.empty((fn, n), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        plan.execute_device(dx.data_ptr(), out.data_ptr(), stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.profile(True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / args.steps
    prof = plan.profile_read()
    plan.profile(False)

    out_bytes = 8.0 * fn * n
    in_bytes = 4.0 * n
    value = out_bytes / (ms_step * 1e-3) / 1e9
    kernel_ms = sum(ms for ms, _ in prof.values()) / args.steps
    achieved = (out_bytes + in_bytes) / (kernel_ms * 1e-3) / 1e9
    kernels = {k: {"ms_per_step": ms / args.steps, "launches_per_step": c / args.steps} for k, (ms, c) in prof.items()}

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cwt 100k x 3000: n=100000 (N=131072) float32, 5 tones + N(0,0.2) noise, fs=1000, "
                               "Morlet sigma=2, 3000 LOGSCALES 1-450 Hz, complex64 output 2.4 GB",
                   "n": n, "fn": fn, "N": plan.N, "samples_scales_per_s": round(fn * n / (ms_step * 1e-3), 1),
                   "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole cwt step (all kernels, summed event time)", "kernels": kernels},
    }
    if not args.no_cpu:
        rows = min(args.cpu_rows, fn)
        gbs, dt = cpu_baseline(x, scs.scales, fb, rows)
        result["cpu_baseline"] = {"value": round(gbs, 4), "unit": "GB/s", "cores": 1, "kind": "port",
                                  "sample": f"first {rows} of {fn} scales of the same workload "
                                            f"({dt:.1f} s; oracle float32 port: numpy float32 FFTs + reference "
                                            f"window multiply in C)"}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
