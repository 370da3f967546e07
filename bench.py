"""Benchmark of FCWT::cwt on one MI355X: the 100k x 3000 headline workload.

Workload: n = 100,000 float32 samples (padded to N = 131,072), a seeded sum of
5 unit sinusoids at frequencies drawn from [1, 400] Hz at fs = 1000 plus
N(0, 0.2) noise; Morlet sigma = 2.0; 3000 LOGSCALES frequencies 1-450 Hz;
output 3000 x 100,000 complex64 (2.4 GB) written to HBM.

A step is one full cwt (forward FFT + every scale) on device pointers: the
input is already in HBM and the output buffer preallocated there.  The value
is output GB/s = 8 * fn * n bytes / step time (HIP events around the K steps,
no per-kernel events inside them).  The default is K = 200 timed steps after
W = 50 untimed ones: the GPU's clocks take some 20-30 steps (about 40 ms of
work) to settle, and a run of 20 steps after 3 reads 1.38 ms per step where
the settled rate is 1.25 ms.

The roofline entry is for the dominant kernel (the one with the most time):
its algorithmic bytes per launch (8 B per complex64 output value it writes,
i.e. 8 * n * rows it owns; 4 n + 4 N for the forward FFT) over its average
launch time from HIP events recorded around each launch on the launch stream
(in a separate pass of up to 10 steps after the timed ones), against
MI355X's 8 TB/s HBM peak.  `traffic` is the HBM bytes per launch of
that kernel from rocprofv3 FETCH_SIZE and WRITE_SIZE counter runs, each scaled
by what the counter reads for a device copy of known size; it is read from
the summary that `--summarize-profiles` wrote under profiles/ (null if absent).

The CPU baseline is the oracle's float32 port (the reference's window
multiply in C, then scipy float32 inverse FFTs) on the host's threads (a power
of two, at most 16: one GPU's share of its host) over the first `--cpu-rows`
scales of the same workload (all 3000 by default).  The reference's own
FFTW+OpenMP build cannot be made here (FFTW exists only as prebuilt binaries
inside the reference).

The metric's other half, the batched and adjoint forms against the plain
per-signal GPU call, is measured after the headline on the batched and
adjoint workloads below (5 timed steps after 3 each) and reported in the same
line under config.secondary (`--no-secondary` skips it).

Other workloads (each prints its own JSON line; not the headline metric):
  --workload long      n = 2^22 AR(2) + 3-segment tone, Morlet sigma 4, 1024
                       LINFREQS over [0.001, 0.5] fs (34.4 GB of output)
  --workload batched   384 channels, n ~ U[20k, 200k], fs 30 kHz, pink noise +
                       biphasic spikes, sigma 6, 200 LOGSCALES 10 Hz-10 kHz
  --workload adjoint   64 x 16,384 signals, 128 LOGSCALES over [0.01, 0.45] fs:
                       the adjoint (input gradient) against the plain forward
  --workload host      the headline workload through the host-pointer call
                       (fcwt_hip_cwt: input copy, kernels, 2.4 GB back to a
                       numpy array), pipelined and as one plain copy

Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-rows R] [--no-cpu] [--no-tune] [--no-secondary]
  python bench.py --copy-probe MIB     one device copy of MIB MiB (counter calibration)
  python bench.py --summarize-profiles DIR
        write profiles/ summaries from the rocprofv3 CSV output under DIR:
        trace/ (--kernel-trace --stats) and pmc_fetch/, pmc_write/,
        copy_fetch/, copy_write/ (--pmc runs)
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


def long_signal(n=1 << 22, fs=1000, seed=2025):
    """AR(2) (x[t] = 1.6 x[t-1] - 0.8 x[t-2] + e[t]) plus a tone whose
    frequency is constant on three equal segments."""
    from scipy.signal import lfilter

    rng = np.random.default_rng(seed)
    ar = lfilter([1.0], [1.0, -1.6, 0.8], rng.standard_normal(n))
    f = np.repeat(rng.uniform(0.01, 0.4, 3) * fs, -(-n // 3))[:n]
    phase = 2 * np.pi * np.cumsum(f) / fs
    x = ar / ar.std() + np.sin(phase)
    return x.astype(np.float32)


def batched_channels(count=384, fs=30_000, seed=2026):
    """Lengths ~ U[20k, 200k]; pink (1/f) noise plus biphasic spikes at
    Poisson 20 Hz with amplitude 5 sigma."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(20_000, 200_001, count)
    chans = []
    for n in lengths:
        spec = np.fft.rfft(rng.standard_normal(n))
        f = np.arange(spec.size)
        spec[1:] /= np.sqrt(f[1:])
        spec[0] = 0
        x = np.fft.irfft(spec, n)
        x /= x.std()
        for t in np.flatnonzero(rng.random(n) < 20.0 / fs):
            x[t: t + 15] += 5.0 * np.sin(np.linspace(0, 2 * np.pi, 15))[: n - t]
        chans.append(x.astype(np.float32))
    return chans


def adjoint_signals(B=64, n=16_384, fn=128, fs=1000, seed=2024):
    """N(0,1) plus a random linear chirp per signal; loss weights U[0,1]."""
    rng = np.random.default_rng(seed)
    t = np.arange(n) / fs
    xs = []
    for _ in range(B):
        f0, f1 = rng.uniform(5, 450, 2)
        xs.append(rng.standard_normal(n) + np.sin(2 * np.pi * (f0 + (f1 - f0) * t / (2 * t[-1])) * t))
    return np.stack(xs).astype(np.float32), rng.uniform(0, 1, (fn, n)).astype(np.float32)


# A batch captures its launches as a HIP graph on the second call with the
# same buffers (a one-time cost of a few ms) and replays it from then on; its
# warm-up therefore runs at least two calls, so the timed steps are replays.
BATCH_WARMUP = 2


def _timed(torch, stream, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _tune(N, effort):
    """fcwt_hip_tune for N into a temporary file, loaded into this process's
    table; returns (the file's fields, seconds spent)."""
    import tempfile

    from fcwt_amd import _native as nat

    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, f"n{N}_t1.fcwt_hip").encode()
        t0 = time.perf_counter()
        nat.check(nat.lib().fcwt_hip_tune(N, 1, effort, path), "fcwt_hip_tune")
        tune_s = time.perf_counter() - t0
        with open(path) as f:
            chosen = dict(line.split(None, 1) for line in f.read().strip().splitlines()[1:])
        nat.check(nat.lib().fcwt_hip_tune_load(path), "fcwt_hip_tune_load")
    return chosen, tune_s


def workload_long(args):
    import torch
    import fcwt_amd

    n, fs, fn, fb = 1 << 22, 1000, 1024, 4.0
    x = long_signal(n, fs)
    sc = fcwt_amd.Scales(fcwt_amd.Morlet(fb), fcwt_amd.FCWT_LINFREQS, fs, 0.001 * fs, 0.5 * fs, fn)
    plan = fcwt_amd.Plan(n, sc.scales, fb)
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((fn, n), dtype=torch.complex64, device="cuda")
    st = torch.cuda.current_stream()
    ms = _timed(torch, st, lambda: plan.execute_device(dx.data_ptr(), out.data_ptr(), st.cuda_stream),
                args.steps, args.warmup)
    plan.check()
    # the per-kernel split from one more step with event timing (kept out of the timed steps)
    plan.profile(True)
    plan.execute_device(dx.data_ptr(), out.data_ptr(), st.cuda_stream)
    prof = plan.profile_read()
    plan.profile(False)
    steps = 1
    tuned = None
    if not args.no_tune:
        # the analogue of create_FFT_optimization_plan for N = 2^22
        # (fcwt_hip_tune, FFTW_MEASURE effort: the column path's scratch per
        # batch, launch fusion and the split 512 x 2^13 or 256 x 2^14), then a
        # new plan timed the same way
        chosen, tune_s = _tune(plan.N, 1)
        tplan = fcwt_amd.Plan(n, sc.scales, fb)
        tms = _timed(torch, st, lambda: tplan.execute_device(dx.data_ptr(), out.data_ptr(), st.cuda_stream),
                     args.steps, args.warmup)
        tplan.check()
        tplan.close()
        tuned = {"ms_per_step": round(tms, 3), "vs_default": round(ms / tms, 4), "tune_s": round(tune_s, 2),
                 "schedule": {k: chosen[k].strip() for k in ("cols_mb", "cols_sets", "cols_log2B", "path") if k in chosen}}
    gbs = 8.0 * fn * n / (ms * 1e-3) / 1e9
    return {"metric": "long-signal CWT output GB/s (n=2^22 x 1024)", "value": round(gbs, 2), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "n=4194304 AR(2)+3-segment tone, Morlet sigma=4, 1024 LINFREQS 1-500 Hz at "
                                   "fs=1000, complex64 output 34.4 GB", "rows": plan.row_split(),
                       "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2), "tuned_schedule": tuned},
            "kernels": {k: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps} for k, v in prof.items()}}


def workload_batched(args):
    import torch
    import fcwt_amd

    fs, fn, fb = 30_000, 200, 6.0
    chans = batched_channels(fs=fs)
    morl = fcwt_amd.Morlet(fb)
    sc = fcwt_amd.Scales(morl, fcwt_amd.FCWT_LOGSCALES, fs, 10.0, 10_000.0, fn).scales
    plans = [fcwt_amd.Plan(c.size, sc, fb) for c in chans]
    total = sum(c.size for c in chans)
    xs = [torch.from_numpy(c).cuda() for c in chans]
    out = torch.empty(fn * total, dtype=torch.complex64, device="cuda")
    offs = np.concatenate([[0], np.cumsum([fn * c.size for c in chans])[:-1]])
    st = torch.cuda.current_stream()
    base = out.data_ptr()

    def plain():
        for p, x, o in zip(plans, xs, offs):
            p.execute_device(x.data_ptr(), base + 8 * int(o), st.cuda_stream)

    ms_plain = _timed(torch, st, plain, args.steps, args.warmup)
    for p in plans:
        p.check()
    for p in plans:
        p.close()
    batch = fcwt_amd.Batch([c.size for c in chans], sc, fb)
    x_ptrs = [x.data_ptr() for x in xs]
    o_ptrs = [base + 8 * int(o) for o in offs]
    ms = _timed(torch, st, lambda: batch.execute_device(x_ptrs, o_ptrs, st.cuda_stream), args.steps,
                max(args.warmup, BATCH_WARMUP))
    batch.check()
    gbs = 8.0 * fn * total / (ms * 1e-3) / 1e9
    return {"metric": "batched CWT output GB/s (384 uneven channels x 200 scales)", "value": round(gbs, 2),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"384 channels, n in [20000, 200000] (sum {total}), fs=30000, pink noise + "
                                   f"spikes, Morlet sigma=6, 200 LOGSCALES 10-10000 Hz, "
                                   f"{8 * fn * total / 1e9:.1f} GB output", "mode": "one batch execute",
                       "plans": batch.shape(), "graph_replays": batch.graph_launches(), "plain_ms": round(ms_plain, 3),
                       "batched_vs_plain_speedup": round(ms_plain / ms, 3)}}


def workload_adjoint(args):
    import torch
    import fcwt_amd

    B, n, fn, fs, fb = 64, 16_384, 128, 1000, 2.0
    x, w = adjoint_signals(B, n, fn, fs)
    sc = fcwt_amd.Scales(fcwt_amd.Morlet(fb), fcwt_amd.FCWT_LOGSCALES, fs, 0.01 * fs, 0.45 * fs, fn).scales
    plan = fcwt_amd.Plan(n, sc, fb)
    dx = torch.from_numpy(x).cuda()
    W = torch.empty((B, fn, n), dtype=torch.complex64, device="cuda")
    G = torch.randn((B, fn, n), dtype=torch.complex64, device="cuda")
    xbar = torch.empty((B, n), dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()

    batch = fcwt_amd.Batch([n] * B, sc, fb)
    x_ptrs = [dx[b].data_ptr() for b in range(B)]
    w_ptrs = [W[b].data_ptr() for b in range(B)]

    def fwd_plain():
        for b in range(B):
            plan.execute_device(dx[b].data_ptr(), W[b].data_ptr(), st.cuda_stream)

    def adj():
        plan.adjoint_batch_device(G.data_ptr(), xbar.data_ptr(), B, st.cuda_stream)

    ms_f = _timed(torch, st, fwd_plain, args.steps, args.warmup)
    ms_fb = _timed(torch, st, lambda: batch.execute_device(x_ptrs, w_ptrs, st.cuda_stream), args.steps,
                   max(args.warmup, BATCH_WARMUP))
    ms_a = _timed(torch, st, adj, args.steps, args.warmup)
    plan.check()
    batch.check()
    gbs = 8.0 * B * fn * n / (ms_a * 1e-3) / 1e9
    return {"metric": "adjoint gradient-read GB/s (64 x 16384 x 128)", "value": round(gbs, 2), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_a, 3),
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "64 signals n=16384, N(0,1)+chirp, Morlet sigma=2, 128 LOGSCALES 10-450 Hz, "
                                   "fs=1000; adjoint reads 1.07 GB of output gradient",
                       "plain_forward_ms": round(ms_f, 3), "batched_forward_ms": round(ms_fb, 3),
                       "adjoint_vs_plain_forward_speedup": round(ms_f / ms_a, 3),
                       "batched_forward_vs_plain_speedup": round(ms_f / ms_fb, 3)}}


def workload_host(args):
    import fcwt_amd

    n, fs, f0, f1, fn, fb = 100_000, 1000, 1.0, 450.0, 3000, 2.0
    x = headline_signal(n, fs)
    sc = fcwt_amd.Scales(fcwt_amd.Morlet(fb), fcwt_amd.FCWT_LOGSCALES, fs, f0, f1, fn)
    plan = fcwt_amd.Plan(n, sc.scales, fb)
    out = np.zeros((fn, n), np.complex64)
    res = {}
    for mode, env in (("pipelined", "0"), ("plain_copy", "1")):
        os.environ["FCWT_HIP_NO_PIPELINE"] = env
        for _ in range(args.warmup):
            plan.execute_host(x, out)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute_host(x, out)
        res[mode] = (time.perf_counter() - t0) / args.steps * 1e3
    os.environ.pop("FCWT_HIP_NO_PIPELINE", None)
    gbs = 8.0 * fn * n / (res["pipelined"] * 1e-3) / 1e9
    return {"metric": "host-pointer CWT output GB/s (100k x 3000, PCIe-inclusive)", "value": round(gbs, 2),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(res["pipelined"], 3), "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "the 100k x 3000 workload through fcwt_hip_cwt into a 2.4 GB numpy array",
                       "plain_copy_ms": round(res["plain_copy"], 3),
                       "pipelined_vs_plain_speedup": round(res["plain_copy"] / res["pipelined"], 3)}}


def host_threads(cap=16):
    """Host threads for the CPU baseline: the largest power of two <= the
    CPUs this process may run on, at most `cap` (a GPU's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    n = max(1, min(cap, n))
    return 1 << (n.bit_length() - 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(x, scales, fb, rows, threads):
    """Oracle float32 port (the reference's window multiply with its O1
    reuse, then a full N-point float32 inverse FFT per scale) over the first
    `rows` scales, on `threads` host threads."""
    from oracle import oracle as O

    out = np.empty((rows, x.size), np.complex64)
    O.cwt_fp32_port(x, scales[:min(rows, 64)], fb, rows=min(rows, 64), out=out, workers=threads)  # warm
    t0 = time.perf_counter()
    O.cwt_fp32_port(x, scales[:rows], fb, rows=rows, out=out, workers=threads)
    dt = time.perf_counter() - t0
    return 8.0 * rows * x.size / dt / 1e9, dt


# rocprofv3 kernel names of each kernel kind the plan profiles
KIND_KERNELS = {
    "forward_fft": ("k_fwd_small", "k_fwd_a", "k_fwd_b", "k_fwd_post"),
    "row_writer_direct": ("k_rows_direct",),
    "row_team": ("k_rows_team",),
    "row_spreader": ("k_spreader",),
    "row_writer_two_pass": ("k_writer",),
    "row_columns": ("k_cols_step",),
    "column_transpose": ("k_spec_transpose",),
}
WORKLOAD_TAG = "100k_x_3000"


def kind_bytes(kind, n, N, split, direct_in_team=False):
    """Algorithmic HBM bytes one launch of `kind` must move in this workload.
    direct_in_team: the direct rows ran inside the team launch (its members
    take them once their team sequence is claimed; the plan then times no
    direct-row launch)."""
    if kind == "forward_fft":
        return 4.0 * n + 4.0 * N  # signal read + N/2 complex64 spectrum written
    rows = {"row_writer_direct": split["direct"],
            "row_team": split["team"] + (split["direct"] if direct_in_team else 0),
            "row_writer_two_pass": split["two_pass"], "row_columns": split["two_pass"]}.get(kind, 0)
    return 8.0 * n * rows


def traffic_summary(kind):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{kind}_{WORKLOAD_TAG}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB: the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    nbytes = int(mib) << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}))


def _pmc_per_dispatch(csv_path, counter, names):
    """Mean over dispatches of kernels whose name contains one of `names` of
    the counter's value (summed over the rows of one dispatch)."""
    import csv

    per = {}
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter or not any(k in r["Kernel_Name"] for k in names):
                continue
            key = r["Dispatch_Id"]
            per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    if not per:
        raise SystemExit(f"{csv_path}: no {counter} rows for {names}")
    return sum(per.values()) / len(per), len(per)


TRACE_CMD = ("cd /tmp && export TMPDIR=/tmp && timeout -k 10 300 rocprofv3 --kernel-trace --stats --output-format csv "
             "-d $GRAFT_REPO_ROOT/gpurun_out/prof/{d} -o run -- python3 $GRAFT_REPO_ROOT/bench.py {a}")

# kernel traces summarized under profiles/: trace directory, file tag, title, bench.py arguments, note
TRACES = [
    ("trace", f"cwt_{WORKLOAD_TAG}", "cwt 100k x 3000", "--steps 100 --warmup 50 --no-cpu --no-tune --no-secondary",
     "Calls include the 50 warm-up steps, the 100 timed steps and the 10 steps of the per-kernel event pass of "
     "bench.py."),
    ("trace_long", "cwt_long_4194304_x_1024", "long signal, n = 2^22 x 1024 LINFREQS (sigma 4)",
     "--workload long --steps 3 --warmup 2 --no-tune",
     "Calls include the 2 warm-up steps, the 3 timed steps and the profiled step of bench.py."),
    ("trace_batched", "cwt_batched_384_channels", "384 channels x 200 LOGSCALES (sigma 6)",
     "--workload batched --steps 3 --warmup 2",
     "Calls include the plain and batched runs of bench.py (warm-up and timed; the batch warms up with two "
     "executes, the second of which captures its HIP graph). Each batched execute makes one row-kernel launch, "
     "`k_rows_team_multi` (the team rows of the four padded sizes and then every direct row); the other row "
     "kernels below are the 384 plain calls'."),
    ("trace_adjoint", "adjoint_64_x_16384_x_128", "adjoint, 64 signals x 16,384 x 128 LOGSCALES (sigma 2)",
     "--workload adjoint --steps 3 --warmup 1",
     "Calls include the plain forward, the batched forward and the adjoint runs of bench.py (warm-up and timed)."),
]


def summarize_trace(root, d, tag, title, args, note):
    """profiles/<tag>_kernel_trace.md from rocprofv3 --kernel-trace --stats
    output in root/<d>/run_kernel_stats.csv."""
    import csv

    path = os.path.join(root, d, "run_kernel_stats.csv")
    with open(path) as f:
        rows = list(csv.DictReader(f))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{tag}_kernel_trace.md")
    with open(out, "w") as f:
        f.write(f"# Kernel times, {title} (rocprofv3 --kernel-trace --stats)\n\n")
        f.write(f"Command (on the MI355X):\n\n    {TRACE_CMD.format(d=d, a=args)}\n\n")
        f.write(f"Summarized by `python bench.py --summarize-profiles gpurun_out/prof`.\n{note}\n\n")
        f.write("| kernel | calls | average (us) | min (us) | max (us) | share (%) |\n|---|---|---|---|---|---|\n")
        for r in rows:
            f.write(f"| `{r['Name'][:80]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                    f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |\n")
    print(f"wrote {out}")


def summarize_profiles(root):
    for spec in TRACES:
        if os.path.exists(os.path.join(root, spec[0], "run_kernel_stats.csv")):
            summarize_trace(root, *spec)
    if os.path.exists(os.path.join(root, "pmc_fetch", "run_counter_collection.csv")):
        summarize_pmc(root)
    if os.path.exists(os.path.join(root, "sq_a", "run_counter_collection.csv")):
        summarize_sq(root)


# instruction-mix counter passes (one rocprofv3 --pmc run each; at most 8 SQ counters per run)
SQ_PASSES = {
    "sq_a": ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR",
             "SQ_INSTS_SMEM", "SQ_WAVES", "SQ_INSTS_BRANCH"],
    "sq_b": ["SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS", "SQ_ACTIVE_INST_SCA", "SQ_WAIT_INST_ANY",
             "SQ_WAVE_CYCLES", "SQ_BUSY_CU_CYCLES", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"],
    "sq_c": ["SQ_WAIT_ANY", "SQ_ACTIVE_INST_ANY", "SQ_WAIT_INST_LDS", "SQ_IFETCH", "SQ_INSTS_VALU_FMA_F32",
             "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_LDS_DATA_FIFO_FULL"],
}
# wave64 VALU instructions one CU issues per clock for the row kernels' mix:
# a probe of the 148-instruction 16-point register DFT (100 two-operand adds
# and subtracts, 48 multiply-adds, 46 of them in the two-operand encoding)
# measured 1.78; three-operand v_fma_f32 alone 0.98, v_add_f32 alone 1.9
VALU_RATE = 1.78
# kernels summarized, with the rows they write in the 100k x 3000 workload
SQ_KERNELS = [("k_rows_team", "team"), ("k_rows_direct<9>", "direct"), ("k_rows_direct<10>", "direct")]


def pruned_ifft_flops(n, scales, fb):
    """5 N sum_i log2 M_i: the floating-point operations of the band-limited
    inverse transforms (radix-2 count, M_i = 2^ceil(log2 band_i)), from the
    library's host planning (fcwt_hip_describe_rows; no device)."""
    import ctypes

    from fcwt_amd import _native as nat

    sc = np.ascontiguousarray(scales, np.float32)
    band = np.zeros(sc.size, np.int32)
    nat.check(nat.lib().fcwt_hip_describe_rows(int(n), nat.fptr(sc), sc.size, float(fb), 1, 0, None,
                                                band.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None),
              "describe_rows")
    N = 1 << int(np.ceil(np.log2(max(int(n), 1))))
    return 5.0 * N * float(np.ceil(np.log2(np.maximum(band, 1))).sum())


def headline_rows_per_kernel():
    """Rows of the 100k x 3000 plan per row kernel, from the library's host
    planning (fcwt_hip_describe_rows; no device): M = 2^ceil(log2 band) rows
    of M <= 512 and 1024 go to the direct kernel's instances, wider ones to
    the team kernel."""
    import ctypes

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from fcwt_amd import _native as nat
    import fcwt_amd

    fn = 3000
    sc = fcwt_amd.Scales(fcwt_amd.Morlet(2.0), fcwt_amd.FCWT_LOGSCALES, 1000, 1.0, 450.0, fn).scales
    band = np.zeros(fn, np.int32)
    nat.check(nat.lib().fcwt_hip_describe_rows(100_000, nat.fptr(sc), fn, 2.0, 1, 0, None,
                                                band.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None),
              "describe_rows")
    log2M = np.ceil(np.log2(np.maximum(band, 1))).astype(int)
    return {"k_rows_team": int((log2M > 10).sum()), "k_rows_direct<10>": int((log2M == 10).sum()),
            "k_rows_direct<9>": int((log2M <= 9).sum())}


def summarize_sq(root):
    """profiles/row_kernels_100k_x_3000_instruction_mix.{json,md} from the
    sq_a/ and sq_b/ counter runs (bench.py --steps 2 --warmup 1 --no-cpu) and
    the trace/ run's average kernel durations: per dispatch counts, and per
    point of the padded row (rows x N points) the VALU instructions of one lane."""
    import csv

    N, n = 131072, 100_000
    rows_of = headline_rows_per_kernel()
    dur = {}
    tpath = os.path.join(root, "trace", "run_kernel_stats.csv")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for r in csv.DictReader(f):
                dur[r["Name"]] = float(r["AverageNs"])
    per = {}
    for d, ctrs in SQ_PASSES.items():
        for kname, _ in SQ_KERNELS:
            for c in ctrs:
                try:
                    v, nd = _pmc_per_dispatch(os.path.join(root, d, "run_counter_collection.csv"), c, (kname,))
                except SystemExit:
                    continue
                per.setdefault(kname, {})[c] = v
    out = {"workload": WORKLOAD_TAG, "N": N, "n": n, "kernels": {}}
    cmd = ("cd /tmp && export TMPDIR=/tmp && timeout -s KILL 90 rocprofv3 --pmc {c} --output-format csv "
           "-d $GRAFT_REPO_ROOT/gpurun_out/prof/{d} -o run -- python3 $GRAFT_REPO_ROOT/bench.py --steps 2 "
           "--warmup 1 --no-cpu --no-tune --no-secondary")
    out["commands"] = [cmd.format(c=" ".join(v), d=k) for k, v in SQ_PASSES.items()]
    for kname, kind in SQ_KERNELS:
        if kname not in per:
            continue
        c = per[kname]
        rows = (rows_of or {}).get(kname)
        ent = {"per_dispatch": c, "rows": rows}
        ns = next((v for k, v in dur.items() if kname.replace("<", "<") in k), None)
        ent["avg_ns"] = ns
        if rows:
            pts = rows * N
            ent["valu_per_point"] = c.get("SQ_INSTS_VALU", 0) * 64 / pts
            ent["salu_per_point"] = c.get("SQ_INSTS_SALU", 0) * 64 / pts
            ent["lds_per_point"] = c.get("SQ_INSTS_LDS", 0) * 64 / pts
        if ns and "SQ_INSTS_VALU" in c:
            # VALU_RATE wave64 instructions per CU per clock (the measured rate of
            # the register DFTs' mix), 256 CUs x 2.4 GHz
            ent["valu_floor_ns"] = c["SQ_INSTS_VALU"] / (256 * 2.4 * VALU_RATE)
            ent["valu_issue_fraction"] = ent["valu_floor_ns"] / ns
        if c.get("SQ_LDS_IDX_ACTIVE"):
            ent["lds_bank_conflict_fraction"] = c.get("SQ_LDS_BANK_CONFLICT", 0) / c["SQ_LDS_IDX_ACTIVE"]
        if c.get("SQ_WAVE_CYCLES"):
            # shares of the waves' time: issuing, waiting on memory/LDS results
            # (s_waitcnt), waiting on anything (also barriers and s_sleep)
            wc = c["SQ_WAVE_CYCLES"]
            for k, name in (("SQ_ACTIVE_INST_ANY", "wave_issuing"), ("SQ_WAIT_INST_ANY", "wave_waiting_results"),
                            ("SQ_WAIT_ANY", "wave_waiting_any")):
                if k in c:
                    ent[name] = c[k] / wc
        out["kernels"][kname] = ent
    base = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"row_kernels_{WORKLOAD_TAG}_instruction_mix")
    with open(base + ".json", "w") as f:
        json.dump(out, f, indent=1)
    with open(base + ".md", "w") as f:
        f.write("# Instruction mix of the row kernels, cwt 100k x 3000 (rocprofv3 --pmc)\n\n")
        f.write("Commands (on the MI355X, one counter set per run):\n\n")
        for c in out["commands"]:
            f.write(f"    {c}\n")
        f.write("\nSummarized by `python bench.py --summarize-profiles gpurun_out/prof` (average kernel durations "
                "from the trace/ run of `profiles/cwt_100k_x_3000_kernel_trace.md`). Counts are per dispatch, "
                "summed over the device; per point = x 64 lanes / (rows x N), N = 131072. The VALU floor is "
                f"SQ_INSTS_VALU / (256 CUs x 2.4 GHz x {VALU_RATE}): {VALU_RATE} wave64 VALU instructions per CU "
                "per clock, the rate measured for the 16-point register DFT's mix of two-operand (VOP2) and "
                "three-operand (VOP3) encodings; VALU issue = floor / duration.\n\n")
        keys = ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_SMEM", "SQ_INSTS_VMEM_RD",
                "SQ_INSTS_VMEM_WR", "SQ_INSTS_BRANCH", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU",
                "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"]
        keys += ["SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_ACTIVE_INST_ANY", "SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_ADD_F32",
                 "SQ_INSTS_VALU_MUL_F32"]
        f.write("| kernel | rows | avg us | " + " | ".join(keys) + " | VALU/pt | SALU/pt | LDS/pt | VALU floor us "
                "| VALU issue | LDS conflict | waves issuing | waves waiting on results | waves waiting on anything |\n")
        f.write("|" + "---|" * (len(keys) + 12) + "\n")
        for kname, e in out["kernels"].items():
            c = e["per_dispatch"]
            cells = [f"{c[k] / 1e6:.2f} M" if k in c else "-" for k in keys]
            def g(k, fmt, scale=1.0):
                return fmt.format(e[k] * scale) if e.get(k) is not None else "-"

            f.write(f"| `{kname}` | {e['rows'] or '-'} | {g('avg_ns', '{:.1f}', 1e-3)} | " + " | ".join(cells)
                    + f" | {g('valu_per_point', '{:.1f}')} | {g('salu_per_point', '{:.1f}')}"
                    + f" | {g('lds_per_point', '{:.1f}')} | {g('valu_floor_ns', '{:.1f}', 1e-3)}"
                    + f" | {g('valu_issue_fraction', '{:.2f}')} | {g('lds_bank_conflict_fraction', '{:.3f}')}"
                    + f" | {g('wave_issuing', '{:.2f}')} | {g('wave_waiting_results', '{:.2f}')}"
                    + f" | {g('wave_waiting_any', '{:.2f}')} |\n")
    print(json.dumps(out)[:2000])


def summarize_pmc(root):
    """Write profiles/<kind>_100k_x_3000_traffic.{json,md} from four rocprofv3
    --pmc runs under `root`: pmc_fetch/ and pmc_write/ (bench.py --steps 2
    --warmup 1 --no-cpu) and copy_fetch/ and copy_write/ (bench.py
    --copy-probe 1024), each holding run_counter_collection.csv."""
    kind = "row_team"
    names = KIND_KERNELS[kind]
    copy_bytes = 1024 << 20
    cf, _ = _pmc_per_dispatch(os.path.join(root, "copy_fetch", "run_counter_collection.csv"), "FETCH_SIZE",
                              ("copyBuffer", "copy", "elementwise"))
    cw, _ = _pmc_per_dispatch(os.path.join(root, "copy_write", "run_counter_collection.csv"), "WRITE_SIZE",
                              ("copyBuffer", "copy", "elementwise"))
    kf, nd = _pmc_per_dispatch(os.path.join(root, "pmc_fetch", "run_counter_collection.csv"), "FETCH_SIZE", names)
    kw, _ = _pmc_per_dispatch(os.path.join(root, "pmc_write", "run_counter_collection.csv"), "WRITE_SIZE", names)
    f_fetch = copy_bytes / (cf * 1024.0)
    f_write = copy_bytes / (cw * 1024.0)
    fetch_b = kf * 1024.0 * f_fetch
    write_b = kw * 1024.0 * f_write
    cmd = ("cd /tmp && export TMPDIR=/tmp && timeout -s KILL 90 rocprofv3 --pmc {c} --output-format csv "
           "-d $GRAFT_REPO_ROOT/gpurun_out/prof/{d} -o run -- python3 $GRAFT_REPO_ROOT/bench.py {a}")
    summary = {
        "kernel": "fcwt_hip::k_rows_team", "kind": kind, "workload": WORKLOAD_TAG,
        "dispatches": nd,
        "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw,
        "copy_probe_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
        "fetch_scale": f_fetch, "write_scale": f_write,
        "fetch_bytes_per_launch": fetch_b, "write_bytes_per_launch": write_b,
        "traffic_bytes_per_launch": fetch_b + write_b,
        "commands": [cmd.format(c="FETCH_SIZE", d="pmc_fetch", a="--steps 2 --warmup 1 --no-cpu --no-tune --no-secondary"),
                     cmd.format(c="WRITE_SIZE", d="pmc_write", a="--steps 2 --warmup 1 --no-cpu --no-tune --no-secondary"),
                     cmd.format(c="FETCH_SIZE", d="copy_fetch", a="--copy-probe 1024"),
                     cmd.format(c="WRITE_SIZE", d="copy_write", a="--copy-probe 1024")],
        "summarized_by": "python bench.py --summarize-profiles gpurun_out/prof",
    }
    out_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    os.makedirs(out_dir, exist_ok=True)
    base = os.path.join(out_dir, f"{kind}_{WORKLOAD_TAG}_traffic")
    with open(base + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    with open(base + ".md", "w") as f:
        f.write(f"# HBM traffic of k_rows_team, cwt 100k x 3000 (rocprofv3 --pmc)\n\n")
        f.write("Commands (one counter per run, on the MI355X):\n\n")
        for c in summary["commands"]:
            f.write(f"    {c}\n")
        f.write(f"\nSummarized by `{summary['summarized_by']}`.\n\n")
        f.write("| quantity | value |\n|---|---|\n")
        f.write(f"| k_rows_team dispatches | {nd} |\n")
        f.write(f"| FETCH_SIZE per dispatch | {kf:.0f} KiB |\n")
        f.write(f"| WRITE_SIZE per dispatch | {kw:.0f} KiB |\n")
        f.write(f"| copy probe (1 GiB) FETCH_SIZE / WRITE_SIZE | {cf:.0f} / {cw:.0f} KiB |\n")
        f.write(f"| scale factors fetch / write | {f_fetch:.4f} / {f_write:.4f} |\n")
        f.write(f"| HBM read bytes per launch | {fetch_b / 1e9:.4f} GB |\n")
        f.write(f"| HBM write bytes per launch | {write_b / 1e9:.4f} GB |\n")
        f.write(f"| traffic per launch | {(fetch_b + write_b) / 1e9:.4f} GB |\n")
    print(json.dumps(summary))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-rows", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the tuned-schedule comparison")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the batched and adjoint workloads the headline line reports beside its metric")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    ap.add_argument("--summarize-profiles", default="", metavar="DIR")
    ap.add_argument("--workload", default="headline", choices=["headline", "long", "batched", "adjoint", "host"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU: FCWT::cwt runs on a single MI355X here "
                 "(the path shards by rows or channels across GPUs, but no multi-GPU run is made); use --gpus 1")
    if args.summarize_profiles:
        summarize_profiles(args.summarize_profiles)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload != "headline":
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        # these steps take 1-90 ms each: fewer of them, with enough untimed
        # work before them for the clocks to settle
        if "--steps" not in sys.argv:
            args.steps = 5
        if "--warmup" not in sys.argv:
            args.warmup = 3
        res = {"long": workload_long, "batched": workload_batched, "adjoint": workload_adjoint,
               "host": workload_host}[args.workload](args)
        print(json.dumps(res))
        return

    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import fcwt_amd

    n, fs, f0, f1, fn, fb = 100_000, 1000, 1.0, 450.0, 3000, 2.0
    x = headline_signal(n, fs)
    morl = fcwt_amd.Morlet(fb)
    scs = fcwt_amd.Scales(morl, fcwt_amd.FCWT_LOGSCALES, fs, f0, f1, fn)
    plan = fcwt_amd.Plan(n, scs.scales, fb, threads=1, normalize=True)
    split = plan.row_split()
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    out = torch.empty((fn, n), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        plan.execute_device(dx.data_ptr(), out.data_ptr(), stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.check()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / args.steps
    plan.check()
    # per-kernel times in a separate pass of the same steps, with HIP events
    # around every launch on the launch stream (kept out of the timed steps)
    prof_steps = max(1, min(args.steps, 10))
    plan.profile(True)
    for _ in range(prof_steps):
        step()
    prof = plan.profile_read()
    plan.profile(False)
    plan.check()

    # the same workload after create_FFT_optimization_plan's analogue: the
    # schedule of N tuned on this GPU (fcwt_hip_tune, FFTW_PATIENT effort),
    # loaded, and a new plan timed the same way (reported beside the default)
    tuned = None
    if not args.no_tune:
        chosen, tune_s = _tune(plan.N, 2)
        tplan = fcwt_amd.Plan(n, scs.scales, fb, threads=1, normalize=True)
        for _ in range(args.warmup):
            tplan.execute_device(dx.data_ptr(), out.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            tplan.execute_device(dx.data_ptr(), out.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        tplan.check()
        tms = e0.elapsed_time(e1) / args.steps
        tuned = {"ms_per_step": round(tms, 4), "vs_default": round(ms_step / tms, 4), "tune_s": round(tune_s, 2),
                 "schedule": {k: chosen[k].strip() for k in ("team_buffers", "team_lag", "direct_max", "path")
                              if k in chosen}}
        tplan.close()

    # what a plain write stream reaches on this GPU: torch's fill kernel over
    # the same 2.4 GB output (reported beside the 8 TB/s peak)
    for _ in range(3):
        out.fill_(0)
    e0.record(stream)
    for _ in range(10):
        out.fill_(0)
    e1.record(stream)
    torch.cuda.synchronize()
    fill_gbs = 10 * 8.0 * fn * n / (e0.elapsed_time(e1) * 1e-3) / 1e9

    out_bytes = 8.0 * fn * n
    value = out_bytes / (ms_step * 1e-3) / 1e9
    kernels = {}
    # every direct row of this workload has a band of 2^9 or 2^10 bins: with a
    # team launch they all run inside it
    direct_in_team = "row_writer_direct" not in prof and "row_team" in prof
    for k, (ms, c) in prof.items():
        b = kind_bytes(k, n, plan.N, split, direct_in_team)
        kernels[k] = {"ms_per_launch": ms / c, "launches_per_step": c / prof_steps,
                      "bytes_per_launch": b, "gbs": b / (ms / c * 1e-3) / 1e9}
    dom = max(prof, key=lambda k: prof[k][0])
    d = kernels[dom]
    tr = traffic_summary(dom)

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
                   "n": n, "fn": fn, "N": plan.N, "rows": split, "direct_rows_in_team_launch": direct_in_team,
                   "samples_scales_per_s": round(fn * n / (ms_step * 1e-3), 1),
                   "pruned_ifft_tflop_s": round(pruned_ifft_flops(n, scs.scales, fb) / (ms_step * 1e-3) / 1e12, 2),
                   "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2),
                   "tuned_schedule": tuned},
        "roofline": {"bound": "hbm", "achieved": round(d["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(d["gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": round(tr["traffic_bytes_per_launch"]) if tr else None,
                     "kernel": f"{dom} ({', '.join(KIND_KERNELS.get(dom, (dom,)))}): "
                               f"{d['bytes_per_launch'] / 1e9:.4f} GB algorithmic per launch, "
                               f"{d['ms_per_launch']:.4f} ms per launch",
                     "traffic_source": f"profiles/{dom}_{WORKLOAD_TAG}_traffic.json" if tr else None,
                     "fill_write_gbs": round(fill_gbs, 1),
                     "kernels": kernels},
    }
    if not args.no_secondary:
        # the metric's other half: the batched and adjoint forms against the
        # plain per-signal GPU call, on their own workloads (5 timed steps
        # after 3 warm-up steps each; `--workload batched|adjoint` prints
        # each workload's full line)
        plan.close()
        del out
        torch.cuda.empty_cache()
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = 5, 3
        bt = workload_batched(sub)
        torch.cuda.empty_cache()
        ad = workload_adjoint(sub)
        torch.cuda.empty_cache()
        result["config"]["secondary"] = {
            "batched": {"workload": "384 channels, n ~ U[20000, 200000], 200 LOGSCALES, sigma 6, 66.7 GB output",
                        "ms": bt["ms_per_step"], "gbs": bt["value"], "plain_ms": bt["config"]["plain_ms"],
                        "batched_vs_plain_speedup": bt["config"]["batched_vs_plain_speedup"],
                        "row_launches_per_execute": bt["config"]["plans"]["row_launches"]},
            "adjoint": {"workload": "64 x 16384 x 128 LOGSCALES, sigma 2: input gradient from 1.07 GB of output "
                                    "gradient", "ms": ad["ms_per_step"],
                        "plain_forward_ms": ad["config"]["plain_forward_ms"],
                        "batched_forward_ms": ad["config"]["batched_forward_ms"],
                        "adjoint_vs_plain_forward_speedup": ad["config"]["adjoint_vs_plain_forward_speedup"],
                        "batched_forward_vs_plain_speedup": ad["config"]["batched_forward_vs_plain_speedup"]}}
    if not args.no_cpu:
        rows = min(args.cpu_rows, fn)
        threads = host_threads()
        gbs, dt = cpu_baseline(x, scs.scales, fb, rows, threads)
        result["cpu_baseline"] = {"value": round(gbs, 4), "unit": "GB/s", "cores": threads, "kind": "port",
                                  "sample": f"first {rows} of {fn} scales of the same workload ({dt:.2f} s on "
                                            f"{threads} threads of {cpu_model()}; oracle float32 port: the "
                                            f"reference's window multiply in C, then a full {plan.N}-point float32 "
                                            f"inverse FFT per scale (scipy.fft), blocks of 32 scales per thread)"}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
