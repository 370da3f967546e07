"""Helpers shared by the GPU tests: synthetic signals and a device-pointer
runner around the C ABI (through fcwt_amd.Plan)."""
from __future__ import annotations

import numpy as np

import fcwt_amd


def chirp_noise(n, fs=1000, seed=0, sigma=0.1):
    """sin(2 pi (1 + 20 t/n) t / fs) + N(0, sigma): the 10k x 300 workload's signal."""
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64)
    x = np.sin(2 * np.pi * (1 + 20 * t / n) * t / fs) + sigma * rng.standard_normal(n)
    return x.astype(np.float32)


def tones_noise(n, fs=1000, seed=0, ntones=5, fmax=400.0, sigma=0.2):
    """Sum of `ntones` unit sinusoids at seeded frequencies in [1, fmax] Hz plus N(0, sigma)."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(1.0, fmax, ntones)
    ph = rng.uniform(0, 2 * np.pi, ntones)
    t = np.arange(n, dtype=np.float64) / fs
    x = np.sin(2 * np.pi * f[:, None] * t[None, :] + ph[:, None]).sum(axis=0)
    return (x + sigma * rng.standard_normal(n)).astype(np.float32)


def run_device(x, scales, fb, threads=1, normalize=True, flags=0, plan=None):
    """Runs the plan on cuda:0 with device pointers; returns (fn, n) complex64 numpy."""
    import torch

    x = np.ascontiguousarray(x, np.float32)
    sc = np.ascontiguousarray(scales, np.float32)
    p = plan or fcwt_amd.Plan(x.size, sc, fb, threads, normalize, flags)
    dx = torch.from_numpy(x).to("cuda:0")
    out = torch.empty((sc.size, x.size), dtype=torch.complex64, device="cuda:0")
    p.execute_device(dx.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    p.check()
    return out.cpu().numpy()
