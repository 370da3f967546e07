"""Convenience API of the reference (src/fcwt/boilerplate.py:5-88)."""
from __future__ import annotations

import numpy as np

from .api import FCWT, FCWT_LINFREQS, FCWT_LOGSCALES, Morlet, Scales


# FCWT objects kept between calls of ``cwt``, one per (nthreads, fast, norm):
# each holds the device plan of its last (n, scales), so repeated calls on
# signals of one shape build the plan (device tables, schedule, staging
# buffers) once.  The reference builds its FFTW plans on every call
# (fcwt.cpp:486-502), which costs little on the CPU; a device plan does not.
_FCWT_CACHE = {}


def clear_plan_cache():
    """Drops the plans ``cwt`` keeps, and their device memory."""
    for f in _FCWT_CACHE.values():
        if f._plan is not None:
            f._plan.close()
    _FCWT_CACHE.clear()


def cwt(input, fs, f0, f1, fn, nthreads=1, scaling="lin", fast=False, norm=True):
    """``fcwt.cwt`` (boilerplate.py:5-34): Morlet(2.0), LINFREQS for
    ``scaling="lin"`` else LOGSCALES; returns ``(freqs, output)`` with output
    a (fn, n) complex64 array computed on the GPU.  The device plan of the
    last call is reused while n and the scales are unchanged."""
    input = np.asarray(input)
    if input.ndim > 1:
        raise ValueError("Input must be a vector")
    if input.dtype != np.float32:
        input = input.astype(np.float32)
    key = (int(nthreads), bool(fast), bool(norm))
    fcwt = _FCWT_CACHE.get(key)
    if fcwt is None:
        fcwt = _FCWT_CACHE[key] = FCWT(Morlet(2.0), nthreads, fast, norm)
    morl = fcwt.wavelet
    scales = Scales(morl, FCWT_LINFREQS if scaling == "lin" else FCWT_LOGSCALES, fs, f0, f1, fn)
    output = np.zeros((fn, input.size), dtype=np.complex64)
    freqs = np.zeros((fn,), dtype=np.float32)
    fcwt.cwt(input, scales, output)
    scales.getFrequencies(freqs)
    return freqs, output


def _plot(input, freqs, output, fs, f0, f1, fn):
    """boilerplate.py:36-78 (matplotlib imported only when plotting)."""
    import matplotlib.pyplot as plt

    fig, (ax1, ax2) = plt.subplots(2, 1, sharex=True)
    ax1.plot(input)
    ax1.set_title("Input signal")
    ax2.imshow(np.abs(output), aspect="auto")
    ax2.set_xlabel("Time (s)")
    ax2.set_ylabel("Frequency (Hz)")
    ax2.set_title("CWT")
    ycount = 9
    yfreqs = np.linspace(freqs[0], freqs[-1], num=ycount)
    xinterval = 10
    ax2.set_xticks(ticks=np.arange(0, input.size, fs * xinterval), labels=np.arange(0, input.size / fs, xinterval))
    ax2.set_yticks(ticks=np.linspace(0, fn, num=ycount), labels=np.round(yfreqs, decimals=1))
    plt.show()


def plot(input, fs, f0=0, f1=0, fn=0, nthreads=1, scaling="lin", fast=False, norm=True):
    """boilerplate.py:81-88, with the call to ``_plot`` given its ``freqs``
    argument (the reference passes one argument too few, boilerplate.py:88)."""
    f0 = f0 if f0 else 1 / (input.size / fs)
    f1 = f1 if f1 else fs / 2
    fn = fn if fn else 100
    freqs, output = cwt(input, fs, f0, f1, fn, nthreads=nthreads, scaling=scaling, fast=fast, norm=norm)
    _plot(input, freqs, output, fs, f0, f1, fn)
