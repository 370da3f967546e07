"""The reference's own behavioural tests (tests/test_fcwt.py:21-250), run
against fcwt_amd through the same Python API.  Peak-frequency checks are the
only known answers the reference's suite holds."""
import os

import numpy as np
import pytest

import fcwt_amd as fcwt
from fcwt_amd import FCWT, FCWT_LINFREQS, FCWT_LOGSCALES, Morlet, Scales


def cwt2(x, fs, f0, f1, fn):
    """tests/test_fcwt.py:8-20"""
    morl = Morlet(2.0)
    out = np.zeros((fn, len(x)), dtype="csingle")
    freqs = np.zeros((fn), dtype="single")
    scales = Scales(morl, FCWT_LOGSCALES, fs, f0, f1, fn)
    scales.getFrequencies(freqs)
    f = FCWT(morl, 8, True, True)
    f.cwt(x, scales, out)
    return out, freqs


@pytest.mark.gpu
def test_odd():
    """tests/test_fcwt.py:21-49"""
    fs, n, hz = 1000, 1000, 5
    x = np.arange(n)
    input = np.sin(2 * np.pi * hz * (x / fs), dtype="single")
    f0, f1, fn = 1, 20, 1000
    morl = Morlet(16.0)
    out = np.zeros((fn, len(x)), dtype="csingle")
    freqs = np.zeros((fn), dtype="single")
    scales = Scales(morl, FCWT_LINFREQS, fs, f0, f1, fn)
    scales.getFrequencies(freqs)
    f = FCWT(morl, 1, True, True)
    f.cwt(input, scales, out)
    outhz = freqs[np.argmax(np.mean(np.abs(out), axis=1))]
    assert np.round(hz * 2) / 2 == np.round(outhz * 2) / 2


@pytest.mark.gpu
def test_shape():
    """tests/test_fcwt.py:51-64"""
    fs, n = 100, 2_000_000
    x = np.arange(n)
    input = np.sin(2 * np.pi * ((1 + (7 * x) / n) * (x / fs)), dtype="single")
    out, _ = cwt2(input, fs, 1, 32, 30)
    assert out.shape[0] == 30 and out.shape[1] == n
    assert np.isfinite(out).all() and np.abs(out).max() > 0


def test_wavelet():
    """tests/test_fcwt.py:91-99 (host only)"""
    morl = Morlet(2.0)
    sup = morl.getSupport(2.0)
    assert sup == 12
    wavelet = np.zeros(sup * 2 + 1, dtype="csingle")
    morl.getWavelet(2.0, wavelet)
    assert np.abs(np.sum(wavelet)) != 0


@pytest.mark.gpu
def test_morlet():
    """tests/test_fcwt.py:101-129"""
    fs, n, hz = 1000, 10000, 5
    x = np.arange(n)
    input = np.sin(2 * np.pi * hz * (x / fs), dtype="single")
    f0, f1, fn = 1, 20, 1000
    morl = Morlet(16.0)
    out = np.zeros((fn, len(x)), dtype="csingle")
    freqs = np.zeros((fn), dtype="single")
    scales = Scales(morl, FCWT_LINFREQS, fs, f0, f1, fn)
    scales.getFrequencies(freqs)
    f = FCWT(morl, 8, True, True)
    f.cwt(input, scales, out)
    outhz = freqs[np.argmax(np.mean(np.abs(out), axis=1))]
    assert np.round(hz * 10) / 10 == np.round(outhz * 10) / 10


@pytest.mark.gpu
def test_quick():
    """tests/test_fcwt.py:132-153"""
    fs, n, hz = 1000, 10000, 5
    x = np.arange(n)
    input = np.sin(2 * np.pi * hz * (x / fs), dtype="single")
    freqs, out = fcwt.cwt(input, fs, 1, 20, 1000)
    outhz = freqs[np.argmax(np.mean(np.abs(out), axis=1))]
    assert np.round(hz * 10) / 10 == np.round(outhz * 10) / 10


@pytest.mark.gpu
def test_lowfrequency():
    """tests/test_fcwt.py:214-226 (seeded here; the reference draws hz unseeded)"""
    fs = 100
    n = 60 * fs
    x = np.arange(n)
    hz = 1 + np.random.default_rng(0).random(1) * 31
    input = np.sin(2 * np.pi * hz * (x / fs), dtype="single")
    out, freqs = cwt2(input, fs, 0.01, 10, 300)
    assert np.isfinite(out).all()


def test_plan(tmp_path, monkeypatch):
    """tests/test_fcwt.py:228-250.  The reference's binding takes only a
    string flag (its test passes 0 and fails); the GPU plan needs the scales,
    so the call only validates its arguments and writes no wisdom file."""
    monkeypatch.chdir(tmp_path)
    morl = Morlet(2.0)
    f = FCWT(morl, 1, True, False)
    f.create_FFT_optimization_plan(2048, "FFTW_MEASURE")
    with pytest.raises(TypeError):
        f.create_FFT_optimization_plan(2048, 0)
    assert not os.listdir(tmp_path)
