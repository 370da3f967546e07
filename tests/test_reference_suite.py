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


@pytest.mark.gpu
def test_complex_input():
    """tests/test_fcwt.py:67-89: ccwt of sin + i sin, n = 2,000,000.  Beyond
    the reference's shape check: the input is (1 + i) x for a real x, and only
    the positive half-spectrum enters a row, so the output is (1 + i) times
    the real path's output for x."""
    fs, n = 100, 2_000_000
    x = np.arange(n)
    real = np.sin(2 * np.pi * ((1 + (7 * x) / n) * (x / fs)), dtype="single")
    input = real + 1j * real
    assert input.dtype == np.complex64
    f0, f1, fn = 1, 32, 30
    morl = Morlet(2.0)
    out = np.zeros((fn, len(x)), dtype="csingle")
    scales = Scales(morl, FCWT_LOGSCALES, fs, f0, f1, fn)
    f = FCWT(morl, 8, True, True)
    f.ccwt(input, scales, out)
    assert out.shape[0] == fn and out.shape[1] == n
    ref = np.zeros((fn, len(x)), dtype="csingle")
    f.cwt(real, scales, ref)
    expect = (1 + 1j) * ref.astype(np.complex128)
    err = np.linalg.norm(out - expect, axis=1) / np.linalg.norm(expect, axis=1)
    assert err.max() <= 2e-5, err.max()


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


def test_plan_flags(tmp_path, monkeypatch, capsys):
    """create_FFT_optimization_plan's argument handling (host only): the
    SWIG binding takes a string flag, so the reference's own test, which
    passes 0 (tests/test_fcwt.py:242), raises TypeError; an unknown string is
    reported on stderr and nothing is written (fcwt.cpp:356-374)."""
    monkeypatch.chdir(tmp_path)
    f = FCWT(Morlet(2.0), 1, True, False)
    with pytest.raises(TypeError):
        f.create_FFT_optimization_plan(2048, 0)
    f.create_FFT_optimization_plan(2048, "FFTW_WHATEVER")
    assert "Unknown flag" in capsys.readouterr().err
    assert not os.listdir(tmp_path)


@pytest.mark.gpu
def test_plan(tmp_path, monkeypatch, capsys):
    """tests/test_fcwt.py:228-250 with the GPU tuning file in place of the
    FFTW wisdom file: n{N}_t{T}.fcwt_hip for every N up to maxsize, read back
    by cwt when optimisation schemes are on."""
    monkeypatch.chdir(tmp_path)
    fs, n = 64, 10000
    x = np.sin(2 * np.pi * 3.3 * (np.arange(n) / fs), dtype="single")
    morl = Morlet(2.0)
    # the untuned result, before any tuning enters this process's table (the
    # table is process-wide, as FFTW's wisdom is in the reference)
    scs = Scales(morl, FCWT_LOGSCALES, fs, 1, 32, 300)
    xx = np.resize(x, 30_000)
    b = np.zeros((300, 30_000), np.complex64)
    FCWT(morl, 8, False, False).cwt(xx, scs, b)
    fc = FCWT(morl, 1, True, False)
    fc.create_FFT_optimization_plan(2048, "FFTW_MEASURE")
    assert os.path.isfile("n2048_t1.fcwt_hip")
    fc = FCWT(morl, 8, True, False)
    fc.create_FFT_optimization_plan(2048, "FFTW_ESTIMATE")
    assert os.path.isfile("n2048_t8.fcwt_hip")
    # a size the XCD-team kernel serves: every candidate is timed
    fc.create_FFT_optimization_plan(1 << 15, "FFTW_MEASURE")
    text = open("n32768_t8.fcwt_hip").read()
    assert "arch gfx950" in text and "team_buffers" in text
    assert "cols_mb" in text and "path" in text
    # cwt with optimisation schemes reads the file (no warning) and agrees
    # with the untuned result (the tuned path may sum in another order)
    a = np.zeros((300, 30_000), np.complex64)
    capsys.readouterr()
    fc.cwt(xx, scs, a)
    assert "WARNING" not in capsys.readouterr().out
    err = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)
    assert err.max() <= 2e-6, err.max()
    FCWT(morl, 3, True, False).cwt(xx, scs, b)  # no file for T = 3
    assert "WARNING: Optimization scheme 'n32768_t3.fcwt_hip' was not found" in capsys.readouterr().out


@pytest.mark.gpu
def test_tuning_file_changes_the_schedule(tmp_path):
    """A tuning file that routes the wide rows of N = 2^15 to the column path
    changes the plan's schedule (row split) but not its result beyond
    summation order.  Run in a child process, so that this process's tuning
    table stays untouched."""
    import subprocess
    import sys

    code = r"""
import numpy as np, os, torch, fcwt_amd
torch.zeros(1, device="cuda:0")
from fcwt_amd import _native as nat
from oracle import oracle as O
from tests.gpu_util import run_device, tones_noise
n = 30_000
sc = O.scales(O.LOGSCALES, 1000, 1, 450, 200)
x = tones_noise(n, seed=12)
p0 = fcwt_amd.Plan(n, sc, 2.0)
s0 = p0.row_split()
ref = run_device(x, sc, 2.0, plan=p0)
path = os.path.join(os.environ["TUNE_DIR"], "n32768_t1.fcwt_hip")
open(path, "w").write(chr(10).join(["fcwt_hip tuning 1", "arch gfx950", "N 32768", "threads 1", "effort 1", "team_buffers 3", "cols_mb 64", "path cols", ""]))
assert nat.lib().fcwt_hip_tune_load(path.encode()) == 0
p1 = fcwt_amd.Plan(n, sc, 2.0)
s1 = p1.row_split()
got = run_device(x, sc, 2.0, plan=p1)
assert s0["team"] > 0 and s1["team"] == 0 and s1["two_pass"] == s0["team"], (s0, s1)
err = O.rel_l2_rows(got, ref).max()
assert err <= 2e-6, err
rows = [0, 100, 199]
assert O.rel_l2_rows(got[rows], O.cwt_fp64(x, sc, 2.0, rows=rows)).max() <= 1e-5
print("ok", s0, s1, err)
"""
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, TUNE_DIR=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
