"""Writes tests/golden/ref_vectors.npz from the reference's own code.

Runs the probes built from the reference's src/fcwt/fcwt.cpp by
oracle/build_ref.sh (oracle/_ref/ref_probe_{avx,scalar}) and stores their
outputs as fixtures, so that the oracle can be checked against the reference
where the reference sources are absent.  Only the non-FFT parts of the
reference are covered (FFTW is not available): scales and frequencies
(fcwt.cpp:111-190), the Morlet mother table (fcwt.cpp:51-67), the per-scale
window multiply with its reuse of O1 (fcwt.cpp:197-311, 514-518), and the
time-domain wavelet (fcwt.cpp:68-105, fcwt.h:88).

Usage: python tests/golden/make_golden.py   (needs the probes; see oracle/Makefile)
"""
from __future__ import annotations

import hashlib
import pathlib
import subprocess
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = ROOT / "oracle" / "_ref"
OUT = HERE / "ref_vectors.npz"

# (scaletype, fs, f0, f1, fn): the workloads this repo measures or tests, the
# reference's own tests (tests/test_fcwt.py) and its benchmark (benchmark.cpp:123-128)
SCALE_CASES = [
    (1, 1000, 1.0, 101.0, 300),     # 10k x 300 chirp workload
    (1, 1000, 1.0, 450.0, 3000),    # 100k x 3000 headline workload
    (1, 30000, 10.0, 10000.0, 200), # multichannel 30 kHz workload
    (1, 1000, 10.0, 450.0, 128),    # adjoint workload
    (2, 1000, 1.0, 500.0, 1024),    # long-signal LINFREQS workload
    (2, 1000, 1.0, 20.0, 1000),     # test_odd / test_morlet / test_quick
    (1, 100, 1.0, 32.0, 30),        # test_shape
    (1, 100, 0.01, 10.0, 300),      # test_lowfrequency
    (1, 64, 1.0, 32.0, 300),        # benchmark.cpp
    (0, 1000, 2.0, 200.0, 77),      # LINSCALES
    (2, 100, 0.1, 5.0, 300),        # tutorial LINFREQS setting
    (1, 1000, 450.0, 1.0, 50),      # descending log scales (f0 > f1)
]
MOTHER_CASES = [(2.0, 1024), (2.0, 16384), (16.0, 2048), (6.0, 32768), (4.0, 4096), (1.0, 2048), (0.5, 2048),
                (2.0, 131072)]
WINDOW_N = 512
WINDOW_FB = [2.0, 0.7]
WINDOW_THREADS = [1, 8, 6, 96]
WAVELET_CASES = [(2.0, 2.0, 25), (2.0, 2.0, 5), (16.0, 3.5, 400), (1.0, 10.0, 64)]


def run(binary, args, stdin=None):
    r = subprocess.run([str(REF / binary), *map(str, args)], input=stdin, capture_output=True, check=True)
    return r.stdout


def window_scales(fn):
    # LOGSCALES from far below 2*fs/N up to 0.45*fs for fs=100: the low rows get
    # zero coverage (E < 4), as in the reference's test_lowfrequency.
    fs, f0, f1 = 100, 0.05, 45.0
    s0, s1 = fs / f1, fs / f0
    return (2.0 ** np.linspace(np.log2(s0), np.log2(s1), fn)).astype(np.float32)


def main():
    if not (REF / "ref_probe_avx").exists():
        sys.exit("reference probes not built: run `make -C oracle ref`")
    fx = {}
    for k, (t, fs, f0, f1, fn) in enumerate(SCALE_CASES):
        raw = np.frombuffer(run("ref_probe_avx", ["scales", t, fs, repr(f0), repr(f1), fn]), np.float32)
        fx[f"scales_{k}_args"] = np.array([t, fs, f0, f1, fn], np.float64)
        fx[f"scales_{k}_scales"] = raw[:fn].copy()
        fx[f"scales_{k}_freqs"] = raw[fn:].copy()
    for k, (fb, N) in enumerate(MOTHER_CASES):
        raw = run("ref_probe_avx", ["mother", repr(fb), N])
        fx[f"mother_{k}_args"] = np.array([fb, N], np.float64)
        fx[f"mother_{k}_sha256"] = np.frombuffer(hashlib.sha256(raw).digest(), np.uint8).copy()
        if N <= 32768:
            fx[f"mother_{k}_values"] = np.frombuffer(raw, np.float32).copy()
        else:
            fx[f"mother_{k}_sample"] = np.frombuffer(raw, np.float32)[:: N // 512].copy()
    rng = np.random.default_rng(20240101)
    fn = 24
    sc = window_scales(fn)
    ihat = (rng.standard_normal(WINDOW_N) + 1j * rng.standard_normal(WINDOW_N)).astype(np.complex64)
    fx["window_N"] = np.array(WINDOW_N)
    fx["window_scales"] = sc
    fx["window_ihat"] = ihat
    k = 0
    for fb in WINDOW_FB:
        for binary in ("ref_probe_avx", "ref_probe_scalar"):
            for T in WINDOW_THREADS:
                raw = run(binary, ["window", repr(fb), WINDOW_N, T, fn], stdin=sc.tobytes() + ihat.tobytes())
                snaps = np.frombuffer(raw, np.complex64).reshape(fn, WINDOW_N)
                fx[f"window_{k}_args"] = np.array([fb, 1 if binary.endswith("avx") else 0, T], np.float64)
                fx[f"window_{k}_o1"] = snaps.copy()
                k += 1
    for k, (fb, scale, pn) in enumerate(WAVELET_CASES):
        sup = np.frombuffer(run("ref_probe_avx", ["support", repr(fb), repr(scale)]), np.int32)[0]
        wav = np.frombuffer(run("ref_probe_avx", ["wavelet", repr(fb), repr(scale), pn]), np.complex64)
        fx[f"wavelet_{k}_args"] = np.array([fb, scale, pn], np.float64)
        fx[f"wavelet_{k}_support"] = np.array(sup)
        fx[f"wavelet_{k}_values"] = wav.copy()
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(fx)} arrays)")


if __name__ == "__main__":
    main()
