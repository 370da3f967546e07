"""CPU oracle for the fCWT ``FCWT::cwt`` hot path (Morlet wavelet, real input).

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only to check or to time a CPU baseline -- never as the product path.

Layout
------
* FP32-exact host arithmetic of the reference (scales, Morlet mother table,
  per-scale endpoint/step/coverage, window multiply with the reference's
  stale-bin reuse of ``O1``) lives in C, ``oracle/fcwt_oracle.c``, built to
  ``oracle/_build/libfcwt_oracle.so`` and loaded here with ctypes.
* The transforms are restated here with numpy:
  - ``forward_fp64``: the reference's r2c FFT of the zero-padded signal
    (src/fcwt/fcwt.cpp:486-500), computed as a float64 DFT.
  - ``cwt_fp64``: the whole path (fcwt.cpp:447-529) in float64 -- "the
    double-precision DFT of the same padded problem": float64 spectrum of the
    float32 input, the reference's float32 window values, float64 inverse DFT.
  - ``cwt_fp32_port``: the same path in float32 (numpy's float32 FFT), used
    only as the timed CPU baseline.
  - ``brute_force_cwt``: an O(n * E) direct evaluation for tiny inputs.

Pinning (tests/test_oracle.py): the C functions are checked bit-exactly against
the reference's own code (oracle/_ref probe, and the committed vectors it wrote
in tests/golden/); the float64 transforms against brute-force DFTs, the
on-bin-cosine closed form and DFT identities.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libfcwt_oracle.so"

# src/fcwt/fcwt.h:65
LINSCALES, LOGSCALES, LINFREQS = 0, 1, 2

_lib = None


def lib():
    """Load (building if needed) the oracle's C restatement."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            import subprocess

            subprocess.run(["make", "-s", "-C", str(HERE), str(LIB_PATH)], check=True)
        L = ctypes.CDLL(str(LIB_PATH))
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_find2power.argtypes = [ctypes.c_int]
        L.orc_find2power.restype = ctypes.c_int
        L.orc_scales.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, fp]
        L.orc_frequencies.argtypes = [fp, ctypes.c_int, ctypes.c_int, fp]
        L.orc_mother.argtypes = [ctypes.c_float, ctypes.c_int, fp]
        L.orc_row_meta.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ip, fp, ip]
        L.orc_rows_meta.argtypes = [ctypes.c_int, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, fp, ip]
        L.orc_window_multiply.argtypes = [fp, fp, fp, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_row_window.argtypes = [fp, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        L.orc_row_window.restype = ctypes.c_int
        L.orc_time_wavelet.argtypes = [ctypes.c_float, fp, fp, ctypes.c_int, ctypes.c_float]
        L.orc_time_wavelet.restype = ctypes.c_int
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


# ---------------------------------------------------------------- host maths
def find2power(n: int) -> int:
    """src/fcwt/fcwt.h:154-165."""
    return lib().orc_find2power(int(n))


def scales(scaletype: int, fs: int, f0: float, f1: float, fn: int) -> np.ndarray:
    """Scales::Scales, src/fcwt/fcwt.cpp:111-190 (float32)."""
    out = np.zeros(fn, np.float32)
    lib().orc_scales(int(scaletype), int(fs), float(f0), float(f1), int(fn), _fp(out))
    return out


def frequencies(sc: np.ndarray, fs: int) -> np.ndarray:
    """Scales::getFrequencies, src/fcwt/fcwt.cpp:133-137."""
    sc = np.ascontiguousarray(sc, np.float32)
    out = np.zeros(sc.size, np.float32)
    lib().orc_frequencies(_fp(sc), int(fs), sc.size, _fp(out))
    return out


def mother(fb: float, N: int) -> np.ndarray:
    """Morlet::generate(int), src/fcwt/fcwt.cpp:51-67 (float32 table of N values)."""
    out = np.zeros(N, np.float32)
    lib().orc_mother(float(fb), int(N), _fp(out))
    return out


def rows_meta(N: int, sc: np.ndarray, threads: int = 1, avx: bool = True):
    """Per-scale (E, step, C): src/fcwt/fcwt.cpp:199-214 (AVX) / 276-280 (scalar)."""
    sc = np.ascontiguousarray(sc, np.float32)
    E = np.zeros(sc.size, np.int32)
    C = np.zeros(sc.size, np.int32)
    st = np.zeros(sc.size, np.float32)
    lib().orc_rows_meta(int(N), _fp(sc), sc.size, int(threads), int(bool(avx)), _ip(E), _fp(st), _ip(C))
    return E, st, C


def row_window(mom: np.ndarray, scale: float, N: int, threads: int = 1, avx: bool = True):
    """Window values mother[idx(q)] for q < C (0 beyond), length N/2, and C."""
    mom = np.ascontiguousarray(mom, np.float32)
    w = np.zeros(N // 2, np.float32)
    C = lib().orc_row_window(_fp(mom), float(scale), int(N), int(threads), int(bool(avx)), _fp(w))
    return w, C


def window_multiply(ihat: np.ndarray, o1: np.ndarray, mom: np.ndarray, scale: float, N: int,
                    threads: int = 1, avx: bool = True) -> None:
    """FCWT::daughter_wavelet_multiplication, src/fcwt/fcwt.cpp:197-311 (in place on o1)."""
    assert ihat.dtype == np.complex64 and o1.dtype == np.complex64 and o1.flags.c_contiguous
    ihat = np.ascontiguousarray(ihat)
    lib().orc_window_multiply(_fp(ihat.view(np.float32)), _fp(o1.view(np.float32)),
                              _fp(np.ascontiguousarray(mom, np.float32)), float(scale), int(N),
                              int(threads), int(bool(avx)))


def time_wavelet(fb: float, scale: float, pn: int) -> np.ndarray:
    """Morlet::getWavelet, src/fcwt/fcwt.cpp:86-105 via generate(real, imag, ...) :68-84."""
    width = int(np.float32(fb) * np.float32(scale) * np.float32(3.0))
    m = max(width * 2 + 1, pn)
    re = np.zeros(m, np.float32)
    im = np.zeros(m, np.float32)
    lib().orc_time_wavelet(float(fb), _fp(re), _fp(im), int(pn), float(scale))
    return (re[:pn] + 1j * im[:pn]).astype(np.complex64)


# ------------------------------------------------------------ the transforms
def _signal(x):
    """float32 (real input) or complex64 (complex input, fcwt.cpp:486-490) copy."""
    return np.asarray(x, np.complex64 if np.iscomplexobj(x) else np.float32)


def forward_fp64(x: np.ndarray, N: int) -> np.ndarray:
    """X[k] = sum_{t<n} x[t] e^{-2 pi i k t / N}, k < N/2, float64: the r2c
    transform of the zero-padded real signal (fcwt.cpp:491-500) or the c2c
    transform of the zero-padded complex signal (fcwt.cpp:486-490).  Only
    k < N/2 is ever read (E_i <= N/2), so the reference's Hermitian fill of
    the upper half (fcwt.cpp:507-510) does not matter here."""
    x = _signal(x)
    if np.iscomplexobj(x):
        xp = np.zeros(N, np.complex128)
        xp[: x.size] = x.astype(np.complex128)
        return np.fft.fft(xp)[: N // 2]
    xp = np.zeros(N, np.float64)
    xp[: x.size] = x.astype(np.float64)
    return np.fft.rfft(xp)[: N // 2]


def o1_rows(X: np.ndarray, sc: np.ndarray, mom: np.ndarray, N: int, threads: int = 1, avx: bool = True):
    """Yields (i, O1_i) for every scale: the first N/2 bins of the reference's
    reused O1 buffer after scale i's window multiply (fcwt.cpp:197-311,
    449, 466-467, 514-518), in float64 (bins >= N/2 are never written: 0).

    Bins q < C_i hold X[q] * mother[idx_i(q)]; bins q >= C_i keep the
    product of the last earlier scale that covered them.  The yielded array is
    reused between iterations.
    """
    O1 = np.zeros(N // 2, np.complex128)
    for i, s in enumerate(np.asarray(sc, np.float32)):
        w, C = row_window(mom, float(s), N, threads, avx)
        O1[:C] = X[:C] * w[:C].astype(np.float64)
        yield i, O1


def cwt_fp64(x: np.ndarray, sc: np.ndarray, fb: float, threads: int = 1, avx: bool = True,
             normalize: bool = True, rows=None) -> np.ndarray:
    """Float64 restatement of FCWT::cwt (src/fcwt/fcwt.cpp:447-529), for real
    (float32) or complex (complex64, fcwt.cpp:536-538) input.

    Row i is c * IDFT_N(O1_i)[0:n] with c = 1/N when normalising (fcwt.cpp:431-445),
    where O1_i holds X[q] * mother[idx_i(q)] for q < C_i and, for q >= C_i, the
    product last written there by an earlier scale (the reference reuses one O1
    buffer across scales, fcwt.cpp:449, 466-467, 514-518; bins >= N/2 stay 0).
    Returns complex128 rows (all rows, or those listed in ``rows``).
    """
    x = _signal(x)
    n = x.size
    N = 1 << find2power(n)
    X = forward_fp64(x, N)
    mom = mother(fb, N)
    sc = np.ascontiguousarray(sc, np.float32)
    want = range(sc.size) if rows is None else sorted(set(int(r) for r in rows))
    want_set = set(want)
    last = max(want) if len(want) else -1
    out = {}
    full = np.zeros(N, np.complex128)
    for i, O1 in o1_rows(X, sc[: last + 1], mom, N, threads, avx):
        if i in want_set:
            full[: N // 2] = O1
            y = np.fft.ifft(full) * N
            if normalize:
                y = y / N
            out[i] = y[:n]
    return np.stack([out[i] for i in want]) if want else np.zeros((0, n), np.complex128)


def cwt_fp32_port(x: np.ndarray, sc: np.ndarray, fb: float, threads: int = 1, avx: bool = True,
                  normalize: bool = True, rows=None, out=None, workers: int = 1, block: int = 32) -> np.ndarray:
    """The same path in float32 (CPU baseline port): the reference's window
    multiply with its O1 reuse (C, fcwt.cpp:197-311) row after row, and the
    full N-point unnormalised inverse FFT of each row (fcwt.cpp:418-429) with
    scipy's float32 FFT, over blocks of ``block`` rows on ``workers`` host
    threads (the reference runs its scale loop on T OpenMP/FFTW threads,
    fcwt.cpp:217-220, 469-477).  The result does not depend on ``workers``."""
    import scipy.fft as sfft

    x = _signal(x)
    n = x.size
    N = 1 << find2power(n)
    xp = np.zeros(N, x.dtype)
    xp[:n] = x
    X = (np.fft.fft(xp) if np.iscomplexobj(x) else np.fft.rfft(xp))[: N // 2].astype(np.complex64)
    mom = mother(fb, N)
    sc = np.ascontiguousarray(sc, np.float32)
    nrow = sc.size if rows is None else int(rows)
    if out is None:
        out = np.zeros((nrow, n), np.complex64)
    from concurrent.futures import ThreadPoolExecutor

    def finish(buf, r0, cnt):  # numpy and scipy release the GIL here
        y = sfft.ifft(buf, axis=1, norm="forward", overwrite_x=True)  # unnormalised backward DFT
        if normalize:
            y *= np.float32(1.0 / N)
        out[r0: r0 + cnt] = y[:, :n]

    O1 = np.zeros(N // 2, np.complex64)
    block = max(1, min(block, nrow))
    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as pool:
        jobs = []
        for r0 in range(0, nrow, block):
            cnt = min(block, nrow - r0)
            buf = np.zeros((cnt, N), np.complex64)
            for k in range(cnt):  # the O1 reuse makes this part sequential
                window_multiply(X, O1, mom, float(sc[r0 + k]), N, threads, avx)
                buf[k, : N // 2] = O1
            jobs.append(pool.submit(finish, buf, r0, cnt))
            if len(jobs) > 2 * workers:
                jobs.pop(0).result()
        for j in jobs:
            j.result()
    return out


def effective_windows(sc: np.ndarray, fb: float, N: int, threads: int = 1, avx: bool = True):
    """Yields (i, w_i): the window each bin of row i is effectively multiplied
    by, w_i[q] = mother[idx_j(q)] of the last scale j <= i that covered q
    (the reference's reused O1, fcwt.cpp:449, 514-518), length N/2, float64.
    The array is reused between iterations."""
    mom = mother(fb, N)
    W = np.zeros(N // 2, np.float64)
    for i, s in enumerate(np.asarray(sc, np.float32)):
        w, C = row_window(mom, float(s), N, threads, avx)
        W[:C] = w[:C]
        yield i, W


def adjoint_fp64(G: np.ndarray, sc: np.ndarray, fb: float, threads: int = 1, avx: bool = True,
                 normalize: bool = True, complex_input: bool = False) -> np.ndarray:
    """Adjoint of ``cwt_fp64`` (no reference counterpart; the forward is
    fcwt.cpp:447-548).  G is (fn, n): G_i[t] = dL/dRe W_i[t] + i dL/dIm W_i[t].
    Returns dL/dx: xbar[t] = c Re sum_{q<N/2} A[q] e^{2 pi i q t/N} with
    A[q] = sum_i w_i(q) sum_t G_i[t] e^{-2 pi i q t/N} (complex, without the
    real part, for complex input)."""
    G = np.asarray(G, np.complex128)
    n = G.shape[1]
    N = 1 << find2power(n)
    A = np.zeros(N // 2, np.complex128)
    for i, w in effective_windows(sc, fb, N, threads, avx):
        gp = np.zeros(N, np.complex128)
        gp[:n] = G[i]
        A += w * np.fft.fft(gp)[: N // 2]
    full = np.zeros(N, np.complex128)
    full[: N // 2] = A
    y = np.fft.ifft(full)[:n] * N
    if normalize:
        y = y / N
    return y if complex_input else y.real


def brute_force_cwt(x: np.ndarray, sc: np.ndarray, fb: float, threads: int = 1, avx: bool = True,
                    normalize: bool = True) -> np.ndarray:
    """Direct O(fn * n * N/2) evaluation for tiny inputs (no FFT anywhere).

    X[k] is summed explicitly; each row is summed explicitly over the bins the
    reference's O1 holds for it, including the stale bins (see ``cwt_fp64``).
    Real or complex input.
    """
    x = _signal(x).astype(np.complex128)
    n = x.size
    N = 1 << find2power(n)
    t = np.arange(n)
    k = np.arange(N // 2)
    X = (x[None, :] * np.exp(-2j * np.pi * np.outer(k, t) / N)).sum(axis=1)
    mom = mother(fb, N)
    owner_val = np.zeros(N // 2, np.complex128)
    rows = []
    tt = np.arange(n)
    for s in np.asarray(sc, np.float32):
        w, C = row_window(mom, float(s), N, threads, avx)
        owner_val[:C] = X[:C] * w[:C]
        y = (owner_val[None, :] * np.exp(2j * np.pi * np.outer(tt, k) / N)).sum(axis=1)
        rows.append(y / N if normalize else y)
    return np.stack(rows)


def rel_l2_rows(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-row relative L2 error ||a_i - b_i|| / ||b_i||."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    num = np.linalg.norm(a - b, axis=-1)
    den = np.linalg.norm(b, axis=-1)
    return num / np.where(den == 0, 1.0, den)
