"""Python classes with the reference's SWIG interface (src/fcwt/fcwt.i:16-82,
src/fcwt/fcwt.py:64-152): ``Morlet``, ``Scales``, ``FCWT`` and the
``FCWT_*`` scale-type constants, running ``FCWT.cwt`` on the GPU through
libfcwt_hip.

Argument conversion and errors follow the SWIG wrapper
(src/fcwt/fcwt_wrap.cxx:4962-5097, src/fcwt/numpy.i:311-416):
  * the input is converted to a contiguous float32 vector (any array-like);
    more than one dimension raises TypeError;
  * the output must be a 2-D, C-contiguous, native-endian complex64 ndarray
    (TypeError otherwise) and is written in place;
  * ``Scales`` raises ValueError for f0, f1 or fn <= 0 (fcwt.i:372-376) and
    TypeError for a non-integer ``fs`` (fs is an int in the C++ API, fcwt.h:98).
Deliberate difference: an output whose size is not nscales * n raises
ValueError (the reference only asserts, compiled out, and overflows).
"""
from __future__ import annotations

import ctypes
import numbers
import sys

import numpy as np

from . import _native as nat

# src/fcwt/fcwt.h:65
FCWT_LINSCALES = 0
FCWT_LOGSCALES = 1
FCWT_LINFREQS = 2


def _as_int(v, name):
    if isinstance(v, (bool, np.bool_)) or not isinstance(v, (numbers.Integral, np.integer)):
        raise TypeError(f"in method '{name}', expected an int, got {type(v).__name__}")
    return int(v)


def _as_float(v, name):
    if not isinstance(v, (numbers.Real, np.floating, np.integer)) or isinstance(v, bool):
        raise TypeError(f"in method '{name}', expected a float, got {type(v).__name__}")
    return float(v)


def _inplace(arr, dtype, ndim, what):
    """numpy.i INPLACE_ARRAY checks (numpy.i:374-416)."""
    if not isinstance(arr, np.ndarray):
        raise TypeError(f"{what}: array required, got {type(arr).__name__}")
    if arr.dtype != dtype:
        raise TypeError(f"{what}: array of type '{np.dtype(dtype).name}' required, got '{arr.dtype.name}'")
    if arr.ndim != ndim:
        raise TypeError(f"{what}: array must have {ndim} dimensions, given array has {arr.ndim}")
    if not arr.flags.c_contiguous:
        raise TypeError(f"{what}: array must be contiguous")
    if not arr.dtype.isnative:
        raise TypeError(f"{what}: array must have native byte order")
    if not arr.flags.writeable:
        raise TypeError(f"{what}: array must be writeable")
    return arr


def _input_vector(x, dtype, what):
    """numpy.i IN_ARRAY1 with conversion (numpy.i:311-372)."""
    a = np.asarray(x)
    if a.ndim != 1:
        raise TypeError(f"{what}: array must have 1 dimensions, given array has {a.ndim}")
    return np.ascontiguousarray(a, dtype=dtype)


class Wavelet:
    """Base of the wavelet family (src/fcwt/fcwt.h:67-79)."""

    width = 0
    four_wavelen = 0.0
    imag_frequency = False
    doublesided = False
    mother = None


class Morlet(Wavelet):
    """Morlet wavelet, src/fcwt/fcwt.h:81-94 and src/fcwt/fcwt.cpp:41-105."""

    def __init__(self, bandwidth):
        self.fb = float(np.float32(_as_float(bandwidth, "new_Morlet")))
        self.four_wavelen = float(np.float32(0.9876))
        self.imag_frequency = False
        self.doublesided = False
        self.mother = None
        self.width = 0

    def generate(self, size):
        """Frequency-domain mother table of ``size`` values (fcwt.cpp:51-67)."""
        size = _as_int(size, "Morlet_generate")
        self.width = size
        m = np.zeros(max(size, 0), np.float32)
        nat.check(nat.lib().fcwt_hip_morlet_generate(self.fb, size, nat.fptr(m)), "Morlet.generate")
        self.mother = m
        return m

    def getSupport(self, scale):
        return int(nat.lib().fcwt_hip_morlet_support(self.fb, _as_float(scale, "Morlet_getSupport")))

    def getWavelet(self, scale, pwav):
        """Fill ``pwav`` (complex64, 1-D, in place) with the time-domain wavelet (fcwt.cpp:86-105)."""
        _inplace(pwav, np.complex64, 1, "Morlet.getWavelet")
        nat.check(nat.lib().fcwt_hip_morlet_wavelet(self.fb, _as_float(scale, "Morlet_getWavelet"),
                                                    nat.fptr(pwav.view(np.float32)), pwav.size),
                  "Morlet.getWavelet")


class Scales:
    """Scale vector, src/fcwt/fcwt.h:96-112 and src/fcwt/fcwt.cpp:111-190."""

    def __init__(self, pwav, st, fs, f0, f1, fn):
        fs = _as_int(fs, "new_Scales")
        st = _as_int(st, "new_Scales")
        f0 = _as_float(f0, "new_Scales")
        f1 = _as_float(f1, "new_Scales")
        fn = _as_int(fn, "new_Scales")
        for v in (f0, f1, fn):  # fcwt.i:372-376
            if v <= 0:
                raise ValueError("Expected a non-zero, positive value.")
        self.fs = fs
        self.nscales = fn
        self.fourwavl = getattr(pwav, "four_wavelen", 0.0)
        self.scales = np.zeros(fn, np.float32)
        nat.check(nat.lib().fcwt_hip_scales(st, fs, f0, f1, fn, nat.fptr(self.scales)), "Scales")

    def getScales(self, pfreqs):
        _inplace(pfreqs, np.float32, 1, "Scales.getScales")
        k = min(pfreqs.size, self.nscales)
        pfreqs[:k] = self.scales[:k]

    def getFrequencies(self, pfreqs):
        _inplace(pfreqs, np.float32, 1, "Scales.getFrequencies")
        k = min(pfreqs.size, self.nscales)
        out = np.zeros(k, np.float32)
        nat.check(nat.lib().fcwt_hip_frequencies(nat.fptr(self.scales), k, self.fs, nat.fptr(out)), "getFrequencies")
        pfreqs[:k] = out


class Plan:
    """A GPU plan for one (n, scales, fb, threads, normalisation) combination."""

    def __init__(self, n, scales, fb, threads=1, normalize=True, flags=0):
        sc = np.ascontiguousarray(scales, np.float32)
        h = ctypes.c_void_p()
        nat.check(nat.lib().fcwt_hip_plan_create(ctypes.byref(h), int(n), nat.fptr(sc), sc.size, float(fb),
                                                 int(threads), int(bool(normalize)), int(flags)),
                  "fcwt_hip_plan_create")
        self._h = h
        self.n = int(n)
        self.fn = sc.size
        N = ctypes.c_int()
        nat.lib().fcwt_hip_plan_shape(h, None, ctypes.byref(N), None)
        self.N = N.value

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        b = ctypes.c_size_t()
        nat.check(nat.lib().fcwt_hip_plan_device_bytes(self._h, ctypes.byref(b)), "plan_device_bytes")
        return b.value

    def row_split(self) -> dict:
        """Rows per row kernel: {"direct": d, "team": t, "two_pass": w}."""
        d, t, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nat.check(nat.lib().fcwt_hip_plan_row_split(self._h, ctypes.byref(d), ctypes.byref(t), ctypes.byref(w)),
                  "plan_row_split")
        return {"direct": d.value, "team": t.value, "two_pass": w.value}

    def check(self) -> None:
        """Wait for the device; raise if an in-kernel synchronisation failed."""
        nat.check(nat.lib().fcwt_hip_plan_check(self._h), "fcwt_hip_plan_check")

    def profile(self, enable: bool = True) -> None:
        """Start (and reset) or stop per-kernel event timing."""
        nat.check(nat.lib().fcwt_hip_plan_profile(self._h, int(bool(enable))), "plan_profile")

    def profile_read(self) -> dict:
        """{kernel kind: (total ms, launches)} since the last ``profile``."""
        names = (ctypes.c_char_p * 8)()
        ms = (ctypes.c_double * 8)()
        launches = (ctypes.c_longlong * 8)()
        cnt = ctypes.c_int()
        nat.check(nat.lib().fcwt_hip_plan_profile_read(self._h, 8, names, ms, launches, ctypes.byref(cnt)),
                  "plan_profile_read")
        return {names[i].decode(): (ms[i], launches[i]) for i in range(min(cnt.value, 8))}

    def execute_host(self, x: np.ndarray, out: np.ndarray) -> None:
        """x: n float32 (real input) or n complex64 (complex input) samples,
        contiguous; out: fn*n complex64, contiguous."""
        if x.dtype == np.complex64:
            nat.check(nat.lib().fcwt_hip_ccwt(self._h, x.ctypes.data, out.ctypes.data), "fcwt_hip_ccwt")
        else:
            nat.check(nat.lib().fcwt_hip_cwt(self._h, x.ctypes.data, out.ctypes.data), "fcwt_hip_cwt")

    def execute_device(self, x_ptr: int, out_ptr: int, stream: int = 0, complex_input: bool = False) -> None:
        """Device pointers: n float32 samples (or n complex64 with
        ``complex_input``) and fn*n complex64 outputs, enqueued on ``stream``."""
        fn = nat.lib().fcwt_hip_ccwt_device if complex_input else nat.lib().fcwt_hip_cwt_device
        nat.check(fn(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream)),
                  "fcwt_hip_ccwt_device" if complex_input else "fcwt_hip_cwt_device")

    def spectrum_device(self, x_ptr: int, spec_ptr: int, stream: int = 0, complex_input: bool = False) -> None:
        """The forward transform alone: X[0, N/2) complex64 of the zero-padded input."""
        nat.check(nat.lib().fcwt_hip_spectrum_device(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(spec_ptr),
                                                     int(bool(complex_input)), ctypes.c_void_p(stream)),
                  "fcwt_hip_spectrum_device")

    def adjoint_device(self, g_ptr: int, xbar_ptr: int, stream: int = 0, complex_input: bool = False) -> None:
        """Input gradient from the output gradient (device pointers): g holds
        fn*n complex64 (dL/dRe W + i dL/dIm W), xbar receives n float32 (or n
        complex64 with ``complex_input``).  See fcwt_hip_adjoint_device."""
        nat.check(nat.lib().fcwt_hip_adjoint_device(self._h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(xbar_ptr),
                                                    int(bool(complex_input)), ctypes.c_void_p(stream)),
                  "fcwt_hip_adjoint_device")

    def adjoint_batch_device(self, g_ptr: int, xbar_ptr: int, signals: int, stream: int = 0,
                             complex_input: bool = False) -> None:
        """``adjoint_device`` for ``signals`` signals at once: g holds
        signals x fn x n complex64, xbar receives signals x n values."""
        nat.check(nat.lib().fcwt_hip_adjoint_batch_device(self._h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(xbar_ptr),
                                                          int(signals), int(bool(complex_input)),
                                                          ctypes.c_void_p(stream)),
                  "fcwt_hip_adjoint_batch_device")

    def adjoint_host(self, g: np.ndarray, complex_input: bool = False) -> np.ndarray:
        """Host-array form of ``adjoint_device``: g is (fn, n) complex64."""
        g = np.ascontiguousarray(g, np.complex64)
        if g.size != self.fn * self.n:
            raise ValueError(f"adjoint: gradient holds {g.size} values, need fn*n = {self.fn}*{self.n}")
        xbar = np.empty(self.n, np.complex64 if complex_input else np.float32)
        nat.check(nat.lib().fcwt_hip_adjoint(self._h, g.ctypes.data, xbar.ctypes.data, int(bool(complex_input))),
                  "fcwt_hip_adjoint")
        return xbar

    def close(self):
        if getattr(self, "_h", None):
            nat.lib().fcwt_hip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many signals of possibly different lengths, each with its own scales,
    transformed by one call (fcwt_hip_batch_*): signals that pad to the same
    N run through the same launches."""

    def __init__(self, lengths, scales, fb, threads=1, normalize=True, flags=0):
        lengths = [int(v) for v in lengths]
        if not lengths:
            raise ValueError("Batch: no signals")
        if isinstance(scales, np.ndarray) and scales.ndim == 1:
            scales = [scales] * len(lengths)
        if len(scales) != len(lengths):
            raise ValueError("Batch: one scale vector per signal (or one for all)")
        self._sc = [np.ascontiguousarray(s, np.float32) for s in scales]
        self.lengths = lengths
        self.fns = [s.size for s in self._sc]
        cnt = len(lengths)
        c_len = (ctypes.c_int * cnt)(*lengths)
        c_fn = (ctypes.c_int * cnt)(*self.fns)
        c_sc = (nat._c_float_p * cnt)(*[nat.fptr(s) for s in self._sc])
        h = ctypes.c_void_p()
        nat.check(nat.lib().fcwt_hip_batch_create(ctypes.byref(h), cnt, c_len, c_sc, c_fn, float(fb), int(threads),
                                                  int(bool(normalize)), int(flags)), "fcwt_hip_batch_create")
        self._h = h
        self._xs = (ctypes.c_void_p * cnt)()
        self._outs = (ctypes.c_void_p * cnt)()

    def shape(self) -> dict:
        """Plans (padded sizes), rows on the XCD-team kernel, and row-kernel
        launches per execute."""
        g, t, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nat.check(nat.lib().fcwt_hip_batch_shape(self._h, ctypes.byref(g), ctypes.byref(t)), "fcwt_hip_batch_shape")
        nat.check(nat.lib().fcwt_hip_batch_row_launches(self._h, ctypes.byref(r)), "fcwt_hip_batch_row_launches")
        return {"plans": g.value, "team_rows": t.value, "row_launches": r.value}

    def execute_device(self, x_ptrs, out_ptrs, stream: int = 0) -> None:
        """x_ptrs[c]: lengths[c] float32 on the device; out_ptrs[c]: fns[c] x
        lengths[c] complex64 on the device."""
        for c, (x, o) in enumerate(zip(x_ptrs, out_ptrs)):
            self._xs[c] = x
            self._outs[c] = o
        nat.check(nat.lib().fcwt_hip_batch_execute_device(self._h, self._xs, self._outs, ctypes.c_void_p(stream)),
                  "fcwt_hip_batch_execute_device")

    def check(self) -> None:
        nat.check(nat.lib().fcwt_hip_batch_check(self._h), "fcwt_hip_batch_check")

    def graph_launches(self) -> int:
        """Executes served by replaying a captured HIP graph so far."""
        n = ctypes.c_longlong()
        nat.check(nat.lib().fcwt_hip_batch_graph_launches(self._h, ctypes.byref(n)), "fcwt_hip_batch_graph_launches")
        return n.value

    def close(self):
        if getattr(self, "_h", None):
            nat.lib().fcwt_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FCWT:
    """src/fcwt/fcwt.h:114-152; ``cwt`` runs on the GPU (fcwt.cpp:447-548)."""

    def __init__(self, pwav, pthreads, puse_optimalization_schemes, puse_normalization):
        if not isinstance(pwav, Morlet):
            raise TypeError("FCWT: only the Morlet wavelet is supported")
        self.wavelet = pwav
        self.threads = _as_int(pthreads, "new_FCWT")
        self.use_optimalization_schemes = bool(puse_optimalization_schemes)
        self.use_normalization = bool(puse_normalization)
        self._plan = None
        self._key = None

    def _tune_file(self, N):
        return f"n{N}_t{self.threads}.fcwt_hip"

    def create_FFT_optimization_plan(self, pmaxsize, poptimizationflags):
        """fcwt.cpp:313-374: for every N = 2^11 .. 2^find2power(pmaxsize),
        tune the GPU schedule (fcwt_hip_tune) and save n{N}_t{T}.fcwt_hip in
        the working directory, where ``cwt`` finds it when optimisation
        schemes are on (fcwt.cpp:376-393)."""
        pmaxsize = _as_int(pmaxsize, "FCWT_create_FFT_optimization_plan")
        if not isinstance(poptimizationflags, str):
            raise TypeError("in method 'FCWT_create_FFT_optimization_plan', argument 3 of type 'std::string'")
        efforts = {"FFTW_ESTIMATE": 0, "FFTW_MEASURE": 1, "FFTW_PATIENT": 2, "FFTW_EXHAUSTIVE": 2}
        if poptimizationflags not in efforts:
            sys.stderr.write(f"Unknown flag: {poptimizationflags}\n")
            return
        nt = nat.lib().fcwt_hip_find2power(pmaxsize)
        if nt <= 10:
            sys.stderr.write("Maxsize is too small (<=1024)... please use a larger number\n")
        for i in range(11, nt + 1):
            N = 1 << i
            print(f"Calculating optimal scheme for N:{N}")
            nat.check(nat.lib().fcwt_hip_tune(N, self.threads, efforts[poptimizationflags],
                                              self._tune_file(N).encode()), "create_FFT_optimization_plan")

    def _load_FFT_optimization_plan(self, n):
        """fcwt.cpp:376-393."""
        if not self.use_optimalization_schemes:
            return
        N = 1 << nat.lib().fcwt_hip_find2power(n)
        if N <= 1024:
            print("Inputsize is too small (N <= 1024) to use optimization.")
            return
        f = self._tune_file(N)
        if nat.lib().fcwt_hip_tune_load(f.encode()) != nat.OK:
            print(f"WARNING: Optimization scheme '{f}' was not found, fallback to calculation without optimization.")

    def plan(self, n: int, scales: Scales) -> Plan:
        key = (int(n), scales.scales.tobytes(), self.wavelet.fb, self.threads, self.use_normalization)
        if self._plan is None or self._key != key:
            if self._plan is not None:
                self._plan.close()
            self._load_FFT_optimization_plan(n)
            self._plan = Plan(n, scales.scales, self.wavelet.fb, self.threads, self.use_normalization)
            self._key = key
        return self._plan

    def _run(self, x, scales, output, what):
        if not isinstance(scales, Scales):
            raise TypeError(f"{what}: expected a Scales object")
        _inplace(output, np.complex64, 2, what)
        if output.size != x.size * scales.nscales:
            raise ValueError(f"{what}: output holds {output.size} values, need nscales*n = "
                             f"{scales.nscales}*{x.size}")
        if x.size == 0:
            raise ValueError(f"{what}: empty input")
        self.plan(x.size, scales).execute_host(x, output)

    def cwt(self, input, scales, output):
        """FCWT::cwt(float*, int, Scales*, complex<float>*, int, int) (fcwt.cpp:540-543)."""
        self._run(_input_vector(input, np.float32, "FCWT.cwt"), scales, output, "FCWT.cwt")

    def ccwt(self, input, scales, output):
        """FCWT::cwt(complex<float>*, int, Scales*, complex<float>*, int, int)
        (fcwt.cpp:545-548; SWIG name ``ccwt``, fcwt.i:17): complex input."""
        self._run(_input_vector(input, np.complex64, "FCWT.ccwt"), scales, output, "FCWT.ccwt")
