"""CPU tests of the product's host side: the C-ABI library loads and exports
every symbol include/fcwt_hip.h declares; its host arithmetic equals the
oracle bit for bit; the per-scale band plan matches the oracle's O1 spectra;
the Python API validates arguments like the reference's SWIG binding; the C++
API header compiles against the library.  No GPU calls."""
import ctypes
import pathlib
import re
import subprocess

import numpy as np
import pytest

import fcwt_amd
from fcwt_amd import _native as nat
from oracle import oracle as O

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLD = np.load(ROOT / "tests" / "golden" / "ref_vectors.npz")


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "fcwt_hip.h").read_text()
    declared = set(re.findall(r"\b(fcwt_hip_\w+)\s*\(", header))
    assert len(declared) >= 15
    L = nat.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(nat.SIGNATURES), declared ^ set(nat.SIGNATURES)
    assert b"gfx950" in L.fcwt_hip_version()


def test_find2power():
    for n in [1, 2, 3, 1000, 1024, 1025, 100000, 1 << 22]:
        assert nat.lib().fcwt_hip_find2power(n) == O.find2power(n)


def test_scales_bit_exact_vs_oracle():
    k = 0
    while f"scales_{k}_args" in GOLD:
        t, fs, f0, f1, fn = GOLD[f"scales_{k}_args"]
        m = fcwt_amd.Morlet(2.0)
        s = fcwt_amd.Scales(m, int(t), int(fs), float(f0), float(f1), int(fn))
        assert np.array_equal(s.scales, O.scales(int(t), int(fs), float(f0), float(f1), int(fn)))
        fr = np.zeros(int(fn), np.float32)
        s.getFrequencies(fr)
        assert np.array_equal(fr, GOLD[f"scales_{k}_freqs"])
        k += 1


@pytest.mark.parametrize("fb,N", [(2.0, 1024), (2.0, 131072), (16.0, 2048), (0.5, 4096), (6.0, 32768)])
def test_mother_bit_exact_vs_oracle(fb, N):
    m = fcwt_amd.Morlet(fb).generate(N)
    assert np.array_equal(m, O.mother(fb, N))


def test_time_wavelet_bit_exact_vs_reference_vectors():
    k = 0
    while f"wavelet_{k}_args" in GOLD:
        fb, scale, pn = GOLD[f"wavelet_{k}_args"]
        m = fcwt_amd.Morlet(float(fb))
        assert m.getSupport(float(scale)) == int(GOLD[f"wavelet_{k}_support"])
        w = np.zeros(int(pn), np.complex64)
        m.getWavelet(float(scale), w)
        assert np.array_equal(w, GOLD[f"wavelet_{k}_values"])
        k += 1


def _describe(n, sc, fb, threads=1, flags=0):
    sc = np.ascontiguousarray(sc, np.float32)
    cov = np.zeros(sc.size, np.int32)
    band = np.zeros(sc.size, np.int32)
    seg = np.zeros(sc.size, np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    nat.check(nat.lib().fcwt_hip_describe_rows(n, nat.fptr(sc), sc.size, fb, threads, flags, ip(cov), ip(band),
                                               ip(seg)), "describe_rows")
    return cov, band, seg


@pytest.mark.parametrize("case", [
    (6000, 100, 0.01, 10.0, 300, 2.0, 8, 0),   # zero-coverage rows (reference test_lowfrequency)
    (10000, 1000, 1.0, 101.0, 300, 2.0, 1, 0),
    (20000, 1000, 1.0, 480.0, 96, 2.0, 6, 0),  # non-power-of-two threads: skipped bins
    (20000, 1000, 1.0, 480.0, 96, 2.0, 6, 1),
    (4000, 1000, 1.0, 400.0, 64, 0.6, 1, 0),   # narrow Gaussian in time: stale bins matter
])
def test_row_bands_hold_the_oracle_spectrum(case):
    """Every bin the plan leaves out of a row's band holds (in the oracle's
    reused O1) less than 1e-9 of the row's energy, and the coverage equals the
    oracle's; rows the reference leaves empty are planned empty."""
    n, fs, f0, f1, fn, fb, T, flags = case
    sc = O.scales(O.LOGSCALES, fs, f0, f1, fn)
    cov, band, seg = _describe(n, sc, fb, T, flags)
    N = 1 << O.find2power(n)
    _, _, C = O.rows_meta(N, sc, T, not flags)
    assert np.array_equal(cov, C)
    rng = np.random.default_rng(0)
    X = rng.standard_normal(N // 2) + 1j * rng.standard_normal(N // 2)
    for i, O1 in O.o1_rows(X, sc, O.mother(fb, N), N, T, not flags):
        tot = np.linalg.norm(O1)
        if tot == 0:
            assert band[i] == 0
            continue
        assert band[i] >= C[i] and band[i] <= N // 2
        assert np.linalg.norm(O1[band[i]:]) <= 1e-9 * tot, i


def test_zero_coverage_rows_keep_earlier_products():
    sc = O.scales(O.LOGSCALES, 100, 0.01, 10, 300)
    cov, band, seg = _describe(6000, sc, 2.0, 8)
    z = np.where(cov == 0)[0]
    assert z.size > 10 and np.all(band[z] > 0) and np.all(seg[z] >= 1)


# ------------------------------------------------------------ argument checks
def test_scales_argument_errors():
    m = fcwt_amd.Morlet(2.0)
    with pytest.raises(TypeError):
        fcwt_amd.Scales(m, fcwt_amd.FCWT_LOGSCALES, 1000.0, 1, 10, 5)  # fs is an int (fcwt.h:98)
    for bad in [(0, 10, 5), (1, -1, 5), (1, 10, 0)]:
        with pytest.raises(ValueError):
            fcwt_amd.Scales(m, fcwt_amd.FCWT_LOGSCALES, 1000, *bad)


def test_cwt_argument_errors_before_any_device_work():
    m = fcwt_amd.Morlet(2.0)
    s = fcwt_amd.Scales(m, fcwt_amd.FCWT_LOGSCALES, 1000, 1, 10, 5)
    f = fcwt_amd.FCWT(m, 1, False, True)
    x = np.zeros(100, np.float32)
    with pytest.raises(TypeError):
        f.cwt(x, s, np.zeros((5, 100), np.complex128))  # wrong dtype (numpy.i:374-416)
    with pytest.raises(TypeError):
        f.cwt(x, s, np.zeros(500, np.complex64))  # wrong ndim
    with pytest.raises(TypeError):
        f.cwt(x, s, np.zeros((100, 5), np.complex64).T)  # not C-contiguous
    with pytest.raises(TypeError):
        f.cwt(np.zeros((2, 50), np.float32), s, np.zeros((5, 100), np.complex64))
    with pytest.raises(ValueError):
        f.cwt(x, s, np.zeros((5, 99), np.complex64))  # wrong size: the reference overflows here
    with pytest.raises(ValueError):
        fcwt_amd.cwt(np.zeros((2, 3)), 1000, 1, 10, 5)  # boilerplate.py:8-9


def test_plan_create_rejects_bad_arguments():
    sc = np.array([1.0, 2.0], np.float32)
    h = ctypes.c_void_p()
    L = nat.lib()
    assert L.fcwt_hip_plan_create(ctypes.byref(h), 0, nat.fptr(sc), 2, 2.0, 1, 1, 0) == nat.EINVAL
    assert L.fcwt_hip_plan_create(ctypes.byref(h), 10, nat.fptr(sc), 2, 2.0, 0, 1, 0) == nat.EINVAL
    bad = np.array([1.0, -2.0], np.float32)
    assert L.fcwt_hip_plan_create(ctypes.byref(h), 10, nat.fptr(bad), 2, 2.0, 1, 1, 0) == nat.EINVAL
    assert L.fcwt_hip_plan_create(ctypes.byref(h), 10, nat.fptr(sc), 2, 2.0, 1, 1, 0x80) == nat.EINVAL
    assert b"flags" in L.fcwt_hip_last_error()
    assert L.fcwt_hip_cwt(None, None, None) == nat.EINVAL


def test_cpp_api_compiles_and_runs_host_parts(tmp_path):
    """A reference-style C++ caller (src/main.cpp) builds against include/fcwt.h."""
    src = tmp_path / "main.cpp"
    src.write_text(r'''
#include "fcwt.h"
#include <vector>
int main() {
  Morlet morl(2.0f);
  Scales scs(&morl, FCWT_LOGSCALES, 1000, 1.0f, 450.0f, 3000);
  std::vector<float> f(3000);
  scs.getFrequencies(f.data(), 3000);
  FCWT fcwt(&morl, 8, true, true);
  fcwt.create_FFT_optimization_plan(4096, "FFTW_ESTIMATE");  // defaults, no device needed
  std::vector<std::complex<float>> w(25);
  morl.getWavelet(2.0f, w.data(), 25);
  printf("%d %.6f %.6f\n", find2power(100000), f[0], f[2999]);
  return 0;
}
''')
    exe = tmp_path / "main"
    lib = nat.LIB_PATH
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src), str(lib), f"-Wl,-rpath,{lib.parent}",
                    "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, cwd=tmp_path).stdout.split()
    out = out[-3:]  # the printf line comes after the tuning messages
    assert (tmp_path / "n2048_t8.fcwt_hip").exists() and (tmp_path / "n4096_t8.fcwt_hip").exists()
    assert out[0] == "17" and abs(float(out[1]) - 450) < 1e-3 and abs(float(out[2]) - 1) < 1e-3
