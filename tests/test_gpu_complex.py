"""GPU parity for complex input: FCWT::cwt(complex<float>*, ...) / FCWT.ccwt
(src/fcwt/fcwt.cpp:486-490, 536-538, 545-548; fcwt.i:17) against the float64
oracle, through the C ABI (fcwt_hip_ccwt / fcwt_hip_ccwt_device).

Bar: relative L2 <= 1e-5 per row (TOL).  The spectrum is the c2c transform of
the zero-padded complex signal; only its positive-frequency half enters any
row, so a signal with only negative frequencies gives all-zero rows.
"""
import pathlib
import subprocess

import numpy as np
import pytest

import fcwt_amd
from fcwt_amd import _native as nat
from oracle import oracle as O
from tests.gpu_util import chirp_noise, tones_noise

pytestmark = pytest.mark.gpu

TOL = 1e-5
ROOT = pathlib.Path(__file__).resolve().parents[1]


def analytic(n, seed, fs=1000, ntones=5, fmax=400.0, sigma=0.2):
    """Complex tones e^{2 pi i f t} at seeded f in [-fmax, fmax] plus complex noise."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(-fmax, fmax, ntones)
    t = np.arange(n) / fs
    x = np.exp(2j * np.pi * f[:, None] * t[None, :]).sum(axis=0)
    x = x + sigma * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return x.astype(np.complex64)


def run_device_complex(x, sc, fb, threads=1, normalize=True, plan=None):
    import torch

    x = np.ascontiguousarray(x, np.complex64)
    sc = np.ascontiguousarray(sc, np.float32)
    p = plan or fcwt_amd.Plan(x.size, sc, fb, threads, normalize)
    dx = torch.from_numpy(x).to("cuda:0")
    out = torch.empty((sc.size, x.size), dtype=torch.complex64, device="cuda:0")
    p.execute_device(dx.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream, complex_input=True)
    torch.cuda.synchronize()
    p.check()
    return out.cpu().numpy()


def _check(x, sc, fb=2.0, threads=1, normalize=True, rows=None):
    got = run_device_complex(x, sc, fb, threads, normalize)
    sel = list(range(len(sc))) if rows is None else list(rows)
    ref = O.cwt_fp64(x, sc, fb, threads, True, normalize, rows=sel)
    err = O.rel_l2_rows(got[sel], ref)
    nz = np.linalg.norm(ref, axis=1) > 0
    assert np.all(np.isfinite(got))
    assert np.all(got[sel][~nz] == 0)
    worst = int(np.argmax(np.where(nz, err, 0))) if nz.any() else 0
    assert err[nz].max(initial=0) <= TOL, f"row {sel[worst]}: rel-L2 {err[worst]:.3e}"
    return got


def test_complex_10k_x_300():
    x = analytic(10_000, seed=1)
    _check(x, O.scales(O.LOGSCALES, 1000, 1, 101, 300))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 1000, 1025, 4097, 16384, 20_000, 50_000])
def test_complex_lengths(n):
    """Single-work-group forward transform (N <= 2^14) and the four-step one."""
    x = analytic(n, seed=n)
    _check(x, O.scales(O.LOGSCALES, 1000, 2, 450, 24))


def test_complex_headline_size_rows():
    """n = 100,000 (N = 2^17), 3000 LOGSCALES 1-450 Hz: every row kernel."""
    x = analytic(100_000, seed=2024)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 3000)
    rows = sorted(set(np.linspace(0, 2999, 30).astype(int)) | {0, 2999})
    _check(x, sc, rows=rows)


def test_negative_frequencies_give_zero_rows():
    n = N = 8192
    t = np.arange(n)
    x = np.exp(-2j * np.pi * 700 * t / N).astype(np.complex64)
    sc = O.scales(O.LOGSCALES, 1000, 5, 450, 32)
    got = run_device_complex(x, sc, 2.0)
    pos = run_device_complex(np.conj(x), sc, 2.0)
    assert np.abs(pos).max() > 1e-2
    assert np.abs(got).max() <= 1e-5 * np.abs(pos).max()


def test_real_signal_as_complex_matches_real_path():
    x = tones_noise(30_000, seed=4)
    sc = O.scales(O.LOGSCALES, 1000, 1, 450, 64)
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    from tests.gpu_util import run_device

    real = run_device(x, sc, 2.0, plan=p)
    cplx = run_device_complex(x.astype(np.complex64), sc, 2.0, plan=p)
    assert O.rel_l2_rows(cplx, real).max() <= 2 * TOL


def test_ccwt_python_api_matches_device():
    x = analytic(12_345, seed=9)
    morl = fcwt_amd.Morlet(2.0)
    scs = fcwt_amd.Scales(morl, fcwt_amd.FCWT_LOGSCALES, 1000, 1, 200, 100)
    fc = fcwt_amd.FCWT(morl, 1, False, True)
    out = np.zeros((100, x.size), np.complex64)
    fc.ccwt(x, scs, out)
    assert np.array_equal(out, run_device_complex(x, scs.scales, 2.0))
    # the real entry point on the same object rebuilds nothing and stays correct
    xr = chirp_noise(12_345, seed=3)
    fc.cwt(xr, scs, out)
    ref = O.cwt_fp64(xr, scs.scales, 2.0)
    assert O.rel_l2_rows(out, ref).max() <= TOL
    with pytest.raises(TypeError):
        fc.ccwt(x, scs, out.astype(np.complex128))


def test_cpp_api_real_and_complex(tmp_path):
    """A reference-style C++ caller runs both cwt overloads on the GPU."""
    src = tmp_path / "main.cpp"
    src.write_text(r'''
#include "fcwt.h"
#include <cmath>
#include <cstdio>
#include <vector>
int main() {
  const int n = 5000, fn = 40;
  std::vector<float> x(n);
  std::vector<std::complex<float>> xc(n);
  for (int t = 0; t < n; ++t) {
    x[t] = std::sin(2 * 3.14159265f * 50.0f * t / 1000.0f);
    xc[t] = std::complex<float>(x[t], 0.0f);
  }
  Morlet morl(2.0f);
  Scales scs(&morl, FCWT_LOGSCALES, 1000, 1.0f, 450.0f, fn);
  FCWT fcwt(&morl, 1, false, true);
  std::vector<std::complex<float>> a(n * fn), b(n * fn);
  fcwt.cwt(x.data(), n, &scs, a.data(), fn, n);
  fcwt.cwt(xc.data(), n, &scs, b.data(), fn, n);
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) { num += std::norm(a[i] - b[i]); den += std::norm(a[i]); }
  printf("%.3e %.3e\n", std::sqrt(num / den), den);
  try { fcwt.cwt(x.data(), n, &scs, a.data(), fn, n - 1); } catch (const std::invalid_argument&) { printf("caught\n"); }
  return 0;
}
''')
    exe = tmp_path / "main"
    lib = nat.LIB_PATH
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src), str(lib), f"-Wl,-rpath,{lib.parent}",
                    "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout.split()
    assert float(out[0]) <= 2 * TOL and float(out[1]) > 0 and out[2] == "caught"


def test_complex_input_team_launch_with_direct_rows():
    """Complex input at n = 30,000 (N = 2^15) with 200 scales: a plan small
    enough that its direct rows run inside the team launch; every row
    against the oracle."""
    x = analytic(30_000, seed=91)
    sc = O.scales(O.LOGSCALES, 1000, 2, 480, 200)  # every direct row's band is 2^8 .. 2^10
    p = fcwt_amd.Plan(x.size, sc, 2.0)
    split = p.row_split()
    assert split["team"] > 0 and split["direct"] > 0, split
    p.profile(True)
    run_device_complex(x, sc, 2.0, plan=p)
    assert "row_writer_direct" not in p.profile_read()
    _check(x, sc)
