"""The register DFTs of the row kernels (fcwt_amd/csrc/fft_reg.h), compiled
for the host and checked against a float64 DFT: both transform signs, every
size 2..32, random inputs.  The tangent-form twiddles must keep float32
accuracy (relative L2 error well under 1e-6)."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HIPCC = "/opt/rocm/bin/hipcc"

SRC = r'''
#include <cmath>
#include <complex>
#include <cstdio>
#include "fft_reg.h"
using namespace fcwt_hip;
template <int R, int S> double worst() {
  double w = 0;
  unsigned seed = 12345u + R;
  for (int trial = 0; trial < 300; ++trial) {
    float2 v[R];
    std::complex<double> x[R];
    for (int i = 0; i < R; ++i) {
      seed = seed * 1103515245u + 12345u; float a = (seed >> 8) / 16777216.0f - 0.5f;
      seed = seed * 1103515245u + 12345u; float b = (seed >> 8) / 16777216.0f - 0.5f;
      v[i] = make_float2(a, b);
      x[i] = {a, b};
    }
    dft_reg<R, S>(v);
    double num = 0, den = 0;
    for (int k = 0; k < R; ++k) {
      std::complex<double> s = 0;
      for (int j = 0; j < R; ++j) s += x[j] * std::polar(1.0, S * 2 * M_PI * j * k / R);
      num += std::norm(s - std::complex<double>(v[k].x, v[k].y));
      den += std::norm(s);
    }
    w = std::fmax(w, std::sqrt(num / den));
  }
  return w;
}
int main() {
  printf("%.3e %.3e %.3e %.3e %.3e\n", worst<2, 1>(), worst<4, 1>(), worst<8, 1>(), worst<16, 1>(), worst<32, 1>());
  printf("%.3e %.3e %.3e %.3e %.3e\n", worst<2, -1>(), worst<4, -1>(), worst<8, -1>(), worst<16, -1>(),
         worst<32, -1>());
}
'''


@pytest.mark.skipif(not shutil.which(HIPCC) and not pathlib.Path(HIPCC).exists(), reason="hipcc not present")
def test_register_dfts_match_float64(tmp_path):
    src = tmp_path / "dft.cpp"
    src.write_text(SRC)
    exe = tmp_path / "dft"
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'fcwt_amd' / 'csrc'}", "--offload-arch=gfx950",
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout.split()
    errs = [float(v) for v in out]
    assert len(errs) == 10 and max(errs) < 3e-7, errs
