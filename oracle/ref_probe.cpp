// ref_probe.cpp -- test-only driver around the reference's own fcwt.cpp.
//
// TEST INFRASTRUCTURE ONLY (the checker of the checker).  Built by
// oracle/build_ref.sh against the reference sources where they lie; the
// binaries go to oracle/_ref/.  It calls only the parts of the reference that
// do not need FFTW (Morlet, Scales, FCWT::daughter_wavelet_multiplication);
// the linker drops the FFTW-dependent functions (-ffunction-sections
// --gc-sections), so no FFTW library is linked or needed.
//
// Commands (binary float32 on stdout, little endian):
//   scales <type> <fs> <f0> <f1> <fn>          -> fn scales, then fn frequencies
//   mother <fb> <N>                            -> N floats (Morlet::generate(N))
//   window <fb> <N> <threads> <fn>             stdin: fn scales, 2N floats Ihat
//                                              -> fn snapshots of O1 (2N floats each)
//   support <fb> <scale>                       -> one int32
//   wavelet <fb> <scale> <pn>                  -> 2*pn floats (getWavelet)
#include <cassert>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <vector>
#include <omp.h>
#include <unistd.h>
#if defined(__AVX__)
#include <immintrin.h>
#endif
#include "fftw3.h"

// daughter_wavelet_multiplication is a private member of FCWT; the probe needs
// to call it directly, so the class is seen with public access here.  Every
// header fcwt.h pulls in is included above first, so only fcwt.h's own
// classes are affected.
#define private public
#include "fcwt.h"
#undef private

static void put(const void* p, size_t bytes) { fwrite(p, 1, bytes, stdout); }

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const char* cmd = argv[1];
  if (!strcmp(cmd, "scales") && argc == 7) {
    int type = atoi(argv[2]);
    int fs = atoi(argv[3]);
    float f0 = strtof(argv[4], nullptr), f1 = strtof(argv[5], nullptr);
    int fn = atoi(argv[6]);
    Morlet morl(2.0f);
    Scales scs(&morl, (SCALETYPE)type, fs, f0, f1, fn);
    std::vector<float> s(fn), f(fn);
    scs.getScales(s.data(), fn);
    scs.getFrequencies(f.data(), fn);
    put(s.data(), 4 * fn);
    put(f.data(), 4 * fn);
    return 0;
  }
  if (!strcmp(cmd, "mother") && argc == 4) {
    float fb = strtof(argv[2], nullptr);
    int N = atoi(argv[3]);
    Morlet morl(fb);
    morl.generate(N);
    put(morl.mother, 4 * (size_t)N);
    return 0;
  }
  if (!strcmp(cmd, "window") && argc == 6) {
    float fb = strtof(argv[2], nullptr);
    int N = atoi(argv[3]);
    int threads = atoi(argv[4]);
    int fn = atoi(argv[5]);
    std::vector<float> scales(fn);
    if (fread(scales.data(), 4, fn, stdin) != (size_t)fn) return 3;
    // 32-byte aligned buffers, as the reference allocates them (fcwt.cpp:461-462)
    fftwf_complex* ihat = (fftwf_complex*)aligned_alloc(32, sizeof(fftwf_complex) * N);
    fftwf_complex* o1 = (fftwf_complex*)aligned_alloc(32, sizeof(fftwf_complex) * N);
    if (fread(ihat, 8, N, stdin) != (size_t)N) return 3;
    memset(o1, 0, sizeof(fftwf_complex) * N);  // fcwt.cpp:467
    Morlet morl(fb);
    morl.generate(N);
    FCWT fcwt(&morl, threads, false, true);
    for (int i = 0; i < fn; i++) {
      fcwt.daughter_wavelet_multiplication(ihat, o1, morl.mother, scales[i], N, morl.imag_frequency, morl.doublesided);
      put(o1, 8 * (size_t)N);
    }
    free(ihat);
    free(o1);
    return 0;
  }
  if (!strcmp(cmd, "support") && argc == 4) {
    Morlet morl(strtof(argv[2], nullptr));
    int s = morl.getSupport(strtof(argv[3], nullptr));
    put(&s, 4);
    return 0;
  }
  if (!strcmp(cmd, "wavelet") && argc == 5) {
    Morlet morl(strtof(argv[2], nullptr));
    float scale = strtof(argv[3], nullptr);
    int pn = atoi(argv[4]);
    std::vector<std::complex<float>> w(pn);
    morl.getWavelet(scale, w.data(), pn);
    put(w.data(), 8 * (size_t)pn);
    return 0;
  }
  return 2;
}
