// fcwt_api.cpp -- the reference's C++ classes (Morlet, Scales, FCWT) on top
// of the C ABI.  Host code only.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fcwt.h"

static void check(int rc, const char* what) {
  if (rc == FCWT_HIP_OK) return;
  std::string msg = std::string(what) + ": " + fcwt_hip_last_error();
  if (rc == FCWT_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// fcwt.cpp:41-49
Morlet::Morlet(float bandwidth) {
  four_wavelen = 0.9876f;
  fb = bandwidth;
  fb2 = 2.0f * fb * fb;
  ifb = 1.0f / fb;
  imag_frequency = false;
  doublesided = false;
  mother = nullptr;
}

Morlet::~Morlet() { free(mother); }

// fcwt.cpp:51-67 (host table; the GPU plan builds its own copy on the device)
void Morlet::generate(int size) {
  width = size;
  free(mother);
  mother = (float*)malloc(sizeof(float) * (size > 0 ? size : 1));
  check(fcwt_hip_morlet_generate(fb, size, mother), "Morlet::generate");
}

// fcwt.cpp:68-84: writes 2*getSupport(scale)+1 taps of real and imag
void Morlet::generate(float* real, float* imag, int size, float scale) {
  width = getSupport(scale);
  check(fcwt_hip_morlet_generate_time(fb, scale, size, real, imag), "Morlet::generate");
}

// fcwt.cpp:86-105
void Morlet::getWavelet(float scale, std::complex<float>* pwav, int pn) {
  check(fcwt_hip_morlet_wavelet(fb, scale, reinterpret_cast<float*>(pwav), pn), "Morlet::getWavelet");
}

// fcwt.cpp:111-125
Scales::Scales(Wavelet* wav, SCALETYPE st, int afs, float af0, float af1, int afn) {
  fs = afs;
  fourwavl = wav ? wav->four_wavelen : 0.f;
  nscales = afn;
  scales = (float*)malloc(sizeof(float) * (afn > 0 ? afn : 1));
  check(fcwt_hip_scales((int)st, afs, af0, af1, afn, scales), "Scales");
}

Scales::~Scales() { free(scales); }

// fcwt.cpp:127-131
void Scales::getScales(float* pfreqs, int pnf) {
  for (int i = 0; i < pnf; ++i) pfreqs[i] = scales[i];
}

// fcwt.cpp:133-137
void Scales::getFrequencies(float* pfreqs, int pnf) { check(fcwt_hip_frequencies(scales, pnf, fs, pfreqs), "getFrequencies"); }

FCWT::~FCWT() {
  if (plan_) fcwt_hip_plan_destroy(plan_);
}

// The tuning file of padded size N and thread count T (the reference's
// wisdom file is n{N}_t{T}.wis, fcwt.cpp:340-341).
static std::string tune_file(int N, int threads) {
  return "n" + std::to_string(N) + "_t" + std::to_string(threads) + ".fcwt_hip";
}

// fcwt.cpp:313-355: for every N = 2^11 .. 2^find2power(maxsize), tune and
// write n{N}_t{T}.fcwt_hip in the working directory.  FFTW_ESTIMATE keeps
// the defaults; FFTW_MEASURE times each candidate once; FFTW_PATIENT and
// FFTW_EXHAUSTIVE take the best of three.
void FCWT::create_FFT_optimization_plan(int pmaxsize, int flags) {
  const int nt = find2power(pmaxsize);
  if (nt <= 10) std::cerr << "Maxsize is too small (<=1024)... please use a larger number" << std::endl;
  const int effort = (flags & FFTW_ESTIMATE) ? 0 : ((flags & (FFTW_PATIENT | FFTW_EXHAUSTIVE)) ? 2 : 1);
  for (int i = 11; i <= nt; ++i) {
    const int N = 1 << i;
    std::cout << "Calculating optimal scheme for N:" << N << std::endl;
    const std::string f = tune_file(N, threads);
    check(fcwt_hip_tune(N, threads, effort, f.c_str()), "FCWT::create_FFT_optimization_plan");
    std::cout << "Optimization schemes for N: " << N << " have been calculated and saved in " << f << "." << std::endl;
  }
}

// fcwt.cpp:356-374
void FCWT::create_FFT_optimization_plan(int pmaxsize, std::string flags) {
  int flag = 0;
  if (flags == "FFTW_MEASURE") {
    flag = FFTW_MEASURE;
  } else if (flags == "FFTW_PATIENT") {
    flag = FFTW_PATIENT;
  } else if (flags == "FFTW_EXHAUSTIVE") {
    flag = FFTW_EXHAUSTIVE;
  } else if (flags == "FFTW_ESTIMATE") {
    flag = FFTW_ESTIMATE;
  } else {
    std::cerr << "Unknown flag: " << flags << std::endl;
    return;
  }
  create_FFT_optimization_plan(pmaxsize, flag);
}

// fcwt.cpp:376-393: with optimisation schemes on, read the size's tuning file.
void FCWT::load_FFT_optimization_plan(int n) {
  if (!use_optimalization_schemes) return;
  const int N = 1 << find2power(n);
  if (N <= 1024) {
    std::cout << "Inputsize is too small (N <= 1024) to use optimization." << std::endl;
    return;
  }
  const std::string f = tune_file(N, threads);
  if (fcwt_hip_tune_load(f.c_str()) != FCWT_HIP_OK)
    std::cout << "WARNING: Optimization scheme '" << f << "' was not found, fallback to calculation without optimization."
              << std::endl;
}

fcwt_hip_plan FCWT::plan_for(int n, Scales* sc) {
  Morlet* morl = dynamic_cast<Morlet*>(wavelet);
  if (!morl) throw std::invalid_argument("FCWT::cwt: only the Morlet wavelet is supported");
  std::string key(reinterpret_cast<const char*>(sc->scales), sizeof(float) * (size_t)sc->nscales);
  key.append(reinterpret_cast<const char*>(&morl->fb), sizeof(float));
  if (plan_ && plan_n_ == n && plan_key_ == key) return plan_;
  if (plan_) fcwt_hip_plan_destroy(plan_);
  plan_ = nullptr;
  load_FFT_optimization_plan(n);
  check(fcwt_hip_plan_create(&plan_, n, sc->scales, sc->nscales, morl->fb, threads, use_normalization ? 1 : 0, 0),
        "FCWT::cwt");
  plan_n_ = n;
  plan_key_ = key;
  // the reference leaves wavelet->mother holding the N-point table after cwt
  // (fcwt.cpp:505)
  wavelet->generate(1 << find2power(n));
  return plan_;
}

// fcwt.cpp:532-534 (core 447-529)
void FCWT::cwt(float* pinput, int psize, std::complex<float>* poutput, Scales* scales) {
  if (!pinput || !poutput || !scales) throw std::invalid_argument("FCWT::cwt: null argument");
  fcwt_hip_plan p = plan_for(psize, scales);
  check(fcwt_hip_cwt(p, pinput, reinterpret_cast<float*>(poutput)), "FCWT::cwt");
}

// fcwt.cpp:536-538 (core 447-529, complex branch 486-490)
void FCWT::cwt(std::complex<float>* pinput, int psize, std::complex<float>* poutput, Scales* scales) {
  if (!pinput || !poutput || !scales) throw std::invalid_argument("FCWT::cwt: null argument");
  fcwt_hip_plan p = plan_for(psize, scales);
  check(fcwt_hip_ccwt(p, reinterpret_cast<const float*>(pinput), reinterpret_cast<float*>(poutput)), "FCWT::cwt");
}

static void check_output_size(int psize, const Scales* scales, int pn1, int pn2) {
  if (!scales) throw std::invalid_argument("FCWT::cwt: null argument");
  if ((long long)psize * scales->nscales != (long long)pn1 * pn2)
    throw std::invalid_argument("FCWT::cwt: output holds " + std::to_string((long long)pn1 * pn2) +
                                " values, need nscales*n = " + std::to_string((long long)psize * scales->nscales));
}

// fcwt.cpp:540-543
void FCWT::cwt(float* pinput, int psize, Scales* scales, std::complex<float>* poutput, int pn1, int pn2) {
  check_output_size(psize, scales, pn1, pn2);
  cwt(pinput, psize, poutput, scales);
}

// fcwt.cpp:545-548
void FCWT::cwt(std::complex<float>* pinput, int psize, Scales* scales, std::complex<float>* poutput, int pn1,
               int pn2) {
  check_output_size(psize, scales, pn1, pn2);
  cwt(pinput, psize, poutput, scales);
}
