// fcwt.h -- C++ API of fcwt_amd, source-compatible with fastlib/fCWT's
// src/fcwt/fcwt.h:65-165 (Wavelet, Morlet, Scales, FCWT, SCALETYPE,
// find2power, PI/sqrt2PI/IPI4), so code written against the reference
// compiles unchanged and runs FCWT::cwt on an MI355X (gfx950) through the C
// ABI in fcwt_hip.h.
//
// Differences from the reference, all deliberate:
//  * no FFTW or OpenMP headers; `threads` only selects the reference's
//    per-scale bin coverage (it is not a CPU thread count here);
//  * FCWT::cwt with an output of the wrong size throws std::invalid_argument
//    (the reference's check is an assert compiled out under NDEBUG,
//    fcwt.cpp:541, so it overflows the buffer instead);
//  * device failures throw std::runtime_error carrying fcwt_hip_last_error();
//  * Morlet::generate(int) frees the table it replaces and Scales frees its
//    array (the reference leaks both, fcwt.cpp:59 and fcwt.h:96-112).
#ifndef FCWT_H
#define FCWT_H

#include <complex>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>

#include "fcwt_hip.h"

#define PI 3.14159265358979323846264338327950288419716939937510582097494459072381640628620899862803482534211706798f
#define sqrt2PI 2.50662827463100050241576528f
#define IPI4 0.75112554446f

// The reference header opens namespace std for its users (fcwt.h:63), and
// code written against it uses complex<float>, vector, cout unqualified.
using namespace std;

enum SCALETYPE { FCWT_LINSCALES, FCWT_LOGSCALES, FCWT_LINFREQS };

// The planner flags create_FFT_optimization_plan(int, int) accepts, with
// FFTW's values (the reference's fcwt.h brings them in through fftw3.h).
#ifndef FFTW_MEASURE
#define FFTW_MEASURE (0U)
#define FFTW_EXHAUSTIVE (1U << 3)
#define FFTW_PATIENT (1U << 5)
#define FFTW_ESTIMATE (1U << 6)
#endif

// src/fcwt/fcwt.h:67-79
class Wavelet {
 public:
  Wavelet() {}
  virtual ~Wavelet() {}
  virtual void generate(float* real, float* imag, int size, float scale) {
    printf("ERROR [generate time complex]: Override this virtual class");
  }
  virtual void generate(int size) { printf("ERROR [generate freq]: Override this virtual class"); }
  virtual int getSupport(float scale) {
    printf("ERROR [getsupport]: Override this virtual class");
    return 0;
  }
  virtual void getWavelet(float scale, std::complex<float>* pwav, int pn) {
    printf("ERROR [getsupport]: Override this virtual class");
  }
  int width = 0;
  float four_wavelen = 0.f;
  bool imag_frequency = false, doublesided = false;
  float* mother = nullptr;
};

// src/fcwt/fcwt.h:81-94, fcwt.cpp:41-105
class Morlet : public Wavelet {
 public:
  Morlet(float bandwidth);
  ~Morlet();
  void generate(int size) override;
  void generate(float* real, float* imag, int size, float scale) override;
  int getSupport(float scale) override { return (int)(fb * scale * 3.0f); }
  void getWavelet(float scale, std::complex<float>* pwav, int pn) override;
  float fb;

 private:
  float ifb, fb2;
};

// src/fcwt/fcwt.h:96-112, fcwt.cpp:111-190
class Scales {
 public:
  Scales(Wavelet* pwav, SCALETYPE st, int fs, float f0, float f1, int fn);
  ~Scales();
  Scales(const Scales&) = delete;
  Scales& operator=(const Scales&) = delete;
  void getScales(float* pfreqs, int pnf);
  void getFrequencies(float* pfreqs, int pnf);
  float* scales;
  int fs;
  float fourwavl;
  int nscales;
};

// src/fcwt/fcwt.h:114-152
class FCWT {
 public:
  FCWT(Wavelet* pwav, int pthreads, bool puse_optimalization_schemes, bool puse_normalization)
      : wavelet(pwav),
        threads(pthreads),
        use_optimalization_schemes(puse_optimalization_schemes),
        use_normalization(puse_normalization) {}
  ~FCWT();
  FCWT(const FCWT&) = delete;
  FCWT& operator=(const FCWT&) = delete;

  // fcwt.cpp:313-374.  The reference writes FFTW wisdom files
  // n{N}_t{T}.wis; here, for every N = 2^11 .. 2^find2power(pmaxsize), the
  // GPU schedule is tuned (fcwt_hip_tune) and saved as n{N}_t{T}.fcwt_hip in
  // the working directory.  With use_optimalization_schemes, cwt loads the
  // file of its size before building a plan (fcwt.cpp:376-393); cwt builds
  // its plan on first use and reuses it while n, the scales and the options
  // are unchanged.
  void create_FFT_optimization_plan(int pmaxsize, int poptimizationflags);
  void create_FFT_optimization_plan(int pmaxsize, std::string poptimizationflags);

  // fcwt.cpp:532-548: real or complex input, output fn x n complex64 row-major
  void cwt(float* pinput, int psize, std::complex<float>* poutput, Scales* scales);
  void cwt(std::complex<float>* pinput, int psize, std::complex<float>* poutput, Scales* scales);
  void cwt(float* pinput, int psize, Scales* scales, std::complex<float>* poutput, int pn1, int pn2);
  void cwt(std::complex<float>* pinput, int psize, Scales* scales, std::complex<float>* poutput, int pn1, int pn2);

  Wavelet* wavelet;

 private:
  fcwt_hip_plan plan_for(int n, Scales* scales);
  void load_FFT_optimization_plan(int n);
  int threads;
  bool use_optimalization_schemes;
  bool use_normalization;
  fcwt_hip_plan plan_ = nullptr;
  int plan_n_ = 0;
  const Scales* plan_scales_ = nullptr;
  std::string plan_key_;
};

inline int find2power(int n) { return fcwt_hip_find2power(n); }

#endif
