// host_math.cpp -- the reference's float32 host arithmetic, bit-exact.
//
// Compiled by the host compiler with -ffp-contract=off and without fast-math,
// calling the same libm functions the reference's std:: float overloads
// resolve to (logf, powf, expf, sqrtf, cosf, sinf), so every value matches the
// reference build on the same machine bit for bit.
#include "host_math.h"

#include <algorithm>
#include <cmath>

namespace fcwt_hip {

// src/fcwt/fcwt.h:59-61
static const float kPi = 3.14159265358979323846264338327950288f;
static const float kIpi4 = 0.75112554446f;

// src/fcwt/fcwt.h:154-165
int find2power(int n) {
  int m = 0;
  long long p = 1;
  while (p < n) {
    ++m;
    p <<= 1;
  }
  return m;
}

// src/fcwt/fcwt.cpp:111-190.  Each branch keeps the reference's float32
// expression order; fs stays an int until the reference converts it.
void scales(int scaletype, int fs, float f0, float f1, int fn, float* out) {
  if (scaletype == 1) {  // FCWT_LOGSCALES, fcwt.cpp:139-158 with base 2.0f (fcwt.cpp:119)
    const float base = 2.0f;
    float s0 = fs / f1, s1 = fs / f0;
    float p0 = logf(s0) / logf(base);
    float p1 = logf(s1) / logf(base);
    float dp = p1 - p0;
    for (int i = 0; i < fn; ++i) {
      float p = p0 + (dp / (fn - 1)) * i;
      out[i] = powf(base, p);
    }
  } else if (scaletype == 0) {  // FCWT_LINSCALES, fcwt.cpp:175-190
    float s0 = fs / f1, s1 = fs / f0;
    float ds = s1 - s0;
    for (int i = 0; i < fn; ++i) out[i] = (s0 + (ds / fn) * i);
  } else {  // FCWT_LINFREQS, fcwt.cpp:160-173
    float df = f1 - f0;
    for (int i = 0; i < fn; ++i) out[fn - i - 1] = ((float)fs) / (f0 + (df / fn) * (float)i);
  }
}

// src/fcwt/fcwt.cpp:133-137
void frequencies(const float* sc, int fn, int fs, float* out) {
  for (int i = 0; i < fn; ++i) out[i] = ((float)fs) / sc[i];
}

// src/fcwt/fcwt.cpp:51-67 (frequency-domain Morlet mother table)
void morlet_mother(float fb, int size, float* mother) {
  float toradians = (2 * kPi) / (float)size;
  float norm = sqrtf(2 * kPi) * kIpi4;
  for (int w = 0; w < size; ++w) {
    float a = (2.0f * ((float)w * toradians) * fb - 2.0f * kPi * fb);
    a = -(a * a) / 2;
    mother[w] = norm * expf(a);
  }
}

// src/fcwt/fcwt.h:88
int morlet_support(float fb, float scale) { return (int)(fb * scale * 3.0f); }

// src/fcwt/fcwt.cpp:68-84: 2*width+1 time-domain taps, normalised by `size`
void morlet_time(float fb, float scale, int size, float* re, float* im) {
  const float ifb = 1.0f / fb, fb2 = 2.0f * fb * fb;
  int width = morlet_support(fb, scale);
  float norm = (float)size * ifb * kIpi4;
  for (int t = 0; t < width * 2 + 1; ++t) {
    float u = (float)(t - width) / scale;
    float g = expf(-(u * u) / (fb2));
    re[t] = norm * g * cosf(u * 2.0f * kPi) / scale;
    im[t] = norm * g * sinf(u * 2.0f * kPi) / scale;
  }
}

// src/fcwt/fcwt.cpp:86-105: first pn values of generate(real, imag, pn, scale)
// over a zeroed buffer of max(2*width+1, pn) values
void morlet_wavelet(float fb, float scale, float* out, int pn) {
  int width = morlet_support(fb, scale);
  int len = std::max(width * 2 + 1, pn);
  std::vector<float> re(std::max(len, 1), 0.0f), im(std::max(len, 1), 0.0f);
  morlet_time(fb, scale, pn, re.data(), im.data());
  for (int t = 0; t < pn; ++t) {
    out[2 * t] = re[t];
    out[2 * t + 1] = im[t];
  }
}

// src/fcwt/fcwt.cpp:199-214 (AVX) and 276-280 (scalar)
void row_meta(int N, float scale, int threads, bool avx, int* E, float* step, int* C) {
  float Nf = (float)N;
  float endf = fmin(Nf / 2.0, (Nf * 2.0) / scale);  // double maths, rounded to float
  *step = (float)(scale / 2.0);
  int end = (int)endf;
  *E = end;
  if (avx) {
    int e4 = end >> 2;
    int a = std::min(threads, std::max(1, e4 / 16));
    *C = 4 * (e4 / a) * a;
  } else {
    int a = std::min(threads, std::max(1, end / 16));
    *C = (end / a) * a;
  }
}

static inline int window_index(float step, int q, int N) {
  float v = step * (float)q;
  float mx = (float)N - 1;
  return (int)(v < mx ? v : mx);
}

std::vector<RowBand> build_bands(int N, const float* sc, int fn, int threads, bool avx, const float* mother,
                                 double rel_floor) {
  std::vector<RowBand> bands(fn);
  std::vector<int> C(fn), E(fn);
  std::vector<float> step(fn);
  for (int i = 0; i < fn; ++i) row_meta(N, sc[i], threads, avx, &E[i], &step[i], &C[i]);
  const double floor = rel_floor * (double)mother[N / 2];
  // Largest window value a scale contributes over bins [lo, hi): idx(q) is
  // non-decreasing in q and the mother table peaks at N/2.
  auto seg_max = [&](int lo, int hi, float st) {
    int wl = window_index(st, lo, N), wh = window_index(st, hi - 1, N);
    if (wl <= N / 2 && wh >= N / 2) return (double)mother[N / 2];
    return (double)std::max(mother[wl], mother[wh]);
  };
  int prefix_max = 0;
  for (int i = 0; i < fn; ++i) {
    RowBand& b = bands[i];
    b.E = E[i];
    b.C = C[i];
    if (C[i] > 0) b.segs.push_back({0, C[i], step[i]});
    // Bins [C_i, max_{j<i} C_j) still hold the product of the last earlier
    // scale that covered them (fcwt.cpp:449, 466-467, 514-518).
    int cur = C[i];
    for (int j = i - 1; j >= 0 && cur < prefix_max; --j) {
      if (C[j] > cur) {
        if (seg_max(cur, C[j], step[j]) >= floor) b.segs.push_back({cur, C[j], step[j]});
        cur = C[j];
      }
    }
    prefix_max = std::max(prefix_max, C[i]);
    b.band = 0;
    for (const Segment& s : b.segs) b.band = std::max(b.band, s.hi);
  }
  return bands;
}

}  // namespace fcwt_hip
