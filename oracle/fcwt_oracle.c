/*
 * fcwt_oracle.c -- CPU restatement of the reference fCWT FP32 host arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the fcwt_amd product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product never links or calls it.
 *
 * Every function restates one piece of the reference (fastlib/fCWT) and cites
 * the file:line it follows.  Arithmetic is kept expression-for-expression in
 * IEEE float32 (compile with -ffp-contract=off, no -ffast-math) and calls the
 * same libm routines (logf, powf, expf, sqrtf) that the reference's
 * std::log/std::pow/std::exp/std::sqrt float overloads resolve to.
 *
 * Pinning: tests/test_oracle.py checks each function bit-exactly against a
 * probe built from the reference's own fcwt.cpp (oracle/build_ref.sh, output
 * in oracle/_ref/), plus closed forms and invariants.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* src/fcwt/fcwt.h:59-61 */
#define ORC_PI 3.14159265358979323846264338327950288419716939937510582097494459072381640628620899862803482534211706798f
#define ORC_IPI4 0.75112554446f

/* src/fcwt/fcwt.h:65 */
enum { ORC_LINSCALES = 0, ORC_LOGSCALES = 1, ORC_LINFREQS = 2 };

/* src/fcwt/fcwt.h:154-165: smallest m with 2^m >= n */
int orc_find2power(int n) {
  int m = 0, m2 = 1;
  while (m2 - n < 0) {
    m++;
    m2 <<= 1;
  }
  return m;
}

/* src/fcwt/fcwt.cpp:139-158 (base is 2.0f, fcwt.cpp:119) */
static void orc_logscales(float base, int fs, float f0, float f1, int fn, float* s) {
  float nf0 = f0, nf1 = f1;
  float s0 = (fs / nf1);
  float s1 = (fs / nf0);
  float power0 = logf(s0) / logf(base);
  float power1 = logf(s1) / logf(base);
  float dpower = power1 - power0;
  for (int i = 0; i < fn; i++) {
    float power = power0 + (dpower / (fn - 1)) * i;
    s[i] = powf(base, power);
  }
}

/* src/fcwt/fcwt.cpp:160-173 */
static void orc_linfreqs(int fs, float f0, float f1, int fn, float* s) {
  float nf0 = f0, nf1 = f1;
  float df = nf1 - nf0;
  for (int i = 0; i < fn; i++) s[fn - i - 1] = (((float)fs) / (nf0 + (df / fn) * (float)i));
}

/* src/fcwt/fcwt.cpp:175-190 */
static void orc_linscales(int fs, float f0, float f1, int fn, float* s) {
  float nf0 = f0, nf1 = f1;
  float s0 = fs / nf1;
  float s1 = fs / nf0;
  float ds = s1 - s0;
  for (int i = 0; i < fn; i++) s[i] = (s0 + (ds / fn) * i);
}

/* src/fcwt/fcwt.cpp:111-125 (Scales::Scales dispatch) */
void orc_scales(int type, int fs, float f0, float f1, int fn, float* s) {
  if (type == ORC_LOGSCALES)
    orc_logscales(2.0f, fs, f0, f1, fn, s);
  else if (type == ORC_LINSCALES)
    orc_linscales(fs, f0, f1, fn, s);
  else
    orc_linfreqs(fs, f0, f1, fn, s);
}

/* src/fcwt/fcwt.cpp:133-137 (Scales::getFrequencies) */
void orc_frequencies(const float* scales, int fs, int fn, float* freqs) {
  for (int i = 0; i < fn; i++) freqs[i] = ((float)fs) / scales[i];
}

/* src/fcwt/fcwt.cpp:51-67 (Morlet::generate(int), frequency-domain mother table) */
void orc_mother(float fb, int size, float* mother) {
  float toradians = (2 * ORC_PI) / (float)size;
  float norm = sqrtf(2 * ORC_PI) * ORC_IPI4;
  for (int w = 0; w < size; w++) {
    float tmp1 = (2.0f * ((float)w * toradians) * fb - 2.0f * ORC_PI * fb);
    tmp1 = -(tmp1 * tmp1) / 2;
    mother[w] = (norm * expf(tmp1));
  }
}

/*
 * src/fcwt/fcwt.cpp:197-214 and 276-280: per-scale endpoint E, index step and
 * covered bin count C.  avx != 0 follows the AVX branch (fcwt.cpp:205-215,
 * the branch every shipped build compiles, setup.py:45 / CMakeLists.txt:56);
 * avx == 0 the scalar branch (fcwt.cpp:276-280).
 */
void orc_row_meta(int isize, float scale, int threads, int avx, int* E, float* step, int* C) {
  float isizef = ((float)(isize));
  float endpointf = fmin(isizef / 2.0, ((isizef * 2.0) / scale));
  float st = (scale / 2.0);
  int endpoint = ((int)endpointf);
  int cov;
  if (avx) {
    int endpoint4 = endpoint >> 2;
    int athreads = threads < (endpoint4 / 16 > 1 ? endpoint4 / 16 : 1) ? threads : (endpoint4 / 16 > 1 ? endpoint4 / 16 : 1);
    int batchsize = (endpoint4 / athreads);
    cov = 4 * batchsize * athreads;
  } else {
    int athreads = threads < (endpoint / 16 > 1 ? endpoint / 16 : 1) ? threads : (endpoint / 16 > 1 ? endpoint / 16 : 1);
    int batchsize = (endpoint / athreads);
    cov = batchsize * athreads;
  }
  *E = endpoint;
  *step = st;
  *C = cov;
}

/* Mother-table index of bin q for a scale step (fcwt.cpp:225-229 AVX, 289-291 scalar). */
static inline int orc_index(float step, int q, int isize) {
  float maximum = (float)isize - 1;
  float v = step * (float)q;
  float t = v < maximum ? v : maximum;
  return (int)t;
}

/*
 * src/fcwt/fcwt.cpp:197-311 for the Morlet wavelet (imaginary=false,
 * doublesided=false, fcwt.cpp:46-47): O1[q] = Ihat[q] * mother[idx(q)] for
 * q < C.  Bins q >= C are left untouched, exactly as the reference leaves the
 * previous scale's product in its reused O1 buffer (fcwt.cpp:449, 466-467).
 * Ihat and O1 are interleaved complex float arrays of isize bins.
 */
void orc_window_multiply(const float* ihat, float* o1, const float* mother, float scale, int isize, int threads, int avx) {
  int E, C;
  float step;
  orc_row_meta(isize, scale, threads, avx, &E, &step, &C);
  for (int q = 0; q < C; q++) {
    float m = mother[orc_index(step, q, isize)];
    o1[2 * q] = ihat[2 * q] * m;
    o1[2 * q + 1] = ihat[2 * q + 1] * m;
  }
}

/*
 * Per-row window values w[q] = mother[idx(q)] for q < C (0 beyond), the real
 * factor the reference applies to bin q of row `scale` (fcwt.cpp:224-248).
 * Writes isize/2 values.  Returns C.
 */
int orc_row_window(const float* mother, float scale, int isize, int threads, int avx, float* w) {
  int E, C;
  float step;
  orc_row_meta(isize, scale, threads, avx, &E, &step, &C);
  int half = isize / 2;
  for (int q = 0; q < half; q++) w[q] = q < C ? mother[orc_index(step, q, isize)] : 0.0f;
  return C;
}

/* Batched row metadata for fn scales. */
void orc_rows_meta(int isize, const float* scales, int fn, int threads, int avx, int* E, float* step, int* C) {
  for (int i = 0; i < fn; i++) orc_row_meta(isize, scales[i], threads, avx, &E[i], &step[i], &C[i]);
}

/*
 * src/fcwt/fcwt.cpp:68-84 (Morlet::generate(real, imag, size, scale)),
 * time-domain Morlet wavelet over 2*width+1 taps with width = getSupport(scale)
 * = (int)(fb*scale*3.0f) (fcwt.h:88).  Returns width.
 */
int orc_time_wavelet(float fb, float* real, float* imag, int size, float scale) {
  float ifb = 1.0f / fb;
  float fb2 = 2.0f * fb * fb;
  int width = (int)(fb * scale * 3.0f);
  float norm = (float)size * ifb * ORC_IPI4;
  for (int t = 0; t < width * 2 + 1; t++) {
    float tmp1 = (float)(t - width) / scale;
    float tmp2 = expf(-(tmp1 * tmp1) / (fb2));
    real[t] = norm * tmp2 * cosf(tmp1 * 2.0f * ORC_PI) / scale;
    imag[t] = norm * tmp2 * sinf(tmp1 * 2.0f * ORC_PI) / scale;
  }
  return width;
}
