// host_math.h -- the reference's float32 host arithmetic (scales, Morlet
// tables, per-scale bins) and the per-scale band description used to plan the
// GPU work.  Internal to libfcwt_hip.
#pragma once
#include <cstdint>
#include <vector>

namespace fcwt_hip {

int find2power(int n);
void scales(int scaletype, int fs, float f0, float f1, int fn, float* out);
void frequencies(const float* scales, int fn, int fs, float* out);
void morlet_mother(float fb, int size, float* mother);
int morlet_support(float fb, float scale);
void morlet_time(float fb, float scale, int size, float* real, float* imag);
void morlet_wavelet(float fb, float scale, float* interleaved, int pn);

// Bins [lo, hi) of one scale's spectrum that hold X[q] * mother[idx(q)] with
// idx(q) = (int)fminf(N-1, step*(float)q).
struct Segment {
  int lo, hi;
  float step;
};

// The spectrum one output row is the inverse DFT of, as a list of segments
// sorted by lo.  `band` is one past the highest bin held (0 for an all-zero row).
struct RowBand {
  std::vector<Segment> segs;
  int band = 0;
  int E = 0;  // endpoint, fcwt.cpp:199-202
  int C = 0;  // bins this scale itself covers
};

// Per-scale (E, step, C) exactly as the reference computes them
// (fcwt.cpp:199-214 AVX, 276-280 scalar).
void row_meta(int N, float scale, int threads, bool avx, int* E, float* step, int* C);

// Builds every row's band, reproducing the reference's reuse of one O1
// buffer across scales (bins a scale does not cover keep the product the last
// earlier scale wrote there).  Segments whose largest window value is below
// rel_floor * (window peak) are dropped: their contribution is far below
// float32 resolution.
std::vector<RowBand> build_bands(int N, const float* scales, int fn, int threads, bool avx, const float* mother,
                                 double rel_floor);

}  // namespace fcwt_hip
