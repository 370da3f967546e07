// fft_rows.h -- batched complex FFTs of up to 2^14 points, one work-group per
// signal, radix-16 register butterflies between LDS passes.
// Device code, included by fcwt_hip.hip.
//
// A 2^LOG2H-point transform runs as ceil(LOG2H / 4) Stockham passes (radix
// 2^(LOG2H mod 4) first when LOG2H is not a multiple of 4, radix 16 after):
// each thread reads a group's R points into registers, multiplies by
// W_{Ns R}^{k j}, runs the R-point DFT in registers, and after a barrier
// writes them to their sorted positions.  Twiddles come from two small LDS
// tables of the root of unity (W_H^r = T2[r >> 7] T1[r & 127]), each entry
// computed in float64 and rounded once.
#pragma once
#include "fft_lds.h"
#include "fft_reg.h"

namespace fcwt_hip {

constexpr int kRowFftThreads = 1024;

// LDS index of point i: one pad word per 16 points keeps the stride-16 and
// stride-4 writes of the first passes on distinct banks.
__device__ __forceinline__ int row_pad(int i) { return i + (i >> 4); }

template <int LOG2H>
struct RowFftLds {
  static constexpr int H = 1 << LOG2H;
  static constexpr int T1N = H < 128 ? H : 128;
  static constexpr int T2N = H / T1N;
  float2 data[H + (H >> 4)];  // point i at row_pad(i)
  float2 t1[T1N];
  float2 t2[T2N];
};

// e^{SIGN 2 pi i r / H} from the LDS tables (r < H)
template <int LOG2H, int SIGN>
__device__ __forceinline__ float2 row_tw(const RowFftLds<LOG2H>& s, int r) {
  constexpr int L1 = LOG2H < 7 ? LOG2H : 7;
  float2 w = cmul(s.t2[r >> L1], s.t1[r & ((1 << L1) - 1)]);
  if (SIGN < 0) w.y = -w.y;
  return w;
}

// Fills the LDS twiddle tables: each entry e^{2 pi i r / H} computed in
// float64 and rounded once.
template <int LOG2H>
__device__ __forceinline__ void row_tw_tables(RowFftLds<LOG2H>& s) {
  using S = RowFftLds<LOG2H>;
  const double inv = ldexp(1.0, -LOG2H);
  for (int e = threadIdx.x; e < S::T1N + S::T2N; e += blockDim.x) {
    const int r = e < S::T1N ? e : (e - S::T1N) * S::T1N;
    double sn, cs;
    sincospi(2.0 * r * inv, &sn, &cs);
    const float2 w = make_float2((float)cs, (float)sn);
    if (e < S::T1N)
      s.t1[e] = w;
    else
      s.t2[e - S::T1N] = w;
  }
}

// One radix-R Stockham pass over s.data; Ns = 2^LOG2NS points already combined.
// DERIVED: the radix-16 twiddles W^k come from two table reads (W^1, W^4)
// and at most three products each, instead of one table pair per k (less
// LDS traffic, a few ulp more error).
template <int LOG2H, int LOG2R, int LOG2NS, int SIGN, bool DERIVED>
__device__ __forceinline__ void row_fft_pass(RowFftLds<LOG2H>& s) {
  constexpr int H = 1 << LOG2H, R = 1 << LOG2R, Ns = 1 << LOG2NS;
  constexpr int GROUPS = H / R;
  constexpr int PER = (GROUPS + kRowFftThreads - 1) / kRowFftThreads;  // groups per thread
  float2 v[PER][R];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int g = threadIdx.x + u * kRowFftThreads;
    if (GROUPS % kRowFftThreads == 0 || g < GROUPS) {
      const int jj = g & (Ns - 1);
#pragma unroll
      for (int k = 0; k < R; ++k) v[u][k] = s.data[row_pad(g + k * GROUPS)];
      if (LOG2NS > 0) {
        constexpr int SH = LOG2H - LOG2NS - LOG2R;
        if constexpr (R == 16 && DERIVED) {
          // W^k for k < 16 from W^1 and W^4 (tables), each at most three
          // products deep
          float2 w[16];
          w[1] = row_tw<LOG2H, SIGN>(s, jj << SH);
          w[4] = row_tw<LOG2H, SIGN>(s, (4 * jj) << SH);
          w[2] = cmul(w[1], w[1]);
          w[3] = cmul(w[2], w[1]);
          w[8] = cmul(w[4], w[4]);
          w[12] = cmul(w[8], w[4]);
#pragma unroll
          for (int b = 4; b < 16; b += 4)
#pragma unroll
            for (int k = 1; k < 4; ++k) w[b + k] = cmul(w[b], w[k]);
#pragma unroll
          for (int k = 1; k < R; ++k) v[u][k] = cmul(v[u][k], w[k]);
        } else {
#pragma unroll
          for (int k = 1; k < R; ++k) v[u][k] = cmul(v[u][k], row_tw<LOG2H, SIGN>(s, (jj * k) << SH));
        }
      }
      dft_reg<R, SIGN>(v[u]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int g = threadIdx.x + u * kRowFftThreads;
    if (GROUPS % kRowFftThreads == 0 || g < GROUPS) {
      const int jj = g & (Ns - 1);
      const int d = ((g >> LOG2NS) << (LOG2NS + LOG2R)) + jj;
#pragma unroll
      for (int k = 0; k < R; ++k) s.data[row_pad(d + k * Ns)] = v[u][k];
    }
  }
  __syncthreads();
}

// In-place 2^LOG2H-point DFT of s.data, out[k] = sum_j in[j] e^{SIGN 2 pi i jk/H}.
// Needs a __syncthreads() after filling s.data and the tables.
template <int LOG2H, int SIGN, bool DERIVED = false>
__device__ void row_fft(RowFftLds<LOG2H>& s) {
  constexpr int FIRST = LOG2H % 4;  // radix of the first pass (0: none)
  if constexpr (FIRST > 0) row_fft_pass<LOG2H, FIRST, 0, SIGN, DERIVED>(s);
  if constexpr (LOG2H >= FIRST + 4) row_fft_pass<LOG2H, 4, FIRST, SIGN, DERIVED>(s);
  if constexpr (LOG2H >= FIRST + 8) row_fft_pass<LOG2H, 4, FIRST + 4, SIGN, DERIVED>(s);
  if constexpr (LOG2H >= FIRST + 12) row_fft_pass<LOG2H, 4, FIRST + 8, SIGN, DERIVED>(s);
  static_assert(LOG2H <= 14, "one work-group holds at most 2^14 points");
}

}  // namespace fcwt_hip
