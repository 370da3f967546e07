// fft_reg.h -- complex DFTs of 2..32 points held in one thread's registers,
// natural order in and out, constant twiddles folded at compile time.
// Device code, internal to libfcwt_hip (host-callable too, for its tests).
//
// Radix-2 decimation in time: the input is renamed into bit-reversed order
// (no instructions), then log2 R stages of butterflies (a, b) -> (a + w b,
// a - w b).  A butterfly with a twiddle w = c + i s that is not 1 or +-i is
// written in tangent form (Linzer & Feig, "Modified FFTs for fused
// multiply-add architectures", Math. Comp. 60, 1993):
//   w b = c (b + i (s/c) b)        when |c| >= |s|,
//   w b = s ((c/s) b + i b)        otherwise,
// so that the product and the two sums take 2 + 4 fused multiply-adds
// instead of 4 multiplies and adds plus 4 adds.  A 16-point DFT is 148
// instructions this way against 168 with plain complex multiplies
// (dft_reg_dif).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.h"

namespace fcwt_hip {

// cos(2 pi k / 32), 0 <= k < 16 (float64 values rounded to float32)
__host__ __device__ constexpr float cos32(int k) {
  constexpr float c[9] = {1.0f,         0.980785251f, 0.923879504f, 0.831469595f, 0.707106769f,
                          0.555570245f, 0.382683426f, 0.195090324f, 0.0f};
  return k <= 8 ? c[k] : -c[16 - k];
}
// sin(2 pi k / 32), 0 <= k < 16
__host__ __device__ constexpr float sin32(int k) { return k <= 8 ? cos32(8 - k) : cos32(k - 8); }
// tan(2 pi k / 32) for 0 <= k <= 4
__host__ __device__ constexpr float tan32(int k) {
  constexpr float t[5] = {0.0f, 0.198912367f, 0.414213568f, 0.668178618f, 1.0f};
  return t[k];
}

template <int R>
__host__ __device__ constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// (a, b) <- (a + w b, a - w b) with w = W_R^{SIGN idx} = e^{SIGN 2 pi i idx / R},
// 0 <= idx < R/2; R, SIGN and idx are compile-time constants after unrolling.
template <int R, int SIGN>
__host__ __device__ __forceinline__ void bfly_dit(float2& a, float2& b, int idx) {
  float2 t;
  if (idx == 0) {
    t = b;
  } else if (4 * idx == R) {  // w = SIGN i
    t = make_float2(-SIGN * b.y, SIGN * b.x);
  } else {
    const int k = idx * (32 / R);  // 0 < k < 16, k != 8: w = cos32(k) + i SIGN sin32(k)
    float sc;                      // w b = sc * u
    float2 u;
    if (k <= 4 || k >= 12) {       // |cos| >= |sin|: u = b + i r b, r = SIGN tan
      const float r = SIGN * (k <= 4 ? tan32(k) : -tan32(16 - k));
      sc = cos32(k);
      u = make_float2(fmaf(-r, b.y, b.x), fmaf(r, b.x, b.y));
    } else {                       // u = r b + i b, r = cot / SIGN
      const float r = SIGN * (k < 8 ? tan32(8 - k) : -tan32(k - 8));
      sc = SIGN * sin32(k);
      u = make_float2(fmaf(r, b.x, -b.y), fmaf(r, b.y, b.x));
    }
    const float2 a0 = a;
    a = make_float2(fmaf(sc, u.x, a0.x), fmaf(sc, u.y, a0.y));
    b = make_float2(fmaf(-sc, u.x, a0.x), fmaf(-sc, u.y, a0.y));
    return;
  }
  const float2 a0 = a;
  a = make_float2(a0.x + t.x, a0.y + t.y);
  b = make_float2(a0.x - t.x, a0.y - t.y);
}

// d * W_R^{SIGN idx}, 0 <= idx < R/2, as a plain complex multiply
template <int R, int SIGN>
__host__ __device__ __forceinline__ float2 tw_const(float2 d, int idx) {
  if (idx == 0) return d;
  if (4 * idx == R) return make_float2(-SIGN * d.y, SIGN * d.x);  // * (SIGN i)
  const int k = idx * (32 / R);
  const float c = cos32(k), s = SIGN * sin32(k);
  return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
}

// The same DFT by radix-2 decimation in frequency with plain complex
// multiplies: more instructions, but the first stage pairs values that are
// ready early, which keeps fewer values live (the 32-point form of the
// direct-row kernel stays within 128 registers this way, at 4 waves per SIMD).
template <int R, int SIGN>
__host__ __device__ __forceinline__ void dft_reg_dif(float2 (&v)[R]) {
#pragma unroll
  for (int len = R; len >= 2; len >>= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int st = 0; st < R; st += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) {
        const float2 a = v[st + k], b = v[st + k + half];
        v[st + k] = make_float2(a.x + b.x, a.y + b.y);
        v[st + k + half] = tw_const<R, SIGN>(make_float2(a.x - b.x, a.y - b.y), k * (R / len));
      }
    }
  }
  float2 t[R];
#pragma unroll
  for (int i = 0; i < R; ++i) t[bitrev<R>(i)] = v[i];
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = t[i];
}

// In-place R-point DFT, out[k] = sum_j in[j] e^{SIGN 2 pi i j k / R}: the
// tangent-form decimation in time, except at R = 32 (see dft_reg_dif).
template <int R, int SIGN>
__host__ __device__ __forceinline__ void dft_reg(float2 (&v)[R]) {
  static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "R must be a power of two <= 32");
  if constexpr (R == 32) {
    dft_reg_dif<R, SIGN>(v);
    return;
  }
  float2 t[R];
#pragma unroll
  for (int i = 0; i < R; ++i) t[i] = v[bitrev<R>(i)];
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int st = 0; st < R; st += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) bfly_dit<R, SIGN>(t[st + k], t[st + k + half], k * (R / len));
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = t[i];
}

}  // namespace fcwt_hip
