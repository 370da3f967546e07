// fft_reg.h -- complex DFTs of 2..32 points held in one thread's registers,
// natural order in and out, constant twiddles folded at compile time.
// Device code, internal to libfcwt_hip.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.h"

namespace fcwt_hip {

// cos and sin of 2*pi*k/32 for k = 0..8 (float64 values rounded to float32)
__device__ constexpr float kCos32[9] = {1.0f,         0.980785251f, 0.923879504f, 0.831469595f, 0.707106769f,
                                        0.555570245f, 0.382683426f, 0.195090324f, 0.0f};

__host__ __device__ constexpr float cos32(int k) {  // cos(2 pi k / 32), 0 <= k < 16
  return k <= 8 ? kCos32[k] : -kCos32[16 - k];
}
__host__ __device__ constexpr float sin32(int k) {  // sin(2 pi k / 32), 0 <= k < 16
  return k <= 8 ? kCos32[8 - k] : kCos32[k - 8];
}

// d * W_R^{SIGN*idx} with W_R = e^{2 pi i / R}, 0 <= idx < R/2, idx and R
// compile-time constants after unrolling.
template <int R, int SIGN>
__device__ __forceinline__ float2 tw_const(float2 d, int idx) {
  if (idx == 0) return d;
  if (4 * idx == R) return make_float2(-SIGN * d.y, SIGN * d.x);  // * (SIGN i)
  const int k = idx * (32 / R);
  const float c = cos32(k), s = SIGN * sin32(k);
  return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
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

// In-place R-point DFT, out[k] = sum_j in[j] e^{SIGN 2 pi i j k / R}.
// Radix-2 decimation in frequency, then the bit-reversal as a register renaming.
template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(float2 (&v)[R]) {
  static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "R must be a power of two <= 32");
#pragma unroll
  for (int len = R; len >= 2; len >>= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int st = 0; st < R; st += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) {
        float2 a = v[st + k], b = v[st + k + half];
        v[st + k] = cadd(a, b);
        v[st + k + half] = tw_const<R, SIGN>(csub(a, b), k * (R / len));
      }
    }
  }
  float2 t[R];
#pragma unroll
  for (int i = 0; i < R; ++i) t[bitrev<R>(i)] = v[i];
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = t[i];
}

}  // namespace fcwt_hip
