// fft_lds.h -- batched power-of-two complex FFTs in LDS for gfx950, and the
// twiddle source they share.  Device code, internal to libfcwt_hip.
//
// Layout: a batch of G transforms of P points each lives in LDS as
// buf[e * G + g] (element e of transform g), so consecutive lanes touch
// consecutive 8-byte words and every access is bank-conflict free.
// The transform is Stockham autosort (natural order in and out), in place:
// each stage reads its butterflies into registers, waits at a barrier, writes
// them back to their sorted positions, waits again.
#pragma once
#include <hip/hip_runtime.h>

namespace fcwt_hip {

// float2 in the global address space: pointer parameters of non-inlined
// device functions keep global_* instructions instead of flat_* ones.
typedef __attribute__((address_space(1))) float2 gfloat2;
typedef __attribute__((address_space(1))) unsigned long long gu64;

// *p = v through a global-address-space pointer (one 8-byte store)
__device__ __forceinline__ void gstore(gfloat2* p, float2 v) {
  *reinterpret_cast<gu64*>(p) = ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
}
// *p through a global-address-space pointer (one 8-byte load)
__device__ __forceinline__ float2 gload(const gfloat2* p) {
  unsigned long long u = *reinterpret_cast<const gu64*>(p);
  return make_float2(__uint_as_float((unsigned)u), __uint_as_float((unsigned)(u >> 32)));
}
// p[0] = a, p[1] = b in one 16-byte store (p 16-byte aligned)
__device__ __forceinline__ void gstore2(gfloat2* p, float2 a, float2 b) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(1))) u32x4 gu32x4;
  const u32x4 u = {__float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(b.x), __float_as_uint(b.y)};
  *reinterpret_cast<gu32x4*>(p) = u;
}
// the same with the non-temporal hint, for output that is written once and not re-read
__device__ __forceinline__ void gstore_nt(gfloat2* p, float2 v) {
  __builtin_nontemporal_store(((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x),
                              reinterpret_cast<gu64*>(p));
}

// Stores of output rows through a buffer resource: the hardware range check
// drops a store at an element offset >= the row length, so a partial run of
// a row is written without a branch or an exec mask per store.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// the resource for `count` complex64 values at p (count * 8 < 2^32)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(float2* p, long long count) {
  return __builtin_amdgcn_make_buffer_rsrc(p, 0, (int)(unsigned)(count * 8), 0x00020000);
}
// row[idx] = v with the non-temporal hint (written once, never re-read)
__device__ __forceinline__ void bstore_nt(__amdgpu_buffer_rsrc_t r, unsigned idx, float2 v) {
  const u32x2 u = {__float_as_uint(v.x), __float_as_uint(v.y)};
  __builtin_amdgcn_raw_buffer_store_b64(u, r, idx * 8u, 0, 2);
}
// row[idx] = v without a cache hint: the L2 keeps the line, so that the
// other half of a 128-byte line, written by another work-group shortly
// after, joins it before the line goes to memory
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned idx, float2 v) {
  const u32x2 u = {__float_as_uint(v.x), __float_as_uint(v.y)};
  __builtin_amdgcn_raw_buffer_store_b64(u, r, idx * 8u, 0, 0);
}
// row[idx] = a, row[idx + 1] = b in one 16-byte store (idx even, the row
// 16-byte aligned; both in range or both out: the caller keeps pairs from
// straddling the end)
__device__ __forceinline__ void bstore2_nt(__amdgpu_buffer_rsrc_t r, unsigned idx, float2 a, float2 b) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 u = {__float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(b.x), __float_as_uint(b.y)};
  __builtin_amdgcn_raw_buffer_store_b128(u, r, idx * 8u, 0, 2);
}

// The value that lane l ^ 16 holds in x (lanes 16 apart, within each
// 32-lane half of the wave), from v_permlane16_swap: that instruction swaps
// the odd 16-lane rows of its first operand with the even rows of its
// second, so with x in both, the first result holds the partner's value in
// the odd rows and the second in the even rows.
__device__ __forceinline__ float lane_xor16(float x, bool odd_row) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(odd_row ? r[0] : r[1]);
}

constexpr int kLdsCap = 16384;  // complex64 elements per work-group (128 KiB)
constexpr int kBlock = 1024;    // threads per LDS-FFT work-group

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by s*i (s = +1 or -1)
__device__ __forceinline__ float2 cmul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// e^{+2 pi i r / N} for 0 <= r < N from two float64-generated tables:
// t1[j] = e^{2 pi i j / N} (j < 2^s1bits), t2[k] = e^{2 pi i k 2^s1bits / N}.
struct Twiddles {
  const float2* t1;
  const float2* t2;
  int s1bits;
  int log2N;
  __device__ __forceinline__ float2 operator()(unsigned r) const {
    float2 a = t2[r >> s1bits];
    float2 b = t1[r & ((1u << s1bits) - 1u)];
    return cmul(a, b);
  }
  // e^{sign * 2 pi i r / N}
  __device__ __forceinline__ float2 get(unsigned r, float sign) const {
    float2 w = (*this)(r);
    w.y *= sign;
    return w;
  }
};

// One Stockham stage of radix R over G interleaved transforms of 2^log2P
// points; Ns = 2^log2Ns is the size of the sub-transforms already done.
template <int R>
__device__ __forceinline__ void stockham_stage(float2* buf, int log2P, int log2Ns, int G, int log2G,
                                               const Twiddles& tw, float sign) {
  constexpr int log2R = R == 2 ? 1 : (R == 4 ? 2 : 3);
  const int P = 1 << log2P;
  const int nbf = (P >> log2R) << log2G;  // butterflies in the stage
  constexpr int kMaxPer = kLdsCap / R / kBlock;
  float2 v[kMaxPer][R];
  const int Ns = 1 << log2Ns;
  const int twshift = tw.log2N - (log2Ns + log2R);
#pragma unroll
  for (int u = 0; u < kMaxPer; ++u) {
    int b = threadIdx.x + u * kBlock;
    if (b < nbf) {
      int g = b & (G - 1);
      int j = b >> log2G;
      int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) v[u][r] = buf[((j + r * (P >> log2R)) << log2G) + g];
      if (log2Ns > 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[u][r] = cmul(v[u][r], tw.get((unsigned)(r * k) << twshift, sign));
      }
      if (R == 2) {
        float2 a = v[u][0], c = v[u][1];
        v[u][0] = cadd(a, c);
        v[u][1] = csub(a, c);
      } else if (R == 4) {
        float2 a0 = cadd(v[u][0], v[u][2]), a1 = csub(v[u][0], v[u][2]);
        float2 b0 = cadd(v[u][1], v[u][3]), b1 = cmul_si(csub(v[u][1], v[u][3]), sign);
        v[u][0] = cadd(a0, b0);
        v[u][2] = csub(a0, b0);
        v[u][1] = cadd(a1, b1);
        v[u][3] = csub(a1, b1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kMaxPer; ++u) {
    int b = threadIdx.x + u * kBlock;
    if (b < nbf) {
      int g = b & (G - 1);
      int j = b >> log2G;
      int k = j & (Ns - 1);
      int d = ((j >> log2Ns) << (log2Ns + log2R)) + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[((d + r * Ns) << log2G) + g] = v[u][r];
    }
  }
  __syncthreads();
}

// G interleaved in-place transforms of 2^log2P points:
//   out[k] = sum_j in[j] e^{sign 2 pi i j k / P}  (unnormalised)
// Requires G * 2^log2P <= kLdsCap, G a power of two, blockDim.x == kBlock, and
// a __syncthreads() between filling buf and the call.
__device__ inline void lds_fft(float2* buf, int log2P, int G, int log2G, const Twiddles& tw, float sign) {
  int log2Ns = 0;
  if (log2P & 1) {
    stockham_stage<2>(buf, log2P, 0, G, log2G, tw, sign);
    log2Ns = 1;
  }
  for (; log2Ns < log2P; log2Ns += 2) stockham_stage<4>(buf, log2P, log2Ns, G, log2G, tw, sign);
}

}  // namespace fcwt_hip
