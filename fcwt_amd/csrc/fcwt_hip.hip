// fcwt_hip.hip -- gfx950 implementation of fCWT's FCWT::cwt hot path behind
// the C ABI declared in include/fcwt_hip.h.
//
// The reference (src/fcwt/fcwt.cpp:447-529) pads the real signal to
// N = 2^find2power(n), takes one r2c FFT, then for every scale multiplies the
// one-sided spectrum by the Morlet window sampled from one mother table and
// runs a full N-point inverse FFT, keeping the first n samples.
//
// Here each scale's inverse DFT is band-limited: its spectrum is zero from
// bin `band` up (band <= N/2), so with M = 2^ceil(log2 band) the row factors
// as N = B * F with B <= kMaxB:
//   y[f + F*m] = sum_{qb<B} W_B^{qb*m} * U[qb][f],
//   U[qb][f]   = W_N^{qb*f} * sum_{qf<A} Z[qb + B*qf] W_F^{qf*f},  A = M/B
// where Z[q] = X[q] * mother[idx(q)] and W_K = e^{2 pi i / K}.
//   * "direct" rows (A == 1): the writer kernel forms U on the fly from the
//     spectrum and does G batched B-point inverse FFTs in LDS for G
//     consecutive f, writing G-contiguous runs of the output row.
//   * "two-pass" rows (A > 1): a spreader kernel does the F-point inverse
//     FFTs over qf and writes U to a scratch buffer ([f][qb] layout); the
//     writer then reads U instead of forming it.
// The forward FFT is a half-length complex FFT of the packed signal plus the
// usual real post-pass (one work-group for N/2 <= kLdsCap, otherwise two
// four-step passes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fcwt_hip.h"
#include "fft_lds.h"
#include "host_math.h"

namespace fcwt_hip {

constexpr int kMaxLog2B = 10;  // writer FFT size cap: G * B <= kLdsCap with G >= 16
constexpr double kStaleFloor = 1e-12;
constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr size_t kColsScratchBytes = size_t(128) << 20;  // column path: scratch slots per batch
// The no-allocation entry point: its plans live in a static device array of
// kNoAllocBytes, hold signals of at most kNoAllocMaxN padded samples, and
// give the two-pass kernels at most kNoAllocScratchBudget of scratch.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr int kNoAllocMaxN = 1 << 17;
constexpr size_t kNoAllocScratchBudget = size_t(16) << 20;
constexpr int kKinds = 8;  // kernel kinds timed by fcwt_hip_plan_profile

struct SegD {
  int lo, hi;
  float step;
  int pad;
};

struct RowD {
  int seg_begin, seg_count;
  int log2B;   // writer FFT size
  int log2A;   // spreader band factor (0: direct row)
  int log2G;   // consecutive f per writer block
  int slot;    // scratch slot (two-pass rows)
  long long out_off;  // offset of the row in its signal's output
  int sig;     // signal of the plan the row belongs to
  int n;       // that signal's length (the row has n outputs)
};

struct Item {
  int row;
  int block;
};

struct KArgs {
  const float2* spec;   // X[0, N/2) of each signal, N/2 apart
  const float* mother;  // N values
  const SegD* segs;
  const RowD* rows;
  const Item* items;
  float2* scratch;
  float2* out;          // output of a one-signal plan
  float2* const* outs;  // several signals: output of signal s at outs[s] (out unused)
  int N;
  int log2N;
  float norm;
  Twiddles tw;
};

__device__ __forceinline__ const float2* row_spec(const KArgs& a, const RowD& r) {
  return a.spec + (size_t)r.sig * (size_t)(a.N >> 1);
}
__device__ __forceinline__ float2* row_out(const KArgs& a, const RowD& r) {
  return (a.outs ? a.outs[r.sig] : a.out) + r.out_off;
}

// Mother table entry w (fcwt.cpp:51-67): the argument in float32 exactly as
// the reference forms it, the exponential in float64 rounded to float32.
// k_mother fills the table with it; the column kernels evaluate it in place.
__device__ __forceinline__ float mother_value(int w, float fb, float toradians, float norm) {
#pragma clang fp contract(off)
  const float kPi = 3.14159265358979323846264338327950288f;
  float a = (2.0f * ((float)w * toradians) * fb - 2.0f * kPi * fb);
  a = -(a * a) / 2;
  return norm * (float)exp((double)a);
}

}  // namespace fcwt_hip

#include "fft_rows.h"
#include "rows_direct.h"
#include "rows_team.h"
#include "rows_big.h"

namespace fcwt_hip {

// phases per work-group of k_rows_direct<log2M> (before clamping to L)
static int direct_phases(int log2M) {
  switch (log2M) {
    case 0: return DirectCfg<0>::G;
    case 1: return DirectCfg<1>::G;
    case 2: return DirectCfg<2>::G;
    case 3: return DirectCfg<3>::G;
    case 4: return DirectCfg<4>::G;
    case 5: return DirectCfg<5>::G;
    case 6: return DirectCfg<6>::G;
    case 7: return DirectCfg<7>::G;
    case 8: return DirectCfg<8>::G;
    case 9: return DirectCfg<9>::G;
    default: return DirectCfg<10>::G;
  }
}
// the same for a block run by the team kernel (kFillMinLog2M <= log2M <= 10)
static int direct_phases_team(int log2M) {
  switch (log2M) {
    case 0: return DirectCfg<0, true>::G;
    case 1: return DirectCfg<1, true>::G;
    case 2: return DirectCfg<2, true>::G;
    case 3: return DirectCfg<3, true>::G;
    case 4: return DirectCfg<4, true>::G;
    case 5: return DirectCfg<5, true>::G;
    case 6: return DirectCfg<6, true>::G;
    case 7: return DirectCfg<7, true>::G;
    case 8: return DirectCfg<8, true>::G;
    case 9: return DirectCfg<9, true>::G;
    default: return DirectCfg<10, true>::G;
  }
}

// Z[q] = X[q] * mother[(int)fminf(N-1, step*q)] for the segment holding q
// (fcwt.cpp:224-248), 0 outside every segment.
__device__ __forceinline__ float2 spectrum_bin(const KArgs& a, const RowD& r, int q) {
#pragma clang fp contract(off)
  for (int s = 0; s < r.seg_count; ++s) {
    SegD sg = a.segs[r.seg_begin + s];
    if (q >= sg.lo && q < sg.hi) {
      float v = sg.step * (float)q;
      float mx = (float)a.N - 1.0f;
      int idx = (int)(v < mx ? v : mx);
      float w = a.mother[idx];
      float2 x = row_spec(a, r)[q];
      return make_float2(x.x * w, x.y * w);
    }
  }
  return make_float2(0.f, 0.f);
}

// Writer: one work-group per (row, block of G consecutive f).
__global__ __launch_bounds__(kBlock) void k_writer(KArgs a) {
  __shared__ float2 lds[kLdsCap];
  const Item it = a.items[blockIdx.x];
  const RowD r = a.rows[it.row];
  const int log2B = r.log2B, B = 1 << log2B;
  const int log2G = r.log2G, G = 1 << log2G;
  const int log2F = a.log2N - log2B;
  const int F = 1 << log2F;
  const int f0 = it.block << log2G;
  const int tot = B << log2G;
  if (r.log2A == 0) {
    for (int e = threadIdx.x; e < tot; e += kBlock) {
      int g = e & (G - 1), qb = e >> log2G;
      float2 z = spectrum_bin(a, r, qb);
      float2 w = a.tw((unsigned)(qb * (f0 + g)));
      lds[e] = cmul(z, w);
    }
  } else {
    const float2* U = a.scratch + (size_t)r.slot * a.N;
    for (int e = threadIdx.x; e < tot; e += kBlock) {
      int qb = e & (B - 1), g = e >> log2B;
      lds[(qb << log2G) + g] = U[(size_t)(f0 + g) * B + qb];
    }
  }
  __syncthreads();
  lds_fft(lds, log2B, G, log2G, a.tw, 1.0f);
  float2* out = row_out(a, r);
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int g = e & (G - 1), m = e >> log2G;
    long long t = (long long)(f0 + g) + (long long)F * m;
    if (t < r.n) out[t] = cscale(lds[e], a.norm);
  }
}

// Spreader: one work-group per (two-pass row, block of Q consecutive qb);
// F-point inverse FFT over qf for each qb, twiddle by W_N^{qb*f}, store
// U[f][qb] in the row's scratch slot.
__global__ __launch_bounds__(kBlock) void k_spreader(KArgs a) {
  __shared__ float2 lds[kLdsCap];
  const Item it = a.items[blockIdx.x];
  const RowD r = a.rows[it.row];
  const int log2B = r.log2B, B = 1 << log2B;
  const int log2F = a.log2N - log2B;
  const int F = 1 << log2F;
  const int A = 1 << r.log2A;
  const int log2Q = min(log2B, 14 - log2F);
  const int Q = 1 << log2Q;
  const int qb0 = it.block << log2Q;
  const int tot = F << log2Q;
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int j = e & (Q - 1), qf = e >> log2Q;
    float2 z = make_float2(0.f, 0.f);
    if (qf < A) z = spectrum_bin(a, r, qb0 + j + (qf << log2B));
    lds[e] = z;
  }
  __syncthreads();
  lds_fft(lds, log2F, Q, log2Q, a.tw, 1.0f);
  float2* U = a.scratch + (size_t)r.slot * a.N;
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int j = e & (Q - 1), f = e >> log2Q;
    int qb = qb0 + j;
    U[(size_t)f * B + qb] = cmul(lds[e], a.tw((unsigned)(qb * f)));
  }
}

// Forward transform of the zero-padded input to X[0, N/2).
//   real input (fcwt.cpp:491-500): z[j] = x[2j] + i x[2j+1], an H = N/2 point
//     complex FFT, then the r2c post-pass;
//   complex input (fcwt.cpp:486-490): z[j] = x[j], an H = N point complex FFT
//     of which the first `keep` outputs are kept (N/2 for the cwt).
// Several signals at once: grid dimension y indexes them, each at its own
// stride in x, spec and atmp.
struct FwdArgs {
  const float* x;  // the input as floats (interleaved re, im for complex input)
  int n;           // floats in x per signal
  const float* const* xs;  // if set: signal y's input at xs[y] with ns[y] * len_mul floats (x, n unused)
  const int* ns;
  int len_mul;
  float2* spec;    // X[0, keep) (real input: X[0, N/2))
  float2* ztmp;    // H per signal (four-step, real input only)
  float2* atmp;    // H per signal (four-step only)
  long long x_stride;     // floats between signals
  long long spec_stride;  // complex values between signals' outputs
  int keep;               // complex input: outputs kept
  int log2H;
  int log2H1, log2H2;
  int tw_shift;    // log2(N / H): W_H^r = W_N^{r << tw_shift}
  int complex_in;
  Twiddles tw;
};

__device__ __forceinline__ const float* fwd_input(const FwdArgs& a, int y) {
  return a.xs ? a.xs[y] : a.x + y * a.x_stride;
}
__device__ __forceinline__ int fwd_len(const FwdArgs& a, int y) { return a.ns ? a.ns[y] * a.len_mul : a.n; }

// z[j] = x[2j] + i x[2j+1] (zero past the n floats of x)
__device__ __forceinline__ float2 packed(const float* x, int n, int j) {
  long long i0 = 2LL * j;
  float re = i0 < n ? x[i0] : 0.f;
  float im = i0 + 1 < n ? x[i0 + 1] : 0.f;
  return make_float2(re, im);
}

// X[k] = (Z[k] + conj Z[H-k]) / 2 - (i/2) e^{-2 pi i k/N} (Z[k] - conj Z[H-k])
__device__ __forceinline__ float2 real_post(float2 zk, float2 zc, float2 w) {
  float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y - zc.y));
  float2 o = make_float2(0.5f * (zk.y + zc.y), -0.5f * (zk.x - zc.x));  // (Z[k] - conj Z[H-k]) / (2i)
  return cadd(e, cmul(w, o));
}

// Whole forward transform of one signal per work-group (H <= 2^14) with the
// radix-16 register passes of fft_rows.h; grid dimension y indexes signals.
// DERIVED selects derived radix-16 twiddles (fft_rows.h): used for the
// adjoint's transforms, not for the signal's spectrum, whose absolute error
// every row inherits.
template <int LOG2H, bool DERIVED>
__global__ __launch_bounds__(kRowFftThreads) void k_rows_fft(FwdArgs a) {
  __shared__ RowFftLds<LOG2H> s;
  constexpr int H = 1 << LOG2H;
  constexpr int PER = (H + kRowFftThreads - 1) / kRowFftThreads;  // points per thread
  const float* x = fwd_input(a, blockIdx.y);
  const int nx = fwd_len(a, blockIdx.y);
  float2* spec = a.spec + blockIdx.y * a.spec_stride;
  row_tw_tables<LOG2H>(s);
  if ((nx & 1) == 0) {
    // whole complex values: every load issued before the first is used;
    // the buffer range (nx floats) reads zero past the input, which is the
    // zero padding
    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void*)x, 0, nx * 4, 0x00020000);
    float2 v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = threadIdx.x + u * kRowFftThreads;
      const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(rx, (unsigned)j * 8u, 0, 0);
      v[u] = make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = threadIdx.x + u * kRowFftThreads;
      if (j < H) s.data[row_pad(j)] = v[u];
    }
  } else {
    for (int j = threadIdx.x; j < H; j += kRowFftThreads) s.data[row_pad(j)] = packed(x, nx, j);
  }
  __syncthreads();
  row_fft<LOG2H, -1, DERIVED>(s);
  if (a.complex_in) {
    if (a.keep == H / 2 && H >= 2 * kRowFftThreads) {
      // every LDS read first, then the stores
      constexpr int KP = H / 2 / kRowFftThreads;
      float2 o[KP > 0 ? KP : 1];
#pragma unroll
      for (int u = 0; u < KP; ++u) o[u] = s.data[row_pad(threadIdx.x + u * kRowFftThreads)];
#pragma unroll
      for (int u = 0; u < KP; ++u) spec[threadIdx.x + u * kRowFftThreads] = o[u];
    } else {
      for (int k = threadIdx.x; k < a.keep; k += kRowFftThreads) spec[k] = s.data[row_pad(k)];
    }
    return;
  }
  for (int k = threadIdx.x; k < H; k += kRowFftThreads) {
    float2 zk = s.data[row_pad(k)], zc = s.data[row_pad((H - k) & (H - 1))];
    spec[k] = real_post(zk, zc, a.tw.get((unsigned)k, -1.0f));
  }
}

template <bool DERIVED>
static void launch_rows_fft_t(const FwdArgs& f, unsigned batch, hipStream_t st) {
  const dim3 g(1, batch);
  switch (f.log2H) {
    case 0: k_rows_fft<0, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 1: k_rows_fft<1, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 2: k_rows_fft<2, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 3: k_rows_fft<3, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 4: k_rows_fft<4, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 5: k_rows_fft<5, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 6: k_rows_fft<6, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 7: k_rows_fft<7, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 8: k_rows_fft<8, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 9: k_rows_fft<9, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 10: k_rows_fft<10, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 11: k_rows_fft<11, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 12: k_rows_fft<12, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    case 13: k_rows_fft<13, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
    default: k_rows_fft<14, DERIVED><<<g, kRowFftThreads, 0, st>>>(f); break;
  }
}

static void launch_rows_fft(const FwdArgs& f, unsigned batch, hipStream_t st, bool derived = false) {
  if (derived)
    launch_rows_fft_t<true>(f, batch, st);
  else
    launch_rows_fft_t<false>(f, batch, st);
}

// Transforms per work-group in the four-step passes: as many as fit the LDS,
// but at most 2^kFwdLog2G, so that one signal's pass spreads over many CUs
// (the passes are latency-bound: 1 MB moves in each).
constexpr int kFwdLog2G = 2;
__host__ __device__ constexpr int fwd_log2g(int log2Hthis, int log2Hother) {
  return log2Hother < 14 - log2Hthis ? (log2Hother < kFwdLog2G ? log2Hother : kFwdLog2G)
                                      : (14 - log2Hthis < kFwdLog2G ? 14 - log2Hthis : kFwdLog2G);
}

// Four-step pass 1: H1-point FFTs over j1 of z[j2 + H2*j1] for G consecutive
// j2, times W_H^{-j2*k1}, stored to atmp[k1*H2 + j2].
__global__ __launch_bounds__(kBlock) void k_fwd_a(FwdArgs a) {
  __shared__ float2 lds[kLdsCap];
  const int H1 = 1 << a.log2H1, H2 = 1 << a.log2H2;
  const int log2G = fwd_log2g(a.log2H1, a.log2H2);
  const int G = 1 << log2G;
  const int j20 = blockIdx.x << log2G;
  const int tot = H1 << log2G;
  const float* x = fwd_input(a, blockIdx.y);
  const int nx = fwd_len(a, blockIdx.y);
  float2* atmp = a.atmp + ((size_t)blockIdx.y << a.log2H);
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int g = e & (G - 1), j1 = e >> log2G;
    lds[e] = packed(x, nx, j20 + g + H2 * j1);
  }
  __syncthreads();
  lds_fft(lds, a.log2H1, G, log2G, a.tw, -1.0f);
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int g = e & (G - 1), k1 = e >> log2G;
    int j2 = j20 + g;
    atmp[(size_t)k1 * H2 + j2] = cmul(lds[e], a.tw.get((unsigned)(j2 * k1) << a.tw_shift, -1.0f));
  }
}

// Four-step pass 2: H2-point FFTs over j2 of atmp[k1*H2 + j2] for G
// consecutive k1, stored to ztmp[k1 + H1*k2] (complex input: straight to
// spec, first N/2 outputs only).
__global__ __launch_bounds__(kBlock) void k_fwd_b(FwdArgs a) {
  __shared__ float2 lds[kLdsCap];
  const int H1 = 1 << a.log2H1, H2 = 1 << a.log2H2;
  const int log2G = fwd_log2g(a.log2H2, a.log2H1);
  const int G = 1 << log2G;
  const int k10 = blockIdx.x << log2G;
  const int tot = H2 << log2G;
  const float2* atmp = a.atmp + ((size_t)blockIdx.y << a.log2H);
  float2* spec = a.spec + blockIdx.y * a.spec_stride;
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int j2 = e & (H2 - 1), g = e >> a.log2H2;
    lds[(j2 << log2G) + g] = atmp[(size_t)(k10 + g) * H2 + j2];
  }
  __syncthreads();
  lds_fft(lds, a.log2H2, G, log2G, a.tw, -1.0f);
  for (int e = threadIdx.x; e < tot; e += kBlock) {
    int g = e & (G - 1), k2 = e >> log2G;
    const size_t k = (size_t)(k10 + g) + (size_t)H1 * k2;
    if (!a.complex_in)
      a.ztmp[((size_t)blockIdx.y << a.log2H) + k] = lds[e];
    else if (k < (size_t)a.keep)
      spec[k] = lds[e];
  }
}

__global__ void k_fwd_post(FwdArgs a) {
  const int H = 1 << a.log2H;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  const float2* z = a.ztmp + ((size_t)blockIdx.y << a.log2H);
  if (k < H) a.spec[blockIdx.y * a.spec_stride + k] = real_post(z[k], z[(H - k) & (H - 1)], a.tw.get((unsigned)k, -1.0f));
}

// ---- adjoint (input gradient) kernels
// Window the forward used for bin q of a row: mother[(int)fminf(N-1, step q)]
// in the row's segment holding q, 0 outside (the same segments as the forward).
__device__ __forceinline__ float row_window(const SegD* segs, const RowD& r, int q, float maxidx,
                                            const float* mother) {
#pragma clang fp contract(off)
  for (int s = 0; s < r.seg_count; ++s) {
    const SegD sg = segs[r.seg_begin + s];
    if (q >= sg.lo && q < sg.hi) {
      float v = sg.step * (float)q;
      return mother[(int)(v < maxidx ? v : maxidx)];
    }
  }
  return 0.f;
}

// For signal s = blockIdx.y of a batch (fn rows each, rows numbered s*fn + i
// overall): A_s[q] (+)= sum of w_i(q) Ghat[s*fn + i][q] over the rows i of s
// that lie in this chunk of rows [row0, row0 + count), q < N/2, with
// Ghat[r] = P[r - row0].  A_s is stored conjugated once its last row is in,
// ready for the final forward transform.
__global__ void k_adj_reduce(const float2* P, long long row0, int count, int fn, const RowD* rows, const SegD* segs,
                             const float* mother, int N, float2* A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int half = N / 2;
  const long long s0 = (long long)blockIdx.y * fn;
  const long long lo = max(row0, s0), hi = min(row0 + count, s0 + fn);
  if (q >= half || lo >= hi) return;
  const float maxidx = (float)N - 1.0f;
  float2* a = A + (size_t)blockIdx.y * half;
  float2 acc = lo == s0 ? make_float2(0.f, 0.f) : a[q];
  long long r = lo;
  // four rows at a time, their loads issued together; the sum keeps row order
  for (; r + 4 <= hi; r += 4) {
    float w[4];
    float2 g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      w[k] = row_window(segs, rows[r + k - s0], q, maxidx, mother);
      g[k] = P[(size_t)(r + k - row0) * half + q];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (w[k] != 0.f) {
        acc.x += w[k] * g[k].x;
        acc.y += w[k] * g[k].y;
      }
    }
  }
  for (; r < hi; ++r) {
    const float w = row_window(segs, rows[r - s0], q, maxidx, mother);
    if (w != 0.f) {
      const float2 g = P[(size_t)(r - row0) * half + q];
      acc.x += w * g.x;
      acc.y += w * g.y;
    }
  }
  a[q] = hi == s0 + fn ? make_float2(acc.x, -acc.y) : acc;
}

// Rows of the fused adjoint per work-group: a work-group transforms this many
// consecutive gradient rows of one signal and sums their windowed spectra in
// registers.  Fixed, so that a signal's sum is grouped the same way in a
// batch as alone.
constexpr int kAdjRows = 4;

// The adjoint's row transforms fused with the window and the sum over rows,
// for N = 2^LOG2H <= 2^14 (one work-group holds a row's transform in LDS).
// Work-group (g, s) takes rows [g R, g R + R) of signal s (R = kAdjRows, fn
// rows per signal, n complex values each at grad + (s fn + i) n):
//   Ghat_i = FFT_N(G_i) (zero-padded), acc[q] += w_i(q) Ghat_i[q] for q < N/2,
// where the window is nonzero (the row's band), and stores acc to
// part[(s groups + g) N/2 + q].  Each gradient row is read once and its
// spectrum never leaves the chip.
template <int LOG2H>
__global__ __launch_bounds__(kRowFftThreads) void k_adj_rows(const float2* grad, int n, int fn, int groups,
                                                            const RowD* rows, const SegD* segs,
                                                            const float* mother, float2* part) {
  __shared__ RowFftLds<LOG2H> s;
  constexpr int H = 1 << LOG2H, HALF = H / 2;
  constexpr int PER = (H + kRowFftThreads - 1) / kRowFftThreads;
  constexpr int KQ = (HALF + kRowFftThreads - 1) / kRowFftThreads;
  const int g = blockIdx.x, sig = blockIdx.y;
  const int i0 = g * kAdjRows, i1 = min(fn, i0 + kAdjRows);
  const float maxidx = (float)H - 1.0f;
  float2 acc[KQ];
#pragma unroll
  for (int u = 0; u < KQ; ++u) acc[u] = make_float2(0.f, 0.f);
  row_tw_tables<LOG2H>(s);
  for (int i = i0; i < i1; ++i) {
    const float2* x = grad + ((size_t)sig * fn + i) * (size_t)n;
    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void*)x, 0, n * 8, 0x00020000);
    float2 v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(rx, (unsigned)(threadIdx.x + u * kRowFftThreads) * 8u, 0, 0);
      v[u] = make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
    }
    __syncthreads();  // the previous row's spectrum has been read
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = threadIdx.x + u * kRowFftThreads;
      if (j < H) s.data[row_pad(j)] = v[u];
    }
    __syncthreads();
    row_fft<LOG2H, -1, true>(s);
    // the window of each of this thread's bins: one scan of the row's
    // segments for all of them (the first segment holding q, as
    // row_window), then every gather of the mother table issued before the
    // first is used; bins outside every segment (beyond the band) are
    // neither windowed nor summed
    const RowD r = rows[i];
    int idx[KQ];
#pragma unroll
    for (int u = 0; u < KQ; ++u) idx[u] = -1;
    for (int k = 0; k < r.seg_count; ++k) {
#pragma clang fp contract(off)
      const SegD sg = segs[r.seg_begin + k];
#pragma unroll
      for (int u = 0; u < KQ; ++u) {
        const int q = threadIdx.x + u * kRowFftThreads;
        if (idx[u] < 0 && q < HALF && q >= sg.lo && q < sg.hi) {
          const float v = sg.step * (float)q;
          idx[u] = (int)(v < maxidx ? v : maxidx);
        }
      }
    }
    float wv[KQ];
#pragma unroll
    for (int u = 0; u < KQ; ++u) wv[u] = idx[u] >= 0 ? mother[idx[u]] : 0.f;
#pragma unroll
    for (int u = 0; u < KQ; ++u) {
      const int q = threadIdx.x + u * kRowFftThreads;
      if (idx[u] >= 0) {
        const float2 z = s.data[row_pad(q)];
        acc[u].x += wv[u] * z.x;
        acc[u].y += wv[u] * z.y;
      }
    }
  }
  float2* out = part + ((size_t)sig * groups + g) * HALF;
#pragma unroll
  for (int u = 0; u < KQ; ++u) {
    const int q = threadIdx.x + u * kRowFftThreads;
    if (q < HALF) out[q] = acc[u];
  }
}

// A_s[q] = conj(sum over the groups g of part[(s groups + g) N/2 + q]), in
// group order, for q < N/2; s = blockIdx.y.
__global__ void k_adj_combine(const float2* part, int groups, int half, float2* A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= half) return;
  const float2* p = part + (size_t)blockIdx.y * groups * half + q;
  float2 a = make_float2(0.f, 0.f);
  for (int g = 0; g < groups; ++g) {
    const float2 v = p[(size_t)g * half];
    a.x += v.x;
    a.y += v.y;
  }
  A[(size_t)blockIdx.y * half + q] = make_float2(a.x, -a.y);
}

// xbar_s[t] = c Re y_s[t] (real input) or c conj(y_s[t]) (complex input),
// t < n, s = blockIdx.y, where y_s = FFT_N(conj A_s) (N apart).
__global__ void k_adj_out(const float2* y, int n, int N, float c, int complex_in, float* xbar) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float2 v = y[(size_t)blockIdx.y * N + t];
  if (complex_in) {
    xbar[2 * ((size_t)blockIdx.y * n + t)] = c * v.x;
    xbar[2 * ((size_t)blockIdx.y * n + t) + 1] = -c * v.y;
  } else {
    xbar[(size_t)blockIdx.y * n + t] = c * v.x;
  }
}

// Mother table (fcwt.cpp:51-67), one entry per thread.
__global__ void k_mother(float* mother, int N, float fb, float toradians, float norm) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < N) mother[w] = mother_value(w, fb, toradians, norm);
}

// Twiddle tables, float64 sincospi rounded to float32.
__global__ void k_twiddles(float2* t1, int n1, float2* t2, int n2, int s1bits, int log2N) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double invN = ldexp(1.0, -log2N);
  if (i < n1) {
    double s, c;
    sincospi(2.0 * i * invN, &s, &c);
    t1[i] = make_float2((float)c, (float)s);
  }
  if (i < n2) {
    double s, c;
    sincospi(2.0 * ldexp((double)i, s1bits) * invN, &s, &c);
    t2[i] = make_float2((float)c, (float)s);
  }
}

}  // namespace fcwt_hip

using namespace fcwt_hip;

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

// Tuned schedule parameters per padded size N (fcwt_hip_tune,
// fcwt_hip_tune_load), and the tuner's own override while it times
// candidates.  Zero fields keep the built-in choice.
struct Tuned {
  int team_buffers = 0;  // XCD-team scratch buffers per team
  int cols_mb = 0;       // column path: scratch MB per batch of rows
  int path = 0;          // wide rows: 0 built-in, 1 XCD-team kernel, 2 column path
  int team_lag = 0;      // steps between a row's pass 1 and its pass 2 in a team's sequence (< team_buffers)
  int direct_max = 0;    // widest band (log2 M) on the single-work-group kernel: 9 or 10
  int cols_sets = 0;     // column path with several batches: 1 separate passes, 2 fused launches
  int cols_log2B = 0;    // column path at N = 2^22: N split as 2^8 x 2^14 (8) or 2^9 x 2^13 (9)
};
static std::mutex g_tune_mu;
static std::map<int, Tuned> g_tuned;
static thread_local bool t_forcing = false;
static thread_local Tuned t_force;

static Tuned tuned(int N) {
  if (t_forcing) return t_force;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(N);
  return it == g_tuned.end() ? Tuned{} : it->second;
}

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(FCWT_HIP_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// -------------------------------------------------------------------- plan
struct fcwt_hip_plan_s {
  int device = 0;
  int n = 0, N = 1, log2N = 0, fn = 0, threads = 1;  // n: signal 0's length; fn: rows of all signals
  // signals of the plan (all padded to the same N): length, rows, first row
  int nsig = 1;
  std::vector<int> sig_n, sig_fn, sig_row0;
  int* d_sig_n = nullptr;
  const float** d_xs = nullptr;   // per-signal input pointers (several signals)
  float2** d_outs = nullptr;      // per-signal output pointers
  std::vector<const void*> up_xs, up_outs;  // the pointer arrays last uploaded
  float fb = 0.f;
  bool normalize = true;
  unsigned flags = 0;
  std::vector<float> scales;
  // schedule
  std::vector<RowD> rows;
  std::vector<SegD> segs;
  std::vector<Item> items;  // direct writer items, then per batch: spreader items, writer items
  // XCD-team path for wide bands (rows_team.h)
  bool team_capable = false;
  int team_size = 0, nteams = 0, team_log2F = 0, team_nbuf = 2, team_lag = 1, seq_stride = 0;
  std::vector<int> team_rows, team_seq, team_blocks;
  std::vector<TeamBlk> team_desc;  // team_seq with each block's row fields
  TeamBlk* d_team_desc = nullptr;
  int* d_team_blocks = nullptr;
  float2* d_team_scratch = nullptr;
  unsigned* d_team_sync = nullptr;
  unsigned* d_error = nullptr;
  size_t team_sync_words = 0;
  size_t direct_items = 0;
  struct Group {
    int log2M;
    size_t begin, count;
    bool fill;  // run by the team kernel's members (TeamArgs::fill_items), not by a launch of its own
  };
  std::vector<Group> direct_groups;
  size_t fill_begin = 0, fill_count = 0;  // the fill groups' items, contiguous in items
  int fill_end[kFillGroups] = {};
  struct Batch {
    size_t s_begin, s_count, w_begin, w_count;
  };
  std::vector<Batch> batches;
  int scratch_slots = 0;
  // column path for wide bands at N = 256 F, F = 2^7 .. 2^14 (rows_big.h):
  // batches of rows cols_rows[begin, begin + count), row k of a batch in
  // scratch slot k
  int cols_log2F = 0;
  int cols_log2B = kTeamLog2B;        // N = 2^cols_log2B x 2^cols_log2F on the column path
  int cols_slots = 0;
  std::vector<int> cols_rows;
  struct ColsBatch {
    int begin, count;
  };
  std::vector<ColsBatch> cols_batches;
  int* d_cols_rows = nullptr;
  int cols_sets = 1;                  // scratch sets: 2 when there are several batches
  float2* d_cols_scratch = nullptr;   // cols_sets x cols_slots x N
  float2* d_xt = nullptr;  // transposed half-spectra
  float m_toradians = 0.f, m_norm = 0.f;  // mother-table parameters
  // device
  float* d_mother = nullptr;
  float2* d_tw1 = nullptr;
  float2* d_tw2 = nullptr;
  int s1bits = 0;
  float2* d_spec = nullptr;
  float2* d_ztmp = nullptr;
  float2* d_atmp = nullptr;
  SegD* d_segs = nullptr;
  RowD* d_rows = nullptr;
  Item* d_items = nullptr;
  float2* d_scratch = nullptr;
  float* d_x = nullptr;
  float2* d_out = nullptr;
  // adjoint workspace (allocated by the first adjoint call)
  int adj_chunk = 0;
  float2* d_adj_P = nullptr;     // adj_chunk x N/2 row spectra
  float2* d_adj_atmp = nullptr;  // adj_chunk x N four-step temporaries (N > 2^14)
  int adj_signals = 0;           // signals the A / y buffers hold
  float2* d_adj_A = nullptr;     // adj_signals x N/2
  float2* d_adj_y = nullptr;     // adj_signals x N
  float* d_grad = nullptr;       // host-pointer adjoint: fn x n complex64 staging
  float* d_xbar = nullptr;
  // host-pointer output pipeline: two pinned staging buffers and their events
  char* h_stage[2] = {nullptr, nullptr};
  size_t stage_bytes = 0;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: the caller's stream (no-allocation entry point)
  size_t bytes = 0;
  // arena mode (fcwt_hip_cwt_device_noalloc): device memory is carved from
  // the library's static device array instead of hipMalloc
  char* arena = nullptr;
  size_t arena_cap = 0;
  size_t wide_budget = kScratchBytes;  // generic two-pass kernels' scratch
  size_t cols_budget = 0;              // column path, per batch of rows (0: tuned or default)
  // per-kernel timing (fcwt_hip_plan_profile): event pairs around launches
  struct Timed {
    int kind;
    hipEvent_t a, b;
  };
  bool profiling = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> event_pool;
  double prof_ms[kKinds] = {};
  long long prof_launches[kKinds] = {};
};

// kernel kinds reported by fcwt_hip_plan_profile_read
static const char* const kKindNames[kKinds] = {"forward_fft", "row_writer_direct", "row_spreader",
                                               "row_writer_two_pass", "row_team", "adjoint",
                                               "row_columns", "column_transpose"};

static hipEvent_t take_event(fcwt_hip_plan p) {
  if (!p->event_pool.empty()) {
    hipEvent_t e = p->event_pool.back();
    p->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Records an event before (begin) / after (end) a launch of `kind` on st.
struct ProfScope {
  fcwt_hip_plan p;
  int kind;
  hipStream_t st;
  hipEvent_t a = nullptr;
  ProfScope(fcwt_hip_plan p_, int k, hipStream_t s) : p(p_), kind(k), st(s) {
    if (p->profiling && (a = take_event(p))) (void)hipEventRecord(a, st);
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = take_event(p);
    if (b && hipEventRecord(b, st) == hipSuccess)
      p->pending.push_back({kind, a, b});
    else
      p->event_pool.push_back(a);
  }
};

static void drain_profile(fcwt_hip_plan p) {
  for (auto& t : p->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      p->prof_ms[t.kind] += ms;
      p->prof_launches[t.kind] += 1;
    }
    p->event_pool.push_back(t.a);
    p->event_pool.push_back(t.b);
  }
  p->pending.clear();
}

static void free_plan(fcwt_hip_plan p) {
  if (!p) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(p->device);
  drain_profile(p);
  for (hipEvent_t e : p->event_pool) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; ++k) {
    if (p->h_stage[k]) (void)hipHostFree(p->h_stage[k]);
    if (p->stage_ev[k]) (void)hipEventDestroy(p->stage_ev[k]);
  }
  if (p->arena) {  // nothing was allocated: the memory belongs to the static array
    if (p->stream && p->own_stream) (void)hipStreamDestroy(p->stream);
    (void)hipSetDevice(cur);
    delete p;
    return;
  }
  void* ptrs[] = {p->d_mother, p->d_tw1, p->d_tw2, p->d_spec, p->d_ztmp, p->d_atmp, p->d_segs,
                  p->d_rows,   p->d_items, p->d_scratch, p->d_x, p->d_out,
                  p->d_team_desc, p->d_team_blocks, p->d_team_scratch, p->d_team_sync, p->d_error,
                  p->d_adj_P, p->d_adj_atmp, p->d_adj_A, p->d_adj_y, p->d_grad, p->d_xbar,
                  p->d_sig_n, p->d_xs, p->d_outs, p->d_cols_rows, p->d_cols_scratch, p->d_xt};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->stream && p->own_stream) (void)hipStreamDestroy(p->stream);
  (void)hipSetDevice(cur);
  delete p;
}

// FCWT_HIP_DEVICE_BUDGET_MB=m caps the device memory one plan may hold (its
// tables, spectra, schedule and scratch; not the caller's buffers): an
// allocation beyond it fails with FCWT_HIP_ENOMEM, as a failed hipMalloc does.
static size_t device_budget() {
  const char* e = std::getenv("FCWT_HIP_DEVICE_BUDGET_MB");
  const long long v = e ? std::atoll(e) : 0;
  return v > 0 ? (size_t)v << 20 : SIZE_MAX;
}

template <class T>
static int dalloc(fcwt_hip_plan p, T** ptr, size_t count) {
  if (count == 0) count = 1;
  if (p->arena) {  // 256-byte aligned pieces of the static array, no allocation
    const size_t at = (p->bytes + 255) & ~size_t(255);
    if (at + count * sizeof(T) > p->arena_cap) {
      *ptr = nullptr;
      return fail(FCWT_HIP_EUNSUPPORTED, "this plan needs more than the " + std::to_string(p->arena_cap >> 20) +
                                             " MiB of device memory the no-allocation entry point holds");
    }
    *ptr = reinterpret_cast<T*>(p->arena + at);
    p->bytes = at + count * sizeof(T);
    return FCWT_HIP_OK;
  }
  if (p->bytes + count * sizeof(T) > device_budget()) {
    *ptr = nullptr;
    return fail(FCWT_HIP_ENOMEM, "device allocation of " + std::to_string(count * sizeof(T)) +
                                     " bytes exceeds the plan's budget (FCWT_HIP_DEVICE_BUDGET_MB)");
  }
  hipError_t e = hipMalloc((void**)ptr, count * sizeof(T));
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(FCWT_HIP_ENOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                     " bytes failed: " + hipGetErrorString(e));
  }
  p->bytes += count * sizeof(T);
  return FCWT_HIP_OK;
}

static Twiddles twiddles_of(const fcwt_hip_plan p) {
  Twiddles t;
  t.t1 = p->d_tw1;
  t.t2 = p->d_tw2;
  t.s1bits = p->s1bits;
  t.log2N = p->log2N;
  return t;
}

// FCWT_HIP_NO_COLS=1 in the environment routes the column path's rows to
// the generic two-pass kernels (used by tests of those).
static bool cols_disabled() {
  const char* e = std::getenv("FCWT_HIP_NO_COLS");
  return e && e[0] == '1';
}
// log2 B of the column path's split N = B F.  B = 256, except at N = 2^22,
// which splits as 512 x 2^13: a pass-1 work-group of 2^13 points leaves room
// for a second one and for pass-2 work-groups on its CU, where one of 2^14
// points fills the CU's LDS alone, and a half band of 2^13 points skips the
// radix-2 pass.  A tuning file's cols_log2B, or FCWT_HIP_COLS_B256=1 in the
// environment, keeps 256 x 2^14.
static int cols_split(int log2N, const Tuned& tn) {
  if (log2N != 22) return kTeamLog2B;
  const char* e = std::getenv("FCWT_HIP_COLS_B256");
  if (e && e[0] == '1') return kTeamLog2B;
  return tn.cols_log2B == kTeamLog2B ? kTeamLog2B : 9;
}
// Whether the direct rows (bands 2^kFillMinLog2M .. 2^10) of a plan with a
// team launch run inside it.  Inside, they run at the team kernel's
// occupancy (3 work-groups per CU, against 4 waves per SIMD in their own
// launches) and, at M = 1024, in 64-byte output runs, but they fill the time
// the team's last rows leave idle and save their own launches' ramps.
// Measured: one-signal plans 12-26% faster with team_rows * N up to 2^25
// (e.g. 200 scales at n = 20k..100k), even at 3e7, 4% slower at 1e8 and 3e8
// (1000 and 3000 scales at n = 100k); the plans of the 384-channel batch
// (several signals each, 19.7k team rows at N = 2^17) 1.3% faster.  So
// plans of several signals always fill, one-signal plans up to 2^25.
// FCWT_HIP_FILL=1 / FCWT_HIP_NO_FILL=1 force either way.
constexpr double kFillMaxTeamPoints = 33554432.0;  // 2^25
static bool fill_rows(size_t team_rows, int N, int signals) {
  const char* off = std::getenv("FCWT_HIP_NO_FILL");
  if (off && off[0] == '1') return false;
  const char* on = std::getenv("FCWT_HIP_FILL");
  if (on && on[0] == '1') return team_rows > 0;
  return team_rows > 0 && (signals > 1 || (double)team_rows * (double)N <= kFillMaxTeamPoints);
}

// Scratch buffers per team by default: as many N-point buffers as fit
// kTeamScratchPerXcd (6 MiB), within [kTeamMinBuffers, kTeamBuffers].  At
// N = 2^18 that is 6 buffers of 2 MiB (12 MiB per team, most of it in the
// memory-side cache): fewer leave the sequence too little slack.
constexpr int kTeamMinBuffers = 6;
static int team_default_buffers(int N) {
  const size_t row_bytes = sizeof(float2) * (size_t)N;
  return (int)std::max<size_t>(kTeamMinBuffers, std::min<size_t>(kTeamBuffers, kTeamScratchPerXcd / row_bytes));
}

static int build_schedule(fcwt_hip_plan p, const std::vector<float>& mother_host) {
  const bool avx = !(p->flags & FCWT_HIP_COVERAGE_SCALAR);
  // each signal's rows, with that signal's reuse of stale bins (one
  // reference cwt call per signal)
  std::vector<RowBand> bands;
  for (int sgi = 0; sgi < p->nsig; ++sgi) {
    std::vector<RowBand> b = build_bands(p->N, p->scales.data() + p->sig_row0[sgi], p->sig_fn[sgi], p->threads, avx,
                                         mother_host.data(), kStaleFloor);
    bands.insert(bands.end(), b.begin(), b.end());
  }
  p->rows.resize(p->fn);
  std::vector<int> wide, team, cols;
  const int team_log2B = team_log2b(p->log2N), team_log2F = p->log2N - team_log2B;
  const Tuned tn = tuned(p->N);
  const bool team_ok = p->team_capable && team_log2F >= kTeamMinLog2F && team_log2F <= kTeamMaxLog2F &&
                       p->log2N <= kTeamMaxLog2N && tn.path != 2;
  const int log2F256 = p->log2N - kTeamLog2B;  // the column path's F at B = 256
  const bool cols_ok = log2F256 >= kColsMinLog2F && log2F256 <= kColsMaxLog2F && !cols_disabled();
  const int cols_lb = cols_split(p->log2N, tn), cols_lf = p->log2N - cols_lb;
  for (int i = 0, sgi = 0; i < p->fn; ++i) {
    while (i >= p->sig_row0[sgi] + p->sig_fn[sgi]) ++sgi;
    RowD& r = p->rows[i];
    r.sig = sgi;
    r.n = p->sig_n[sgi];
    r.seg_begin = (int)p->segs.size();
    r.seg_count = (int)bands[i].segs.size();
    for (const Segment& s : bands[i].segs) p->segs.push_back({s.lo, s.hi, s.step, 0});
    int log2M = find2power(std::max(bands[i].band, 1));
    log2M = std::min(log2M, p->log2N);
    // the tuned crossover moves M = 1024 bands to the team kernel (A = 4)
    const int direct_max = tn.direct_max == 9 ? 9 : kMaxLog2B;
    if (log2M > direct_max && log2M > team_log2B && team_ok && r.seg_count <= kTeamSegLds) {
      r.log2B = team_log2B;
      r.log2A = log2M - team_log2B;
      team.push_back(i);
    } else if (log2M > kMaxLog2B && cols_ok && r.seg_count <= kTeamSegLds) {
      r.log2B = cols_lb;
      r.log2A = log2M - cols_lb;
      cols.push_back(i);
    } else {
      r.log2B = std::min(log2M, kMaxLog2B);
      r.log2A = log2M - r.log2B;
      if (r.log2A > 0) wide.push_back(i);
    }
    int log2F = p->log2N - r.log2B;
    r.log2G = std::min(log2F, 14 - r.log2B);
    r.slot = -1;
    r.out_off = (long long)(i - p->sig_row0[sgi]) * r.n;
  }
  if (!team.empty()) {
    // one team per XCD; rows dealt round-robin
    p->team_size = kTeamWgPerXcd;
    p->nteams = 8;
    p->team_log2F = team_log2F;
    // scratch buffers per team: as many N-point buffers as fit the per-team
    // scratch budget, and at least kTeamMinBuffers, unless tuned
    // (fcwt_hip_tune_load); FCWT_HIP_TEAM_BUFFERS overrides both
    p->team_nbuf = team_default_buffers(p->N);
    if (tn.team_buffers) p->team_nbuf = tn.team_buffers;
    if (const char* e = std::getenv("FCWT_HIP_TEAM_BUFFERS")) {
      int v = std::atoi(e);
      if (v >= 2 && v <= kTeamBuffers) p->team_nbuf = v;
    }
    const int nt = p->nteams;
    const int per = ((int)team.size() + nt - 1) / nt;
    p->team_rows.assign((size_t)per * nt, 0);
    std::vector<int> count(nt, 0);
    for (size_t k = 0; k < team.size(); ++k) {
      int t = (int)(k % nt), i = (int)(k / nt);
      p->team_rows[t + (size_t)nt * i] = team[k];
      count[t] = i + 1;
    }
    // item-block sequence per team: step s = pass 1 of row s, then pass 2 of
    // row s - lag; a pass 1 into buffer b waits for the pass 2 of the row
    // nbuf earlier, which sits nbuf - lag steps back
    int lag = std::max(1, p->team_nbuf / 2);
    if (tn.team_lag >= 1 && tn.team_lag < p->team_nbuf) lag = tn.team_lag;
    p->team_lag = lag;
    p->seq_stride = 2 * per;
    p->team_seq.assign((size_t)p->seq_stride * nt, 0);
    p->team_blocks.assign(nt, 0);
    for (int t = 0; t < nt; ++t) {
      int len = 0;
      for (int st = 0; st < count[t] + lag; ++st) {
        if (st < count[t]) p->team_seq[(size_t)t * p->seq_stride + len++] = 2 * st;
        if (st >= lag) p->team_seq[(size_t)t * p->seq_stride + len++] = 2 * (st - lag) + 1;
      }
      p->team_blocks[t] = len;
    }
    p->team_desc.assign(p->team_seq.size(), TeamBlk{});
    for (int t = 0; t < nt; ++t)
      for (int s = 0; s < p->team_blocks[t]; ++s) {
        TeamBlk& d = p->team_desc[(size_t)t * p->seq_stride + s];
        d.sq = p->team_seq[(size_t)t * p->seq_stride + s];
        const RowD& r = p->rows[p->team_rows[t + (size_t)nt * (d.sq >> 1)]];
        d.seg_begin = r.seg_begin;
        d.seg_count = r.seg_count;
        d.log2A = r.log2A;
        d.sig = r.sig;
        d.n = r.n;
        d.out_off = r.out_off;
        for (int k = 0; k < 2 && k < r.seg_count; ++k) d.seg[k] = p->segs[(size_t)r.seg_begin + k];
      }
  }
  if (!cols.empty()) {
    // batches of rows, one N-point scratch slot per row; the slots of one
    // batch are meant to stay in the memory-side cache between the passes
    p->cols_log2F = cols_lf;
    p->cols_log2B = cols_lb;
    const size_t per_row = sizeof(float2) * (size_t)p->N;
    size_t budget = tn.cols_mb ? (size_t)tn.cols_mb << 20 : kColsScratchBytes;
    if (const char* e = std::getenv("FCWT_HIP_COLS_SCRATCH_MB")) {
      long v = std::atol(e);
      if (v >= 1 && v <= 4096) budget = (size_t)v << 20;
    }
    if (p->cols_budget) budget = p->cols_budget;
    p->cols_slots = (int)std::max<size_t>(1, std::min<size_t>(cols.size(), budget / per_row));
    p->cols_rows = cols;
    for (size_t b0 = 0; b0 < cols.size(); b0 += p->cols_slots) {
      const int cnt = (int)std::min<size_t>(cols.size() - b0, p->cols_slots);
      p->cols_batches.push_back({(int)b0, cnt});
      for (int k = 0; k < cnt; ++k) p->rows[cols[b0 + k]].slot = k;
    }
    // with several batches and F <= 2^13 at B = 256, batch k uses scratch
    // set k & 1 and pass 1 of batch k runs in the launch of pass 2 of batch
    // k - 1 (at F = 2^14 a pass-1 work-group fills a CU's LDS, so the passes
    // alternate; at B = 512 they alternate too: fused launches measured
    // 2% faster to 14% slower than separate ones, depending on the GPU)
    p->cols_sets = (p->cols_batches.size() > 1 && cols_lf < 14 && cols_lb == kTeamLog2B) ? 2 : 1;
    if (tn.cols_sets == 1) p->cols_sets = 1;
    if (const char* e = std::getenv("FCWT_HIP_COLS_SETS")) {  // 1: separate passes, 2: fused (several batches)
      if (e[0] == '1') p->cols_sets = 1;
      if (e[0] == '2' && p->cols_batches.size() > 1) p->cols_sets = 2;
    }
  }
  // direct items, grouped by band size (one kernel instance per size); with
  // a team launch, the bands of 2^kFillMinLog2M and up run inside it, after
  // each member's team sequence (TeamArgs::fill_items)
  const bool fill = p->nteams > 0 && fill_rows(team.size(), p->N, p->nsig);
  p->fill_begin = p->fill_count = 0;
  for (int lm = 0; lm <= kMaxLog2B; ++lm) {
    const bool f = fill && lm >= kFillMinLog2M;
    size_t begin = p->items.size();
    if (f && p->fill_count == 0) p->fill_begin = begin;
    int L = 1 << (p->log2N - lm);
    int G = std::min(f ? direct_phases_team(lm) : direct_phases(lm), L);
    for (int i = 0; i < p->fn; ++i) {
      const RowD& r = p->rows[i];
      if (r.log2A != 0 || r.log2B != lm) continue;
      for (int b = 0; b < L / G; ++b) p->items.push_back({i, b});
    }
    if (p->items.size() > begin) p->direct_groups.push_back({lm, begin, p->items.size() - begin, f});
    if (f) {
      p->fill_count = p->items.size() - p->fill_begin;
      p->fill_end[lm - kFillMinLog2M] = (int)p->fill_count;
    }
  }
  p->direct_items = p->items.size();
  if (!wide.empty()) {
    if (p->log2N - kMaxLog2B > 14)
      return fail(FCWT_HIP_EUNSUPPORTED, "padded size N = 2^" + std::to_string(p->log2N) +
                                             " exceeds this build's limit of 2^24 for wide scales");
    size_t per_row = sizeof(float2) * (size_t)p->N;
    p->scratch_slots = (int)std::max<size_t>(1, std::min<size_t>(wide.size(), p->wide_budget / per_row));
    for (size_t b0 = 0; b0 < wide.size(); b0 += p->scratch_slots) {
      size_t b1 = std::min(wide.size(), b0 + p->scratch_slots);
      fcwt_hip_plan_s::Batch bt;
      bt.s_begin = p->items.size();
      for (size_t k = b0; k < b1; ++k) {
        RowD& r = p->rows[wide[k]];
        r.slot = (int)(k - b0);
        int log2F = p->log2N - r.log2B;
        int log2Q = std::min(r.log2B, 14 - log2F);
        int nblk = 1 << (r.log2B - log2Q);
        for (int b = 0; b < nblk; ++b) p->items.push_back({wide[k], b});
      }
      bt.s_count = p->items.size() - bt.s_begin;
      bt.w_begin = p->items.size();
      for (size_t k = b0; k < b1; ++k) {
        const RowD& r = p->rows[wide[k]];
        int nblk = 1 << (p->log2N - r.log2B - r.log2G);
        for (int b = 0; b < nblk; ++b) p->items.push_back({wide[k], b});
      }
      bt.w_count = p->items.size() - bt.w_begin;
      p->batches.push_back(bt);
    }
  }
  return FCWT_HIP_OK;
}

static int check_device(int* dev) {
  HIP_TRY(hipGetDevice(dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, *dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(FCWT_HIP_EDEVICE, std::string("libfcwt_hip is built for gfx950; current device is ") + prop.gcnArchName);
  return FCWT_HIP_OK;
}

// The team path launches kTeamWgPerCu work-groups per CU (an ordinary
// launch) and makes one team of the work-groups each XCD receives, whose
// scratch stays in that XCD's L2: it is meant for the full 256-CU device in
// one partition.  FCWT_HIP_NO_TEAMS=1 in the environment turns it off.
static const void* team_kernel(int log2N) {
  switch (log2N) {
    case 15: return (const void*)k_rows_team<7, 8>;
    case 16: return (const void*)k_rows_team<8, 8>;
    case 17: return (const void*)k_rows_team<9, 8>;
    case 18: return (const void*)k_rows_team<9, 9>;
    default: return nullptr;
  }
}

static bool team_capable(int dev, int log2N) {
  const char* env = std::getenv("FCWT_HIP_NO_TEAMS");
  if (env && env[0] == '1') return false;
  const void* kern = team_kernel(log2N);
  if (!kern) return false;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
  if (prop.multiProcessorCount != 256) return false;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kTeamThreads, 0) != hipSuccess) return false;
  return per_cu >= kTeamWgPerCu;
}

__global__ void k_fill_u32(unsigned* p, size_t n, unsigned v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

#define TRY(expr)                 \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != FCWT_HIP_OK) {     \
      free_plan(p);               \
      return rc_;                 \
    }                             \
  } while (0)

template <int LOG2M>
static void launch_direct_t(unsigned blocks, const KArgs& a, hipStream_t st) {
  k_rows_direct<LOG2M><<<blocks, DirectCfg<LOG2M>::NT, 0, st>>>(a);
}

static void launch_direct(int log2M, unsigned blocks, const KArgs& a, hipStream_t st) {
  switch (log2M) {
    case 0: launch_direct_t<0>(blocks, a, st); break;
    case 1: launch_direct_t<1>(blocks, a, st); break;
    case 2: launch_direct_t<2>(blocks, a, st); break;
    case 3: launch_direct_t<3>(blocks, a, st); break;
    case 4: launch_direct_t<4>(blocks, a, st); break;
    case 5: launch_direct_t<5>(blocks, a, st); break;
    case 6: launch_direct_t<6>(blocks, a, st); break;
    case 7: launch_direct_t<7>(blocks, a, st); break;
    case 8: launch_direct_t<8>(blocks, a, st); break;
    case 9: launch_direct_t<9>(blocks, a, st); break;
    default: launch_direct_t<10>(blocks, a, st); break;
  }
}

// Forward transforms of `batch` complex signals of f.n floats each (strides
// in f), keeping f.keep outputs: one work-group per signal up to 2^14 points
// (with derived twiddles: these are the adjoint's transforms), the two
// four-step kernels beyond.
static void launch_fwd_complex(FwdArgs f, int batch, hipStream_t st) {
  f.complex_in = 1;
  f.log2H1 = f.log2H / 2;
  f.log2H2 = f.log2H - f.log2H1;
  if (f.log2H <= 14) {
    launch_rows_fft(f, (unsigned)batch, st, true);
  } else {
    int ga = 1 << (f.log2H2 - fwd_log2g(f.log2H1, f.log2H2));
    int gb = 1 << (f.log2H1 - fwd_log2g(f.log2H2, f.log2H1));
    k_fwd_a<<<dim3(ga, batch), kBlock, 0, st>>>(f);
    k_fwd_b<<<dim3(gb, batch), kBlock, 0, st>>>(f);
  }
}

constexpr size_t kAdjScratchBytes = 256u << 20;  // adjoint row-spectrum workspace

static int adjoint_workspace(fcwt_hip_plan p, int signals) {
  const size_t N = (size_t)p->N;
  if (!p->d_adj_P) {
    const size_t per_row = sizeof(float2) * (N / 2 + (p->log2N > 14 ? N : 0));
    const size_t rows = (size_t)p->fn * std::max(1, signals);
    p->adj_chunk = (int)std::max<size_t>(1, std::min<size_t>(rows, kAdjScratchBytes / per_row));
    p->adj_chunk = std::min(p->adj_chunk, 65535);  // grid dimension y of the row transforms
    int rc = dalloc(p, &p->d_adj_P, (size_t)p->adj_chunk * (N / 2));
    if (rc == FCWT_HIP_OK && p->log2N > 14) rc = dalloc(p, &p->d_adj_atmp, (size_t)p->adj_chunk * N);
    if (rc != FCWT_HIP_OK) return rc;
  }
  if (signals > p->adj_signals) {
    if (p->d_adj_A) (void)hipFree(p->d_adj_A);
    if (p->d_adj_y) (void)hipFree(p->d_adj_y);
    p->d_adj_A = nullptr;
    p->d_adj_y = nullptr;
    p->adj_signals = 0;
    int rc = dalloc(p, &p->d_adj_A, (size_t)signals * (N / 2));
    if (rc == FCWT_HIP_OK) rc = dalloc(p, &p->d_adj_y, (size_t)signals * N);
    if (rc != FCWT_HIP_OK) return rc;
    p->adj_signals = signals;
  }
  return FCWT_HIP_OK;
}

template <int LOG2H>
static void launch_adj_rows_t(dim3 grid, const float2* g, int n, int fn, int groups, fcwt_hip_plan p, float2* part,
                              hipStream_t st) {
  k_adj_rows<LOG2H><<<grid, kRowFftThreads, 0, st>>>(g, n, fn, groups, p->d_rows, p->d_segs, p->d_mother, part);
}

static void launch_adj_rows(int log2H, dim3 grid, const float2* g, int n, int fn, int groups, fcwt_hip_plan p,
                            float2* part, hipStream_t st) {
  switch (log2H) {
    case 1: launch_adj_rows_t<1>(grid, g, n, fn, groups, p, part, st); break;
    case 2: launch_adj_rows_t<2>(grid, g, n, fn, groups, p, part, st); break;
    case 3: launch_adj_rows_t<3>(grid, g, n, fn, groups, p, part, st); break;
    case 4: launch_adj_rows_t<4>(grid, g, n, fn, groups, p, part, st); break;
    case 5: launch_adj_rows_t<5>(grid, g, n, fn, groups, p, part, st); break;
    case 6: launch_adj_rows_t<6>(grid, g, n, fn, groups, p, part, st); break;
    case 7: launch_adj_rows_t<7>(grid, g, n, fn, groups, p, part, st); break;
    case 8: launch_adj_rows_t<8>(grid, g, n, fn, groups, p, part, st); break;
    case 9: launch_adj_rows_t<9>(grid, g, n, fn, groups, p, part, st); break;
    case 10: launch_adj_rows_t<10>(grid, g, n, fn, groups, p, part, st); break;
    case 11: launch_adj_rows_t<11>(grid, g, n, fn, groups, p, part, st); break;
    case 12: launch_adj_rows_t<12>(grid, g, n, fn, groups, p, part, st); break;
    case 13: launch_adj_rows_t<13>(grid, g, n, fn, groups, p, part, st); break;
    default: launch_adj_rows_t<14>(grid, g, n, fn, groups, p, part, st); break;
  }
}

// Input gradients of `signals` signals from their output gradients d_grad
// (signals x fn rows of n complex64, contiguous): Ghat_r = FFT_N(G_r)[0, N/2)
// and A_s = sum over s's rows of w_i Ghat_r (at N <= 2^14 fused in
// k_adj_rows, beyond that per chunk of rows through the workspace); then
// xbar_s = c Re IFFT_N(A_s)[0, n) (the complex value for complex input), n
// apart.
static int launch_adjoint(fcwt_hip_plan p, const float* d_grad, int signals, bool complex_in, float* d_xbar,
                          hipStream_t st) {
  const int n = p->n, N = p->N, fn = p->fn;
  const float c = p->normalize ? 1.0f / (float)N : 1.0f;
  const size_t xbar_words = (size_t)(complex_in ? 2 : 1) * n * signals;
  if (N < 2) {  // no spectrum bin: the gradient is zero
    k_fill_u32<<<(unsigned)((xbar_words + 255) / 256), 256, 0, st>>>((unsigned*)d_xbar, xbar_words, 0u);
    HIP_TRY(hipGetLastError());
    return FCWT_HIP_OK;
  }
  int rc = adjoint_workspace(p, signals);
  if (rc != FCWT_HIP_OK) return rc;
  FwdArgs f;
  f.xs = nullptr;
  f.ns = nullptr;
  f.len_mul = 1;
  f.n = 2 * n;
  f.x_stride = 2LL * n;
  f.spec_stride = N / 2;
  f.keep = N / 2;
  f.ztmp = nullptr;
  f.atmp = p->d_adj_atmp;
  f.log2H = p->log2N;
  f.tw_shift = 0;
  f.tw = twiddles_of(p);
  ProfScope ps(p, 5, st);
  const long long rows = (long long)fn * signals;
  const int groups = (fn + kAdjRows - 1) / kAdjRows;
  const bool fused = p->log2N <= 14 && groups <= p->adj_chunk;
  if (fused) {
    // fused: row transforms, windows and the sum over rows in one kernel,
    // partial sums per group of rows in the workspace, signals in chunks
    const int per = std::max(1, std::min(p->adj_chunk / groups, 65535));
    for (int s0 = 0; s0 < signals; s0 += per) {
      const int cnt = std::min(per, signals - s0);
      const float2* g0 = (const float2*)d_grad + (size_t)s0 * fn * n;
      launch_adj_rows(p->log2N, dim3(groups, cnt), g0, n, fn, groups, p, p->d_adj_P, st);
      k_adj_combine<<<dim3((N / 2 + 255) / 256, cnt), 256, 0, st>>>(p->d_adj_P, groups, N / 2,
                                                                    p->d_adj_A + (size_t)s0 * (N / 2));
    }
  }
  // otherwise, per chunk of rows: their spectra into the workspace, then
  // the windowed sum over rows
  for (long long row0 = 0; row0 < rows && !fused; row0 += p->adj_chunk) {
    const int count = (int)std::min<long long>(p->adj_chunk, rows - row0);
    f.x = d_grad + (size_t)row0 * 2 * n;
    f.spec = p->d_adj_P;
    launch_fwd_complex(f, count, st);
    k_adj_reduce<<<dim3((N / 2 + 255) / 256, signals), 256, 0, st>>>(p->d_adj_P, row0, count, fn, p->d_rows,
                                                                      p->d_segs, p->d_mother, N, p->d_adj_A);
  }
  // y_s = FFT_N(conj A_s), the first n outputs
  f.x = (const float*)p->d_adj_A;
  f.n = N;  // N/2 complex values
  f.x_stride = N;
  f.spec = p->d_adj_y;
  f.spec_stride = N;
  f.keep = n;
  f.atmp = p->d_adj_atmp;
  for (int s0 = 0; s0 < signals; s0 += p->adj_chunk) {
    const int cnt = std::min(p->adj_chunk, signals - s0);
    FwdArgs g = f;
    g.x = f.x + (size_t)s0 * N;
    g.spec = f.spec + (size_t)s0 * N;
    launch_fwd_complex(g, cnt, st);
  }
  k_adj_out<<<dim3((n + 255) / 256, signals), 256, 0, st>>>(p->d_adj_y, n, N, c, complex_in ? 1 : 0, d_xbar);
  HIP_TRY(hipGetLastError());
  return FCWT_HIP_OK;
}

// Uploads the per-signal pointer arrays of a several-signal plan when they
// changed since the last call (stream-ordered, so earlier launches keep
// reading the arrays they were given).
static int upload_pointers(fcwt_hip_plan p, const float* const* xs, float2* const* outs, hipStream_t st) {
  std::vector<const void*> nx(xs, xs + p->nsig), no(outs, outs + p->nsig);
  if (nx != p->up_xs) {
    p->up_xs = nx;
    HIP_TRY(hipMemcpyAsync(p->d_xs, p->up_xs.data(), sizeof(void*) * p->nsig, hipMemcpyHostToDevice, st));
  }
  if (no != p->up_outs) {
    p->up_outs = no;
    HIP_TRY(hipMemcpyAsync(p->d_outs, p->up_outs.data(), sizeof(void*) * p->nsig, hipMemcpyHostToDevice, st));
  }
  return FCWT_HIP_OK;
}

// Every kernel of one execute: xs[s] / outs[s] are signal s's device input
// and output (arrays of nsig pointers).
// The forward transform of every signal into spec (N/2 values per signal).
static void launch_forward(fcwt_hip_plan p, const float* const* xs, bool complex_in, float2* spec, hipStream_t st) {
  const bool multi = p->nsig > 1;
  FwdArgs f;
  f.x = xs[0];
  f.n = complex_in ? 2 * p->n : p->n;
  f.xs = multi ? p->d_xs : nullptr;
  f.ns = multi ? p->d_sig_n : nullptr;
  f.len_mul = complex_in ? 2 : 1;
  f.spec = spec;
  f.ztmp = p->d_ztmp;
  f.atmp = p->d_atmp;
  f.log2H = complex_in ? p->log2N : p->log2N - 1;
  f.log2H1 = f.log2H / 2;
  f.log2H2 = f.log2H - f.log2H1;
  f.tw_shift = p->log2N - f.log2H;
  f.complex_in = complex_in ? 1 : 0;
  f.x_stride = 0;
  f.spec_stride = p->N / 2;
  f.keep = p->N / 2;
  f.tw = twiddles_of(p);
  const unsigned ns = (unsigned)p->nsig;
  if (complex_in && p->N == 1) {
    // N = 1: X[0, N/2) is empty; nothing reads the spectrum
  } else if (p->N >= 2) {
    ProfScope ps(p, 0, st);
    if (f.log2H <= 14) {
      launch_rows_fft(f, ns, st);
    } else {
      int ga = 1 << (f.log2H2 - fwd_log2g(f.log2H1, f.log2H2));
      int gb = 1 << (f.log2H1 - fwd_log2g(f.log2H2, f.log2H1));
      k_fwd_a<<<dim3(ga, ns), kBlock, 0, st>>>(f);
      k_fwd_b<<<dim3(gb, ns), kBlock, 0, st>>>(f);
      if (!complex_in) {
        int H = 1 << f.log2H;
        k_fwd_post<<<dim3((H + 255) / 256, ns), 256, 0, st>>>(f);
      }
    }
  }
}

// One column-path launch for N = 2^LB x 2^LF: pass 1 of n1 rows (b1) beside
// pass 2 of n2 rows (b2); pass 2 alone (n1 = 0) runs as work-groups of one
// item (256 threads at B = 256, 512 at B = 512).
template <int LF, int LB = kTeamLog2B>
static void launch_cols_t(const BigArgs& b1, int n1, const BigArgs& b2, int n2, hipStream_t st) {
  using C = ColsCfg<LF>;
  using G = ColsGrid<LF, LB>;
  if (n1 == 0) {
    if constexpr (LB == 9)
      k_cols_out512<LF><<<dim3((1 << LF) / 16, n2), 512, 0, st>>>(b2);
    else
      k_cols_out<LF><<<dim3((1 << LF) / 16, n2), 256, 0, st>>>(b2);
    return;
  }
  const unsigned wgs = (unsigned)(G::X1 * n1 + G::X2 * n2);
  k_cols_step<LF, LB><<<wgs, C::T, 0, st>>>(b1, n1, b2, n2);
}

static void launch_cols(int log2F, int log2B, const BigArgs& b1, int n1, const BigArgs& b2, int n2,
                        hipStream_t st) {
  if (log2B == 9) {  // only N = 2^22 splits this way (cols_split)
    launch_cols_t<13, 9>(b1, n1, b2, n2, st);
    return;
  }
  switch (log2F) {
    case 5: launch_cols_t<5>(b1, n1, b2, n2, st); break;
    case 6: launch_cols_t<6>(b1, n1, b2, n2, st); break;
    case 7: launch_cols_t<7>(b1, n1, b2, n2, st); break;
    case 8: launch_cols_t<8>(b1, n1, b2, n2, st); break;
    case 9: launch_cols_t<9>(b1, n1, b2, n2, st); break;
    case 10: launch_cols_t<10>(b1, n1, b2, n2, st); break;
    case 11: launch_cols_t<11>(b1, n1, b2, n2, st); break;
    case 12: launch_cols_t<12>(b1, n1, b2, n2, st); break;
    case 13: launch_cols_t<13>(b1, n1, b2, n2, st); break;
    default: launch_cols_t<14>(b1, n1, b2, n2, st); break;
  }
}

// Polls before a team wait is declared failed: kTeamSpin, or fewer from
// FCWT_HIP_TEAM_SPIN (0: every wait fails at once), read at every launch so
// that a test can force one failed launch and then run the same plan
// normally.
static long team_spin() {
  if (const char* e = std::getenv("FCWT_HIP_TEAM_SPIN")) {
    const long v = std::atol(e);
    if (v >= 0) return std::min(v, kTeamSpin);
  }
  return kTeamSpin;
}

// The team kernel's arguments for plan p (its sync words at `sync`).
static TeamArgs team_args_of(fcwt_hip_plan p, const KArgs& a, unsigned* sync) {
  TeamArgs ta;
  ta.k = a;
  ta.team_desc = p->d_team_desc;
  ta.team_blocks = p->d_team_blocks;
  ta.seq_stride = p->seq_stride;
  ta.nteams = p->nteams;
  ta.log2F = p->team_log2F;
  ta.nbuf = p->team_nbuf;
  ta.spin = team_spin();
  ta.scratch = p->d_team_scratch;
  ta.sync = sync;
  ta.error = p->d_error;
  ta.fill_items = p->d_items + p->fill_begin;
  ta.nfill = (int)p->fill_count;
  for (int g = 0; g < kFillGroups; ++g) ta.fill_end[g] = p->fill_end[g];
  return ta;
}

// Every launch of one execute of plan p on st.  With `defer_team` non-null
// the team launch is not made: its arguments go there, for a launch that
// takes the team rows of several plans (k_rows_team_multi).
static int launch_all(fcwt_hip_plan p, const float* const* xs, bool complex_in, float2* const* outs,
                      hipStream_t st, TeamArgs* defer_team = nullptr) {
  const bool multi = p->nsig > 1;
  if (multi) {
    int rc = upload_pointers(p, xs, outs, st);
    if (rc != FCWT_HIP_OK) return rc;
  }
  launch_forward(p, xs, complex_in, p->d_spec, st);
  KArgs a;
  a.spec = p->d_spec;
  a.mother = p->d_mother;
  a.segs = p->d_segs;
  a.rows = p->d_rows;
  a.scratch = p->d_scratch;
  a.out = outs[0];
  a.outs = multi ? p->d_outs : nullptr;
  a.N = p->N;
  a.log2N = p->log2N;
  a.norm = p->normalize ? 1.0f / (float)p->N : 1.0f;
  a.tw = twiddles_of(p);
  if (p->direct_items > p->fill_count) {
    ProfScope ps(p, 1, st);
    for (const auto& gr : p->direct_groups) {
      if (gr.fill) continue;
      a.items = p->d_items + gr.begin;
      launch_direct(gr.log2M, (unsigned)gr.count, a, st);
    }
  }
  if (p->nteams && defer_team) {
    *defer_team = team_args_of(p, a, nullptr);
  } else if (p->nteams) {
    ProfScope ps(p, 4, st);
    size_t words = p->team_sync_words;
    k_fill_u32<<<(unsigned)((words + 255) / 256), 256, 0, st>>>(p->d_team_sync, words, 0u);
    TeamArgs ta = team_args_of(p, a, p->d_team_sync);
    void* args[] = {&ta};
    HIP_TRY(hipLaunchKernel(team_kernel(p->log2N), dim3(256 * kTeamWgPerCu), dim3(kTeamThreads), args, 0, st));
  }
  if (!p->cols_batches.empty()) {
    const int lf = p->cols_log2F, lb = p->cols_log2B;
    BigArgs b;
    b.k = a;
    b.xt = p->d_xt;
    b.scratch = p->d_cols_scratch;
    b.fb = p->fb;
    b.toradians = p->m_toradians;
    b.mnorm = p->m_norm;
    {
      ProfScope ps(p, 7, st);
      k_spec_transpose<<<dim3((1 << (lb - 4)) * std::max(1, (1 << lf) / 128), (unsigned)p->nsig), 256, 0, st>>>(
          p->d_spec, p->d_xt, lf, lb);
    }
    const size_t set_elems = (size_t)p->cols_slots * p->N;
    const size_t nb = p->cols_batches.size();
    auto batch_args = [&](size_t k) {
      BigArgs x = b;
      x.rows = p->d_cols_rows + p->cols_batches[k].begin;
      x.scratch = p->d_cols_scratch + (k & 1) * (p->cols_sets - 1) * set_elems;
      return x;
    };
    ProfScope ps(p, 6, st);
    if (p->cols_sets == 2) {
      // launch k: pass 1 of batch k (scratch set k & 1) beside pass 2 of batch k - 1
      for (size_t k = 0; k <= nb; ++k) {
        const BigArgs b1 = k < nb ? batch_args(k) : b, b2 = k > 0 ? batch_args(k - 1) : b;
        launch_cols(lf, lb, b1, k < nb ? p->cols_batches[k].count : 0, b2, k > 0 ? p->cols_batches[k - 1].count : 0,
                    st);
      }
    } else {
      for (size_t k = 0; k < nb; ++k) {
        const BigArgs bk = batch_args(k);
        launch_cols(lf, lb, bk, p->cols_batches[k].count, b, 0, st);
        launch_cols(lf, lb, b, 0, bk, p->cols_batches[k].count, st);
      }
    }
  }
  for (const auto& bt : p->batches) {
    {
      ProfScope ps(p, 2, st);
      a.items = p->d_items + bt.s_begin;
      k_spreader<<<(unsigned)bt.s_count, kBlock, 0, st>>>(a);
    }
    ProfScope ps(p, 3, st);
    a.items = p->d_items + bt.w_begin;
    k_writer<<<(unsigned)bt.w_count, kBlock, 0, st>>>(a);
  }
  HIP_TRY(hipGetLastError());
  return FCWT_HIP_OK;
}

// ------------------------------------------------------------------- C ABI
extern "C" {

const char* fcwt_hip_last_error(void) { return g_last_error.c_str(); }

const char* fcwt_hip_version(void) { return "fcwt_hip 0.1.0 (gfx950)"; }

int fcwt_hip_find2power(int n) { return find2power(n); }

int fcwt_hip_scales(int scaletype, int fs, float f0, float f1, int fn, float* out) {
  if (!out || fn < 1) return fail(FCWT_HIP_EINVAL, "fcwt_hip_scales: need fn >= 1 and an output buffer");
  if (scaletype < 0 || scaletype > 2) return fail(FCWT_HIP_EINVAL, "fcwt_hip_scales: unknown scale type");
  scales(scaletype, fs, f0, f1, fn, out);
  return FCWT_HIP_OK;
}

int fcwt_hip_frequencies(const float* sc, int fn, int fs, float* out) {
  if (!sc || !out || fn < 0) return fail(FCWT_HIP_EINVAL, "fcwt_hip_frequencies: bad arguments");
  frequencies(sc, fn, fs, out);
  return FCWT_HIP_OK;
}

int fcwt_hip_morlet_generate(float fb, int size, float* mother) {
  if (!mother || size < 0) return fail(FCWT_HIP_EINVAL, "fcwt_hip_morlet_generate: bad arguments");
  morlet_mother(fb, size, mother);
  return FCWT_HIP_OK;
}

int fcwt_hip_morlet_support(float fb, float scale) { return morlet_support(fb, scale); }

int fcwt_hip_morlet_generate_time(float fb, float scale, int size, float* real, float* imag) {
  if (!real || !imag) return fail(FCWT_HIP_EINVAL, "fcwt_hip_morlet_generate_time: bad arguments");
  morlet_time(fb, scale, size, real, imag);
  return FCWT_HIP_OK;
}

int fcwt_hip_morlet_wavelet(float fb, float scale, float* out, int pn) {
  if (!out || pn < 0) return fail(FCWT_HIP_EINVAL, "fcwt_hip_morlet_wavelet: bad arguments");
  morlet_wavelet(fb, scale, out, pn);
  return FCWT_HIP_OK;
}

int fcwt_hip_describe_rows(int n, const float* sc, int fn, float fb, int threads, unsigned flags, int* covered,
                           int* band, int* segments) {
  if (n < 1 || fn < 1 || !sc || threads < 1 || find2power(n) > 28)
    return fail(FCWT_HIP_EINVAL, "fcwt_hip_describe_rows: bad arguments");
  int N = 1 << find2power(n);
  std::vector<float> mother(N);
  morlet_mother(fb, N, mother.data());
  std::vector<RowBand> b = build_bands(N, sc, fn, threads, !(flags & FCWT_HIP_COVERAGE_SCALAR), mother.data(),
                                       kStaleFloor);
  for (int i = 0; i < fn; ++i) {
    if (covered) covered[i] = b[i].C;
    if (band) band[i] = b[i].band;
    if (segments) segments[i] = (int)b[i].segs.size();
  }
  return FCWT_HIP_OK;
}

}  // extern "C"

// A plan for nsig signals that all pad to the same N: signal s has
// lengths[s] samples and fns[s] scales at scales[s].  One-signal plans are
// the public fcwt_hip_plan_create; batches hold one plan per padded size.
static int create_plan(fcwt_hip_plan* out, int nsig, const int* lengths, const float* const* sc, const int* fns,
                       float fb, int threads, int normalize, unsigned flags, const char* what,
                       char* arena = nullptr, size_t arena_cap = 0, hipStream_t stream = nullptr) {
  if (!out) return fail(FCWT_HIP_EINVAL, std::string(what) + ": null plan pointer");
  *out = nullptr;
  if (nsig < 1 || !lengths || !sc || !fns) return fail(FCWT_HIP_EINVAL, std::string(what) + ": null argument");
  if (threads < 1) return fail(FCWT_HIP_EINVAL, std::string(what) + ": threads must be >= 1");
  if (!std::isfinite(fb)) return fail(FCWT_HIP_EINVAL, std::string(what) + ": wavelet bandwidth must be finite");
  if (flags & ~FCWT_HIP_COVERAGE_SCALAR) return fail(FCWT_HIP_EINVAL, std::string(what) + ": unknown flags");
  long long rows = 0;
  for (int s = 0; s < nsig; ++s) {
    if (lengths[s] < 1) return fail(FCWT_HIP_EINVAL, std::string(what) + ": signal length must be >= 1");
    if (fns[s] < 1 || !sc[s]) return fail(FCWT_HIP_EINVAL, std::string(what) + ": need fn >= 1 scales");
    if (find2power(lengths[s]) != find2power(lengths[0]))
      return fail(FCWT_HIP_EINVAL, std::string(what) + ": signals of one plan must pad to the same size");
    for (int i = 0; i < fns[s]; ++i)
      if (!(sc[s][i] > 0.f) || !std::isfinite(sc[s][i]))
        return fail(FCWT_HIP_EINVAL, std::string(what) + ": scales must be finite and > 0");
    rows += fns[s];
  }
  if (rows > (1LL << 30)) return fail(FCWT_HIP_EUNSUPPORTED, std::string(what) + ": too many rows");
  int log2N = find2power(lengths[0]);
  if (log2N > 28) return fail(FCWT_HIP_EUNSUPPORTED, std::string(what) + ": signal too long");
  int dev = 0;
  int rc = check_device(&dev);
  if (rc != FCWT_HIP_OK) return rc;

  fcwt_hip_plan p = new (std::nothrow) fcwt_hip_plan_s;
  if (!p) return fail(FCWT_HIP_ENOMEM, std::string(what) + ": out of host memory");
  p->device = dev;
  p->nsig = nsig;
  p->n = lengths[0];
  p->log2N = log2N;
  p->N = 1 << log2N;
  p->fn = (int)rows;
  p->fb = fb;
  p->threads = threads;
  p->normalize = normalize != 0;
  p->flags = flags;
  p->team_capable = team_capable(dev, log2N);
  if (arena) {  // everything in the static array, on the caller's stream
    p->arena = arena;
    p->arena_cap = arena_cap;
    p->own_stream = false;
    p->stream = stream;
    p->wide_budget = kNoAllocScratchBudget;
    p->cols_budget = kNoAllocScratchBudget;
  }
  for (int s = 0; s < nsig; ++s) {
    p->sig_row0.push_back((int)p->scales.size());
    p->sig_n.push_back(lengths[s]);
    p->sig_fn.push_back(fns[s]);
    p->scales.insert(p->scales.end(), sc[s], sc[s] + fns[s]);
  }

  // host copy of the mother table, used only to decide which stale bins matter
  std::vector<float> mother_host(p->N);
  morlet_mother(fb, p->N, mother_host.data());
  TRY(build_schedule(p, mother_host));

  if (!p->arena && hipStreamCreate(&p->stream) != hipSuccess) {
    free_plan(p);
    return fail(FCWT_HIP_EDEVICE, std::string(what) + ": hipStreamCreate failed");
  }
  const int N = p->N;
  p->s1bits = (log2N + 1) / 2;
  int n1 = 1 << p->s1bits, n2 = N >> p->s1bits;
  TRY(dalloc(p, &p->d_mother, N));
  TRY(dalloc(p, &p->d_tw1, n1));
  TRY(dalloc(p, &p->d_tw2, n2));
  TRY(dalloc(p, &p->d_spec, (size_t)nsig * std::max(1, N / 2)));
  // four-step temporaries: N/2 points for real input, N for complex input
  if (log2N > 14) {
    TRY(dalloc(p, &p->d_ztmp, (size_t)nsig * N / 2));
    TRY(dalloc(p, &p->d_atmp, (size_t)nsig * N));
  }
  if (nsig > 1) {
    TRY(dalloc(p, &p->d_sig_n, (size_t)nsig));
    TRY(dalloc(p, &p->d_xs, (size_t)nsig));
    TRY(dalloc(p, &p->d_outs, (size_t)nsig));
  }
  TRY(dalloc(p, &p->d_segs, p->segs.size()));
  TRY(dalloc(p, &p->d_rows, p->rows.size()));
  TRY(dalloc(p, &p->d_items, p->items.size()));
  if (p->scratch_slots) TRY(dalloc(p, &p->d_scratch, (size_t)p->scratch_slots * N));
  if (p->cols_slots) {
    TRY(dalloc(p, &p->d_cols_rows, p->cols_rows.size()));
    TRY(dalloc(p, &p->d_cols_scratch, (size_t)p->cols_sets * p->cols_slots * N));
    TRY(dalloc(p, &p->d_xt, (size_t)nsig * (N / 2)));
  }
  TRY(dalloc(p, &p->d_error, 1));
  if (p->nteams) {
    TRY(dalloc(p, &p->d_team_desc, p->team_desc.size()));
    TRY(dalloc(p, &p->d_team_blocks, p->team_blocks.size()));
    TRY(dalloc(p, &p->d_team_scratch, (size_t)p->nteams * p->team_nbuf * N));
    p->team_sync_words = (size_t)64 * (1 + p->nteams * kTeamBuffers + p->nteams);
    TRY(dalloc(p, &p->d_team_sync, p->team_sync_words));
  }

  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return FCWT_HIP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->stream));
    return FCWT_HIP_OK;
  };
  TRY(up(p->d_segs, p->segs.data(), p->segs.size() * sizeof(SegD)));
  TRY(up(p->d_rows, p->rows.data(), p->rows.size() * sizeof(RowD)));
  TRY(up(p->d_items, p->items.data(), p->items.size() * sizeof(Item)));
  if (nsig > 1) TRY(up(p->d_sig_n, p->sig_n.data(), sizeof(int) * nsig));
  if (p->cols_slots) TRY(up(p->d_cols_rows, p->cols_rows.data(), p->cols_rows.size() * sizeof(int)));
  if (p->nteams) {
    TRY(up(p->d_team_desc, p->team_desc.data(), p->team_desc.size() * sizeof(TeamBlk)));
    TRY(up(p->d_team_blocks, p->team_blocks.data(), p->team_blocks.size() * sizeof(int)));
  }
  k_fill_u32<<<1, 64, 0, p->stream>>>(p->d_error, 1, 0u);

  // mother-table scalars in float32 exactly as fcwt.cpp:56-57 forms them
  const float kPi = 3.14159265358979323846264338327950288f;
  float toradians = (2 * kPi) / (float)N;
  float norm = sqrtf(2 * kPi) * 0.75112554446f;
  p->m_toradians = toradians;
  p->m_norm = norm;
  k_mother<<<(N + 255) / 256, 256, 0, p->stream>>>(p->d_mother, N, fb, toradians, norm);
  int nt = std::max(n1, n2);
  k_twiddles<<<(nt + 255) / 256, 256, 0, p->stream>>>(p->d_tw1, n1, p->d_tw2, n2, p->s1bits, log2N);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess) {
    free_plan(p);
    return fail(FCWT_HIP_EDEVICE, std::string(what) + ": table kernels failed");
  }
  *out = p;
  return FCWT_HIP_OK;
}

extern "C" {

int fcwt_hip_plan_create(fcwt_hip_plan* out, int n, const float* sc, int fn, float fb, int threads, int normalize,
                         unsigned flags) {
  return create_plan(out, 1, &n, &sc, &fn, fb, threads, normalize, flags, "fcwt_hip_plan_create");
}

int fcwt_hip_plan_destroy(fcwt_hip_plan p) {
  free_plan(p);
  return FCWT_HIP_OK;
}

int fcwt_hip_plan_shape(fcwt_hip_plan p, int* n, int* N, int* fn) {
  if (!p) return fail(FCWT_HIP_EINVAL, "null plan");
  if (n) *n = p->n;
  if (N) *N = p->N;
  if (fn) *fn = p->fn;
  return FCWT_HIP_OK;
}

static int read_error_word(fcwt_hip_plan p) {
  unsigned h = 0;
  HIP_TRY(hipMemcpy(&h, p->d_error, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (h != 0) {
    unsigned zero = 0;
    HIP_TRY(hipMemcpy(p->d_error, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
    return fail(FCWT_HIP_EDEVICE, h & 2u ? "wide-band kernel: an XCD's team left rows unclaimed"
                                         : "wide-band kernel: a team wait timed out");
  }
  return FCWT_HIP_OK;
}

int fcwt_hip_plan_check(fcwt_hip_plan p) {
  if (!p) return fail(FCWT_HIP_EINVAL, "null plan");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  hipError_t e = hipDeviceSynchronize();
  int rc = e == hipSuccess ? read_error_word(p)
                           : fail(FCWT_HIP_EDEVICE, std::string("device error: ") + hipGetErrorString(e));
  if (cur != p->device) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_plan_profile(fcwt_hip_plan p, int enable) {
  if (!p) return fail(FCWT_HIP_EINVAL, "null plan");
  drain_profile(p);
  p->profiling = enable != 0;
  for (int k = 0; k < kKinds; ++k) {
    p->prof_ms[k] = 0;
    p->prof_launches[k] = 0;
  }
  return FCWT_HIP_OK;
}

int fcwt_hip_plan_profile_read(fcwt_hip_plan p, int max, const char** names, double* ms, long long* launches,
                               int* count) {
  if (!p || max < 0) return fail(FCWT_HIP_EINVAL, "fcwt_hip_plan_profile_read: bad arguments");
  drain_profile(p);
  int k = 0;
  for (int i = 0; i < kKinds; ++i) {
    if (p->prof_launches[i] == 0) continue;
    if (k < max) {
      if (names) names[k] = kKindNames[i];
      if (ms) ms[k] = p->prof_ms[i];
      if (launches) launches[k] = p->prof_launches[i];
    }
    ++k;
  }
  if (count) *count = k;
  return FCWT_HIP_OK;
}

int fcwt_hip_plan_row_split(fcwt_hip_plan p, int* direct_rows, int* team_rows, int* two_pass_rows) {
  if (!p) return fail(FCWT_HIP_EINVAL, "null plan");
  int d = 0, t = 0, w = 0;
  for (const RowD& r : p->rows) {
    if (r.log2A == 0) ++d;
    else if (r.slot < 0) ++t;  // team rows hold no scratch slot
    else ++w;
  }
  if (direct_rows) *direct_rows = d;
  if (team_rows) *team_rows = t;
  if (two_pass_rows) *two_pass_rows = w;
  return FCWT_HIP_OK;
}

int fcwt_hip_plan_device_bytes(fcwt_hip_plan p, size_t* bytes) {
  if (!p || !bytes) return fail(FCWT_HIP_EINVAL, "null argument");
  *bytes = p->bytes;
  return FCWT_HIP_OK;
}

static int run_device(fcwt_hip_plan p, const float* d_input, bool complex_in, float* d_output, void* stream,
                      const char* what) {
  if (!p || !d_input || !d_output) return fail(FCWT_HIP_EINVAL, std::string(what) + ": null argument");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  float2* o = (float2*)d_output;
  int rc = launch_all(p, &d_input, complex_in, &o, (hipStream_t)stream);
  if (cur != p->device) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_cwt_device(fcwt_hip_plan p, const float* d_input, float* d_output, void* stream) {
  return run_device(p, d_input, false, d_output, stream, "fcwt_hip_cwt_device");
}

int fcwt_hip_ccwt_device(fcwt_hip_plan p, const float* d_input, float* d_output, void* stream) {
  return run_device(p, d_input, true, d_output, stream, "fcwt_hip_ccwt_device");
}

}  // extern "C"

// ------------------------------------------------ no-allocation entry point
// The device state of fcwt_hip_cwt_device_noalloc: one plan's tables,
// spectrum, schedule and scratch, carved from this array (each GPU the
// library runs on holds its own copy until the process ends).  In host code
// the name is not a device address: the per-device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
struct NoAllocSlot {
  bool ready = false;      // ptr looked up on this device
  char* ptr = nullptr;     // the device address of `scratch` on this device
  fcwt_hip_plan plan = nullptr;  // the plan whose state the array holds
  std::string key;         // what that plan was built for
};
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
NoAllocSlot g_noalloc[kNoAllocDevices];
std::atomic<unsigned> g_tune_generation{0};  // bumped when the tuning table changes
}  // namespace

extern "C" {

int fcwt_hip_cwt_device_noalloc(const float* d_input, int n, const float* sc, int fn, float fb, int threads,
                                int normalize, unsigned flags, float* d_output, void* stream) {
  const char* what = "fcwt_hip_cwt_device_noalloc";
  if (!d_input || !d_output || !sc) return fail(FCWT_HIP_EINVAL, std::string(what) + ": null argument");
  if (n < 1 || fn < 1) return fail(FCWT_HIP_EINVAL, std::string(what) + ": need n >= 1 and fn >= 1");
  if ((1LL << find2power(n)) > kNoAllocMaxN)
    return fail(FCWT_HIP_EUNSUPPORTED, std::string(what) + ": n = " + std::to_string(n) +
                                           " exceeds the limit of " + std::to_string(kNoAllocMaxN) +
                                           " samples (padded size 2^17) of the no-allocation entry point");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return fail(FCWT_HIP_EUNSUPPORTED, std::string(what) + ": device index");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoAllocSlot& slot = g_noalloc[dev];
  if (!slot.ready) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(FCWT_HIP_EDEVICE, std::string(what) + ": hipGetSymbolAddress: " + hipGetErrorString(e));
    slot.ptr = (char*)ptr;
    slot.ready = true;
  }
  // the plan the array holds is reused while the call asks for the same one
  std::string key;
  {
    std::ostringstream ks;
    ks << n << ' ' << fn << ' ' << fb << ' ' << threads << ' ' << (normalize != 0) << ' ' << flags << ' '
       << g_tune_generation.load() << ' ';
    key = ks.str();
    key.append(reinterpret_cast<const char*>(sc), sizeof(float) * (size_t)fn);
  }
  if (!slot.plan || slot.key != key) {
    if (slot.plan) {
      free_plan(slot.plan);
      slot.plan = nullptr;
      slot.key.clear();
    }
    fcwt_hip_plan p = nullptr;
    int rc = create_plan(&p, 1, &n, &sc, &fn, fb, threads, normalize, flags, what, slot.ptr, kNoAllocBytes,
                         (hipStream_t)stream);
    if (rc != FCWT_HIP_OK) return rc;
    slot.plan = p;
    slot.key = key;
  }
  float2* o = (float2*)d_output;
  return launch_all(slot.plan, &d_input, false, &o, (hipStream_t)stream);
}

int fcwt_hip_spectrum_device(fcwt_hip_plan p, const float* d_input, float* d_spec, int complex_input, void* stream) {
  if (!p || !d_input || !d_spec) return fail(FCWT_HIP_EINVAL, "fcwt_hip_spectrum_device: null argument");
  if (p->nsig != 1) return fail(FCWT_HIP_EINVAL, "fcwt_hip_spectrum_device: one-signal plans only");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  launch_forward(p, &d_input, complex_input != 0, (float2*)d_spec, (hipStream_t)stream);
  hipError_t e = hipGetLastError();
  if (cur != p->device) (void)hipSetDevice(cur);
  if (e != hipSuccess) return fail(FCWT_HIP_EDEVICE, std::string("fcwt_hip_spectrum_device: ") + hipGetErrorString(e));
  return FCWT_HIP_OK;
}

// Host-pointer execute: input copy, kernels, output copy, error-word check.
constexpr size_t kStageBytes = 64u << 20;  // per pinned staging buffer

// Copies `bytes` from src to dst with up to `threads` host threads.
static void parallel_memcpy(char* dst, const char* src, size_t bytes, int threads) {
  if (threads <= 1 || bytes < (4u << 20)) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> pool;
  const size_t part = (bytes + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const size_t lo = part * t, hi = std::min(bytes, lo + part);
    if (lo >= hi) break;
    pool.emplace_back([=] { std::memcpy(dst + lo, src + lo, hi - lo); });
  }
  for (auto& th : pool) th.join();
}

// Device output -> caller's (pageable) host buffer in row blocks through two
// pinned staging buffers: the copy engine fills one while host threads empty
// the other.  FCWT_HIP_NO_PIPELINE=1 makes one plain hipMemcpy instead.
static int copy_out(fcwt_hip_plan p, float* output, size_t bytes) {
  const char* env = std::getenv("FCWT_HIP_NO_PIPELINE");
  if ((env && env[0] == '1') || bytes <= kStageBytes) {
    HIP_TRY(hipMemcpyAsync(output, p->d_out, bytes, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return FCWT_HIP_OK;
  }
  if (!p->h_stage[0]) {
    for (int k = 0; k < 2; ++k) {
      if (hipHostMalloc((void**)&p->h_stage[k], kStageBytes, hipHostMallocDefault) != hipSuccess)
        return fail(FCWT_HIP_ENOMEM, "pinned staging allocation failed");
      HIP_TRY(hipEventCreateWithFlags(&p->stage_ev[k], hipEventDisableTiming));
    }
    p->stage_bytes = kStageBytes;
  }
  const unsigned hw = std::thread::hardware_concurrency();
  const int threads = (int)std::max(1u, std::min(8u, hw ? hw : 1u));
  const char* src = (const char*)p->d_out;
  char* dst = (char*)output;
  const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t c) -> int {
    const size_t off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
    HIP_TRY(hipMemcpyAsync(p->h_stage[c & 1], src + off, len, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipEventRecord(p->stage_ev[c & 1], p->stream));
    return FCWT_HIP_OK;
  };
  int rc = issue(0);
  for (size_t c = 0; rc == FCWT_HIP_OK && c < chunks; ++c) {
    // the buffer chunk c + 1 lands in was emptied by the host in step c - 1
    if (c + 1 < chunks) rc = issue(c + 1);
    if (rc != FCWT_HIP_OK) break;
    if (hipEventSynchronize(p->stage_ev[c & 1]) != hipSuccess) return fail(FCWT_HIP_EDEVICE, "output copy failed");
    const size_t off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
    parallel_memcpy(dst + off, p->h_stage[c & 1], len, threads);
  }
  if (rc == FCWT_HIP_OK) HIP_TRY(hipStreamSynchronize(p->stream));
  return rc;
}

static int run_host(fcwt_hip_plan p, const float* input, bool complex_in, float* output, const char* what) {
  if (!p || !input || !output) return fail(FCWT_HIP_EINVAL, std::string(what) + ": null argument");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  int rc = FCWT_HIP_OK;
  const size_t out_elems = (size_t)p->fn * p->n;
  const size_t in_floats = (complex_in ? 2 : 1) * (size_t)p->n;
  if (!p->d_x) rc = dalloc(p, &p->d_x, 2 * (size_t)p->n);  // room for either input type
  if (rc == FCWT_HIP_OK && !p->d_out) rc = dalloc(p, &p->d_out, out_elems);
  if (rc == FCWT_HIP_OK) {
    hipError_t e = hipMemcpyAsync(p->d_x, input, sizeof(float) * in_floats, hipMemcpyHostToDevice, p->stream);
    if (e != hipSuccess) rc = fail(FCWT_HIP_EDEVICE, std::string("input copy failed: ") + hipGetErrorString(e));
  }
  const float* dx = p->d_x;
  if (rc == FCWT_HIP_OK) rc = launch_all(p, &dx, complex_in, &p->d_out, p->stream);
  if (rc == FCWT_HIP_OK) rc = copy_out(p, output, sizeof(float2) * out_elems);
  if (rc == FCWT_HIP_OK && p->nteams) rc = read_error_word(p);
  if (cur != p->device) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_cwt(fcwt_hip_plan p, const float* input, float* output) {
  return run_host(p, input, false, output, "fcwt_hip_cwt");
}

int fcwt_hip_adjoint_device(fcwt_hip_plan p, const float* d_grad, float* d_xbar, int complex_input, void* stream) {
  if (!p || !d_grad || !d_xbar) return fail(FCWT_HIP_EINVAL, "fcwt_hip_adjoint_device: null argument");
  if (p->nsig != 1) return fail(FCWT_HIP_EINVAL, "fcwt_hip_adjoint_device: one-signal plans only");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  int rc = launch_adjoint(p, d_grad, 1, complex_input != 0, d_xbar, (hipStream_t)stream);
  if (cur != p->device) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_adjoint_batch_device(fcwt_hip_plan p, const float* d_grad, float* d_xbar, int signals, int complex_input,
                                  void* stream) {
  if (!p || !d_grad || !d_xbar || signals < 1)
    return fail(FCWT_HIP_EINVAL, "fcwt_hip_adjoint_batch_device: bad arguments");
  if (p->nsig != 1) return fail(FCWT_HIP_EINVAL, "fcwt_hip_adjoint_batch_device: one-signal plans only");
  if (signals > 65535) return fail(FCWT_HIP_EUNSUPPORTED, "fcwt_hip_adjoint_batch_device: at most 65535 signals");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  int rc = launch_adjoint(p, d_grad, signals, complex_input != 0, d_xbar, (hipStream_t)stream);
  if (cur != p->device) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_adjoint(fcwt_hip_plan p, const float* grad, float* xbar, int complex_input) {
  if (!p || !grad || !xbar) return fail(FCWT_HIP_EINVAL, "fcwt_hip_adjoint: null argument");
  if (p->nsig != 1) return fail(FCWT_HIP_EINVAL, "fcwt_hip_adjoint: one-signal plans only");
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  if (cur != p->device) HIP_TRY(hipSetDevice(p->device));
  int rc = FCWT_HIP_OK;
  const size_t g_floats = 2 * (size_t)p->fn * p->n;
  const size_t x_floats = (complex_input ? 2 : 1) * (size_t)p->n;
  if (!p->d_grad) rc = dalloc(p, &p->d_grad, g_floats);
  if (rc == FCWT_HIP_OK && !p->d_xbar) rc = dalloc(p, &p->d_xbar, 2 * (size_t)p->n);
  if (rc == FCWT_HIP_OK) {
    hipError_t e = hipMemcpyAsync(p->d_grad, grad, sizeof(float) * g_floats, hipMemcpyHostToDevice, p->stream);
    if (e != hipSuccess) rc = fail(FCWT_HIP_EDEVICE, std::string("gradient copy failed: ") + hipGetErrorString(e));
  }
  if (rc == FCWT_HIP_OK) rc = launch_adjoint(p, p->d_grad, 1, complex_input != 0, p->d_xbar, p->stream);
  if (rc == FCWT_HIP_OK) {
    hipError_t e = hipMemcpyAsync(xbar, p->d_xbar, sizeof(float) * x_floats, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = fail(FCWT_HIP_EDEVICE, std::string("adjoint failed: ") + hipGetErrorString(e));
  }
  if (cur != p->device) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_ccwt(fcwt_hip_plan p, const float* input, float* output) {
  return run_host(p, input, true, output, "fcwt_hip_ccwt");
}

// ------------------------------------------------------------------ batches
struct fcwt_hip_batch_s {
  struct Group {
    fcwt_hip_plan plan = nullptr;
    std::vector<int> members;  // signals of the batch in this plan, in plan order
  };
  std::vector<Group> groups;
  int count = 0;
  // Graph replay: once the same input and output pointers come twice in a
  // row, that execute's launches are captured (on `cap`) into a graph, which
  // later executes with those pointers replay on the caller's stream.  Any
  // execute with other pointers drops the graph (the plans' device pointer
  // arrays then change).  Capture failing, or FCWT_HIP_NO_GRAPHS=1, leaves
  // the batch eager.
  std::vector<const void*> last_key, graph_key;
  int repeats = 0;
  bool graphs = true;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  long long graph_launches = 0;
  // One team launch for the team rows of every group (k_rows_team_multi),
  // when two or more groups have them: its control words, then each such
  // group's sync words (at msync_off[group]), all zeroed by one fill kernel
  // before the launch.
  bool team_multi = false;
  unsigned* d_msync = nullptr;
  size_t msync_words = 0;
  std::vector<size_t> msync_off;
};

// Whether a batch runs the team rows of all its padded sizes in one launch
// (k_rows_team_multi) rather than one launch per size.
// FCWT_HIP_BATCH_TEAM_LAUNCHES=separate keeps one launch per size.
static bool batch_team_multi() {
  const char* e = std::getenv("FCWT_HIP_BATCH_TEAM_LAUNCHES");
  return !(e && std::strcmp(e, "separate") == 0);
}

// Every launch of one execute of batch b on st: each group's forward
// transforms and row launches, and (team_multi) one team launch for all of
// them after the last group's other launches.
static int launch_batch(fcwt_hip_batch b, const std::vector<std::vector<const float*>>& xs,
                        const std::vector<std::vector<float2*>>& outs, hipStream_t st) {
  TeamMulti m;
  std::memset(&m, 0, sizeof(m));
  fcwt_hip_plan prof = nullptr;  // the plan whose profile counts the team launch
  for (size_t k = 0; k < b->groups.size(); ++k) {
    fcwt_hip_plan p = b->groups[k].plan;
    const bool defer = b->team_multi && p->nteams > 0;
    TeamArgs ta;
    int rc = launch_all(p, xs[k].data(), false, outs[k].data(), st, defer ? &ta : nullptr);
    if (rc != FCWT_HIP_OK) return rc;
    if (defer) {
      const int s = kTeamMaxLog2N - p->log2N;  // seg[0] holds N = 2^18
      if (s < 0 || s >= kTeamMaxSegs || (m.used >> s & 1))  // team plans are 2^15 .. 2^18, one per size
        return fail(FCWT_HIP_EUNSUPPORTED, "fcwt_hip_batch_execute_device: no team-launch slot for N = 2^" +
                                               std::to_string(p->log2N));
      ta.sync = b->d_msync + b->msync_off[k];
      m.seg[s] = ta;
      m.used |= 1 << s;
      if (!prof) prof = p;
    }
  }
  if (m.used) {
    ProfScope ps(prof, 4, st);
    m.ctl = b->d_msync;
    k_fill_u32<<<(unsigned)((b->msync_words + 255) / 256), 256, 0, st>>>(b->d_msync, b->msync_words, 0u);
    void* args[] = {&m};
    HIP_TRY(hipLaunchKernel((const void*)k_rows_team_multi, dim3(256 * kTeamWgPerCu), dim3(kTeamThreads), args, 0,
                            st));
  }
  return FCWT_HIP_OK;
}

static void drop_graph(fcwt_hip_batch b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

int fcwt_hip_batch_create(fcwt_hip_batch* out, int count, const int* lengths, const float* const* scales,
                          const int* fns, float fb, int threads, int normalize, unsigned flags) {
  if (!out) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_create: null batch pointer");
  *out = nullptr;
  if (count < 1 || !lengths || !scales || !fns) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_create: bad arguments");
  for (int c = 0; c < count; ++c)
    if (lengths[c] < 1) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_create: signal length must be >= 1");
  fcwt_hip_batch b = new (std::nothrow) fcwt_hip_batch_s;
  if (!b) return fail(FCWT_HIP_ENOMEM, "fcwt_hip_batch_create: out of host memory");
  b->count = count;
  if (const char* e = std::getenv("FCWT_HIP_NO_GRAPHS")) b->graphs = e[0] != '1';
  std::vector<int> order(count);
  for (int c = 0; c < count; ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return find2power(lengths[a]) < find2power(lengths[c]); });
  for (size_t k = 0; k < order.size();) {
    size_t e = k;
    while (e < order.size() && find2power(lengths[order[e]]) == find2power(lengths[order[k]])) ++e;
    fcwt_hip_batch_s::Group g;
    g.members.assign(order.begin() + k, order.begin() + e);
    std::vector<int> len, fnv;
    std::vector<const float*> sc;
    for (int c : g.members) {
      len.push_back(lengths[c]);
      fnv.push_back(fns[c]);
      sc.push_back(scales[c]);
    }
    int rc = create_plan(&g.plan, (int)g.members.size(), len.data(), sc.data(), fnv.data(), fb, threads, normalize,
                         flags, "fcwt_hip_batch_create");
    if (rc != FCWT_HIP_OK) {
      for (auto& h : b->groups) free_plan(h.plan);
      delete b;
      return rc;
    }
    b->groups.push_back(std::move(g));
    k = e;
  }
  int team_groups = 0;
  for (auto& g : b->groups) team_groups += g.plan->nteams > 0;
  if (team_groups >= 2 && batch_team_multi()) {
    b->msync_words = 64;  // the launch's abort and exit words
    b->msync_off.assign(b->groups.size(), 0);
    for (size_t k = 0; k < b->groups.size(); ++k)
      if (b->groups[k].plan->nteams > 0) {
        b->msync_off[k] = b->msync_words;
        b->msync_words += b->groups[k].plan->team_sync_words;
      }
    hipError_t e = hipMalloc((void**)&b->d_msync, b->msync_words * sizeof(unsigned));
    if (e != hipSuccess) {
      for (auto& h : b->groups) free_plan(h.plan);
      delete b;
      return fail(FCWT_HIP_ENOMEM, std::string("fcwt_hip_batch_create: hipMalloc of the team launch's words: ") +
                                       hipGetErrorString(e));
    }
    b->team_multi = true;
  }
  *out = b;
  return FCWT_HIP_OK;
}

int fcwt_hip_batch_destroy(fcwt_hip_batch b) {
  if (!b) return FCWT_HIP_OK;
  drop_graph(b);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->d_msync) (void)hipFree(b->d_msync);
  for (auto& g : b->groups) free_plan(g.plan);
  delete b;
  return FCWT_HIP_OK;
}

int fcwt_hip_batch_execute_device(fcwt_hip_batch b, const float* const* d_inputs, float* const* d_outputs,
                                  void* stream) {
  if (!b || !d_inputs || !d_outputs) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_execute_device: null argument");
  for (int c = 0; c < b->count; ++c)
    if (!d_inputs[c] || !d_outputs[c]) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_execute_device: null pointer");
  hipStream_t st = (hipStream_t)stream;
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  const int dev = b->groups[0].plan->device;
  if (cur != dev) HIP_TRY(hipSetDevice(dev));
  std::vector<const void*> key(d_inputs, d_inputs + b->count);
  key.insert(key.end(), d_outputs, d_outputs + b->count);
  bool profiling = false;
  for (auto& g : b->groups) profiling |= g.plan->profiling;
  int rc = FCWT_HIP_OK;
  if (b->exec && key == b->graph_key && !profiling) {
    hipError_t e = hipGraphLaunch(b->exec, st);
    if (e == hipSuccess) ++b->graph_launches;
    rc = e == hipSuccess ? FCWT_HIP_OK : fail(FCWT_HIP_EDEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    if (cur != dev) (void)hipSetDevice(cur);
    return rc;
  }
  if (key != b->graph_key) drop_graph(b);
  if (key == b->last_key) {
    ++b->repeats;
  } else {
    b->repeats = 0;
    b->last_key = key;
  }
  // each group's signals, in plan order
  std::vector<std::vector<const float*>> xs(b->groups.size());
  std::vector<std::vector<float2*>> outs(b->groups.size());
  for (size_t k = 0; k < b->groups.size(); ++k)
    for (int c : b->groups[k].members) {
      xs[k].push_back(d_inputs[c]);
      outs[k].push_back((float2*)d_outputs[c]);
    }
  if (b->graphs && b->repeats >= 1 && !profiling) {
    // pointer arrays first, on the caller's stream (not part of the graph)
    for (size_t k = 0; k < b->groups.size() && rc == FCWT_HIP_OK; ++k)
      if (b->groups[k].plan->nsig > 1) rc = upload_pointers(b->groups[k].plan, xs[k].data(), outs[k].data(), st);
    if (rc != FCWT_HIP_OK) {
      if (cur != dev) (void)hipSetDevice(cur);
      return rc;
    }
    bool ok = b->cap || hipStreamCreate(&b->cap) == hipSuccess;
    ok = ok && hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      const int lrc = launch_batch(b, xs, outs, b->cap);
      hipGraph_t g = nullptr;
      ok = hipStreamEndCapture(b->cap, &g) == hipSuccess && lrc == FCWT_HIP_OK && g;
      if (ok) {
        b->graph = g;
        ok = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) == hipSuccess;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    if (ok && hipGraphLaunch(b->exec, st) == hipSuccess) {
      b->graph_key = key;
      ++b->graph_launches;
      if (cur != dev) (void)hipSetDevice(cur);
      return FCWT_HIP_OK;
    }
    // capture is not available here: stay eager from now on
    drop_graph(b);
    b->graphs = false;
    (void)hipGetLastError();
  }
  rc = launch_batch(b, xs, outs, st);
  if (cur != dev) (void)hipSetDevice(cur);
  return rc;
}

int fcwt_hip_batch_graph_launches(fcwt_hip_batch b, long long* launches) {
  if (!b || !launches) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_graph_launches: null argument");
  *launches = b->graph_launches;
  return FCWT_HIP_OK;
}

int fcwt_hip_batch_check(fcwt_hip_batch b) {
  if (!b) return fail(FCWT_HIP_EINVAL, "null batch");
  // every plan's report is read (and cleared); the first failure is returned
  int first = FCWT_HIP_OK;
  std::string msg;
  for (auto& g : b->groups) {
    const int rc = fcwt_hip_plan_check(g.plan);
    if (rc != FCWT_HIP_OK && first == FCWT_HIP_OK) {
      first = rc;
      msg = g_last_error;
    }
  }
  if (first != FCWT_HIP_OK) g_last_error = msg;
  return first;
}

// ------------------------------------------------------------- tuning
// Times one plan of n = N - N/4 samples and fn LOGSCALES 1-450 Hz at
// fs = 1000 (3000 rows, fewer where the output would pass 3 GB) with the
// schedule parameters `cand`; ms per execute.
static int tune_time(int N, int threads, const Tuned& cand, int reps, float* ms) {
  const int n = std::max(1, N - N / 4);
  const int fn = (int)std::max<long long>(16, std::min<long long>(3000, 3000000000LL / (8LL * n)));
  std::vector<float> sc(fn), x(n);
  scales(1, 1000, 1.0f, 450.0f, fn, sc.data());
  for (int t = 0; t < n; ++t) x[t] = std::sin(0.05f * t) + 0.5f * std::sin(0.9f * t + 1.0f);
  t_force = cand;
  t_forcing = true;
  fcwt_hip_plan p = nullptr;
  int rc = fcwt_hip_plan_create(&p, n, sc.data(), fn, 2.0f, threads, 1, 0);
  t_forcing = false;
  if (rc != FCWT_HIP_OK) return rc;
  float* dx = nullptr;
  float* dout = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = FCWT_HIP_OK;
  if (hipMalloc(&dx, sizeof(float) * n) != hipSuccess || hipMalloc(&dout, sizeof(float2) * (size_t)fn * n) != hipSuccess)
    rc = fail(FCWT_HIP_ENOMEM, "fcwt_hip_tune: device allocation failed");
  if (rc == FCWT_HIP_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(FCWT_HIP_EDEVICE, "fcwt_hip_tune: event creation failed");
  if (rc == FCWT_HIP_OK && hipMemcpy(dx, x.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(FCWT_HIP_EDEVICE, "fcwt_hip_tune: input copy failed");
  float best = 1e30f;
  for (int r = 0; rc == FCWT_HIP_OK && r <= reps; ++r) {
    (void)hipEventRecord(e0, p->stream);
    rc = fcwt_hip_cwt_device(p, dx, dout, p->stream);
    (void)hipEventRecord(e1, p->stream);
    if (rc == FCWT_HIP_OK) rc = fcwt_hip_plan_check(p);
    float t = 0.f;
    if (rc == FCWT_HIP_OK && hipEventElapsedTime(&t, e0, e1) == hipSuccess && r > 0) best = std::min(best, t);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (dx) (void)hipFree(dx);
  if (dout) (void)hipFree(dout);
  fcwt_hip_plan_destroy(p);
  *ms = best;
  return rc;
}

int fcwt_hip_tune(int N, int threads, int effort, const char* path) {
  if (N < 2 || (N & (N - 1)) || threads < 1 || effort < 0 || effort > 2 || !path)
    return fail(FCWT_HIP_EINVAL, "fcwt_hip_tune: bad arguments");
  const int log2N = find2power(N);
  const int log2F = log2N - kTeamLog2B;  // at B = 256 (the column path's split below 2^22)
  const int team_log2F = log2N - team_log2b(log2N);
  int rc = FCWT_HIP_OK;
  bool teams = false;
  if (effort > 0) {  // effort 0 writes the defaults and needs no device
    int dev = 0;
    rc = check_device(&dev);
    if (rc != FCWT_HIP_OK) return rc;
    teams = team_capable(dev, log2N) && team_log2F >= kTeamMinLog2F && team_log2F <= kTeamMaxLog2F &&
            log2N <= kTeamMaxLog2N;
  }
  const bool cols = log2F >= kColsMinLog2F && log2F <= kColsMaxLog2F;
  // the built-in choice, written as is for effort 0 or sizes with nothing to choose
  Tuned best;
  best.team_buffers = team_default_buffers(N);
  best.cols_mb = (int)(kColsScratchBytes >> 20);
  best.path = 0;  // the built-in routing unless a timed candidate wins
  float best_ms = -1.f, default_ms = -1.f;
  if (effort > 0 && (teams || cols)) {
    // candidates: the team kernel's buffer counts, and the column path's
    // scratch per batch (beside the team kernel where it runs too)
    // candidates: the team kernel's buffer count, the lag between a row's
    // two passes and the direct/team crossover; the column path's scratch
    // per batch and whether its passes share launches
    std::vector<Tuned> cand;
    auto add = [&](int b, int mb, int path, int lag, int dmax, int sets, int lb = 0) {
      Tuned t;
      t.team_buffers = b;
      t.cols_mb = mb;
      t.path = path;
      t.team_lag = lag;
      t.direct_max = dmax;
      t.cols_sets = sets;
      t.cols_log2B = lb;
      cand.push_back(t);
    };
    if (teams)
      for (int b : {2, 3, 4, 6, 8, 12}) {
        if (effort < 2 && (b < 4 || b > 8)) continue;
        for (int lag : {b / 2, b / 2 - 1, b / 2 + 1})
          if (lag >= 1 && lag < b && (effort == 2 || lag == b / 2 || b == 6))
            for (int dmax : {10, 9})
              if (effort == 2 || dmax == 10 || (b == 6 && lag == 3)) add(b, best.cols_mb, 1, lag, dmax, 0);
      }
    for (int mb : {32, 64, 128, 256, 512})
      if (effort == 2 || (mb >= 64 && mb <= 256))
        for (int sets : {0, 1})
          if (effort == 2 || sets == 0 || mb == 128)
            for (int lb : {0, kTeamLog2B})  // at N = 2^22 also the 256 x 2^14 split (cols_split)
              if (lb == 0 || log2N == 22) add(best.team_buffers, mb, 2, 0, 0, sets, lb);
    const int reps = effort == 1 ? 3 : 7;  // timed runs per candidate (the fastest counts)
    // the built-in schedule first: it is kept unless a candidate beats it
    rc = tune_time(N, threads, Tuned{}, reps, &default_ms);
    if (rc != FCWT_HIP_OK) return rc;
    best_ms = default_ms;
    for (const Tuned& c : cand) {
      float ms = 0.f;
      rc = tune_time(N, threads, c, reps, &ms);
      if (rc != FCWT_HIP_OK) return rc;
      if (best_ms < 0.f || ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  std::ofstream f(path);
  if (!f) return fail(FCWT_HIP_EIO, std::string("fcwt_hip_tune: cannot write ") + path);
  f << "fcwt_hip tuning 1\narch gfx950\nN " << N << "\nthreads " << threads << "\neffort " << effort
    << "\nteam_buffers " << best.team_buffers << "\ncols_mb " << best.cols_mb << "\npath "
    << (best.path == 1 ? "team" : best.path == 2 ? "cols" : "default") << "\nteam_lag " << best.team_lag
    << "\ndirect_max " << best.direct_max << "\ncols_sets " << best.cols_sets << "\ncols_log2B " << best.cols_log2B
    << "\ntime_ms " << best_ms
    << "\ndefault_ms " << default_ms << "\n";
  if (!f) return fail(FCWT_HIP_EIO, std::string("fcwt_hip_tune: cannot write ") + path);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[N] = best;
  ++g_tune_generation;
  return FCWT_HIP_OK;
}

int fcwt_hip_tune_load(const char* path) {
  if (!path) return fail(FCWT_HIP_EINVAL, "fcwt_hip_tune_load: null path");
  std::ifstream f(path);
  if (!f) return fail(FCWT_HIP_EIO, std::string("fcwt_hip_tune_load: cannot read ") + path);
  std::string line, key, arch, pth;
  int N = 0, version = 0;
  Tuned t;
  if (!std::getline(f, line) || std::sscanf(line.c_str(), "fcwt_hip tuning %d", &version) != 1 || version != 1)
    return fail(FCWT_HIP_EIO, std::string("fcwt_hip_tune_load: not a tuning file: ") + path);
  while (std::getline(f, line)) {
    std::istringstream is(line);
    is >> key;
    if (key == "arch") is >> arch;
    else if (key == "N") is >> N;
    else if (key == "team_buffers") is >> t.team_buffers;
    else if (key == "cols_mb") is >> t.cols_mb;
    else if (key == "path") is >> pth;
    else if (key == "team_lag") is >> t.team_lag;
    else if (key == "direct_max") is >> t.direct_max;
    else if (key == "cols_sets") is >> t.cols_sets;
    else if (key == "cols_log2B") is >> t.cols_log2B;
  }
  t.path = pth == "team" ? 1 : pth == "cols" ? 2 : 0;
  if (arch != "gfx950") return fail(FCWT_HIP_EIO, std::string("fcwt_hip_tune_load: tuned for another GPU: ") + path);
  if (N < 2 || (N & (N - 1)) || (t.team_buffers && (t.team_buffers < 2 || t.team_buffers > kTeamBuffers)) ||
      t.cols_mb < 0 || t.cols_mb > 4096 || !(pth.empty() || pth == "team" || pth == "cols" || pth == "default") ||
      t.team_lag < 0 || (t.team_buffers && t.team_lag >= t.team_buffers) ||
      !(t.direct_max == 0 || t.direct_max == 9 || t.direct_max == 10) || t.cols_sets < 0 || t.cols_sets > 2 ||
      !(t.cols_log2B == 0 || t.cols_log2B == kTeamLog2B || t.cols_log2B == 9))
    return fail(FCWT_HIP_EIO, std::string("fcwt_hip_tune_load: bad contents: ") + path);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[N] = t;
  ++g_tune_generation;
  return FCWT_HIP_OK;
}

// Row-kernel launches one execute of plan p makes, its team launch counted
// only when `team` (launch_all).
static int plan_row_launches(const fcwt_hip_plan p, bool team) {
  int n = 0;
  for (const auto& gr : p->direct_groups) n += !gr.fill;
  if (p->nteams && team) ++n;
  if (!p->cols_batches.empty()) {
    const int nb = (int)p->cols_batches.size();
    n += 1 + (p->cols_sets == 2 ? nb + 1 : 2 * nb);  // the spectrum transpose, then the passes
  }
  n += 2 * (int)p->batches.size();
  return n;
}

int fcwt_hip_batch_row_launches(fcwt_hip_batch b, int* launches) {
  if (!b || !launches) return fail(FCWT_HIP_EINVAL, "fcwt_hip_batch_row_launches: null argument");
  int n = 0;
  bool any_team = false;
  for (auto& g : b->groups) {
    n += plan_row_launches(g.plan, !b->team_multi);
    any_team |= g.plan->nteams > 0;
  }
  *launches = n + (b->team_multi && any_team ? 1 : 0);
  return FCWT_HIP_OK;
}

int fcwt_hip_batch_shape(fcwt_hip_batch b, int* groups, int* team_rows) {
  if (!b) return fail(FCWT_HIP_EINVAL, "null batch");
  if (groups) *groups = (int)b->groups.size();
  if (team_rows) {
    int t = 0;
    for (auto& g : b->groups) {
      int tr = 0;
      fcwt_hip_plan_row_split(g.plan, nullptr, &tr, nullptr);
      t += tr;
    }
    *team_rows = t;
  }
  return FCWT_HIP_OK;
}

}  // extern "C"
