// rows_big.h -- band-limited inverse FFT for wide scales (band M > 1024 bins)
// at padded sizes N = B F with B = 256 (2^5 <= F <= 2^14) or B = 512
// (F = 2^13, N = 2^22), in two passes per batch of rows with the
// intermediate in one global scratch slot per row; each launch runs pass 1
// of one batch beside pass 2 of the previous one.
// Device code, included by fcwt_hip.hip after rows_team.h.
//
// With the band M = B A (A <= F/2):
//   y[f + F m] = c sum_{qb<B} W_B^{qb m} W_N^{qb f} U[qb][f]
//   U[qb][f]   = sum_{qf<A} Z[qb + B qf] W_F^{qf f}
//
// Pass 1: a work-group transforms Qc = W/F consecutive columns qb of
// one row (W = 2^13 points, 2^14 when F = 2^14).  It reads each column's
// band from the signal's transposed spectrum Xt[qb][qf] (contiguous along
// qf), multiplies it by the window, whose value it computes with the mother
// table's own float32 formula (a gather from the table would touch one cache
// line per bin), runs F-point inverse FFTs as radix-16 Stockham passes in
// LDS, and stores U[qb][0, F) of its columns as one contiguous run of the
// row's scratch slot.
// Pass 2 (B = 256): 256 threads take 16 consecutive f of one row.  They read
// U[qb][f0, f0 + 16) (128-byte runs), apply W_N^{qb f} and the 256-point
// inverse FFT over qb = j + 16 k as two register radix-16 steps around one
// LDS exchange, and write y[f0 + g + F m] as 128-byte runs.
// Pass 2 (B = 512): 512 threads take 16 consecutive f; the 512-point FFT
// over qb = j + 32 k is a register radix-16 step, one LDS exchange, a
// register radix-16 step and a radix-2 step between lane pairs
// (cols_pass2_b512).
#pragma once
#include "fft_lds.h"
#include "fft_reg.h"

namespace fcwt_hip {

constexpr int kColsMinLog2F = 5;
constexpr int kColsMaxLog2F = 14;

struct BigArgs {
  KArgs k;
  const float2* xt;  // transposed half-spectra, signal s at xt + s N/2: [qb][qf], F/2 values per qb
  const int* rows;   // the batch's rows: blockIdx.y -> row rows[blockIdx.y], scratch slot blockIdx.y
  float2* scratch;   // slot y at scratch + y N, laid out U[qb][f]
  float fb, toradians, mnorm;  // the mother table's parameters (k_mother)
};

// Xt[s][qb][qf] = X_s[qb + B qf] for qb < B = 2^log2B, qf < F/2, through
// 64 x 16 LDS tiles; grid (B/16 * ceil(F/128), signals).  (F/2 < 64 leaves
// part of the one tile column unused.)
__global__ __launch_bounds__(256) void k_spec_transpose(const float2* __restrict__ spec, float2* __restrict__ xt,
                                                        int log2F, int log2B) {
  __shared__ float2 tile[64 * 17];
  const int Fh = 1 << (log2F - 1);
  const int qb0 = (blockIdx.x & ((1 << (log2B - 4)) - 1)) << 4, qf0 = (blockIdx.x >> (log2B - 4)) << 6;
  const size_t sig = (size_t)blockIdx.y * ((size_t)Fh << log2B);
  const float2* X = spec + sig;
  float2* Y = xt + sig;
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int qb = e & 15, qf = e >> 4;
    if (qf0 + qf < Fh) tile[qf * 17 + qb] = X[((size_t)(qf0 + qf) << log2B) + qb0 + qb];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int qf = e & 63, qb = e >> 6;
    if (qf0 + qf < Fh) Y[(size_t)(qb0 + qb) * Fh + qf0 + qf] = tile[qf * 17 + qb];
  }
}

template <int LOG2F>
struct ColsCfg {
  static constexpr int LOG2W = LOG2F > 13 ? LOG2F : 13;  // points per work-group
  static constexpr int W = 1 << LOG2W;
  static constexpr int T = W >> 4;                        // threads: 16 points each
  static constexpr int LOG2QC = LOG2W - LOG2F;            // columns per work-group
  static constexpr int PADW = W + (W >> 4);
};

// LDS index of point i: one pad word per 16 points keeps the radix-16 writes
// (stride 16) on distinct banks.
__device__ __forceinline__ int cpad(int i) { return i + (i >> 4); }

// Twiddle tables of the column transforms: t1 holds W_F^r for r < min(F, 128),
// t2 holds W_F^{128 s} for s < max(1, F/128).
template <int LOG2F>
struct ColsTw {
  static constexpr int T1N = LOG2F < 7 ? (1 << LOG2F) : 128;
  static constexpr int T2N = LOG2F < 7 ? 1 : (1 << LOG2F) / 128;
};

// W_F^r (r < F) from the LDS tables t1[r & 127] = W_F^{r mod 128}, t2[r >> 7] = W_F^{128 (r >> 7)}
__device__ __forceinline__ float2 cols_tw(const float2* t1, const float2* t2, int r) {
  return cmul(t2[r >> 7], t1[r & 127]);
}

// The value that the neighbouring lane (lane l ^ 1) holds in x, by DPP
// (quad_perm 1, 0, 3, 2); every lane of the wave must be active.
__device__ __forceinline__ float lane_xor1(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// The value that lane Q of this lane's quad holds in x, by DPP (quad_perm
// Q, Q, Q, Q); every lane of the wave must be active.
template <int Q>
__device__ __forceinline__ float lane_quad(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), Q * 0x55, 0xF, 0xF, false));
}

// The windowed band of one pass-1 work-group's columns: column c holds
// z_c[g] = Xt[qb0 + c][g] psi(qb0 + c + B g) for g < A, 0 beyond.
struct BandSrc {
  const float2* X;  // the row's signal's Xt: column qb at X + qb F/2
  const SegD* sseg;  // the row's segments (LDS)
  int qb0, A, log2S, nseg;  // log2S: the first log2S radix-2 factors only replicate (cols_skip)
  float maxidx, fb, toradians, mnorm;
};

// One radix-2^LOG2R Stockham pass over the work-group's columns (each F
// points, natural order in and out); Ns = 2^LOG2NS points already combined.
// IN_G: the pass's input is read from the band instead of LDS: position p of
// column c after the skipped passes is z_c[p >> log2S] (they only replicate;
// cols_fft takes this form for S = 2^log2S <= 4, where lane pairs or quads
// share their values and split the window evaluations).
// OUT_G: the last pass stores its output to the row's scratch U[qb][f]
// instead of LDS; its lanes write consecutive f.
template <int LOG2F, int LOG2R, int LOG2NS, bool IN_G, bool OUT_G, int LOG2B>
__device__ __forceinline__ void cols_pass(float2* d, const float2* t1, const float2* t2, const BandSrc& src,
                                          gfloat2* U) {
  using C = ColsCfg<LOG2F>;
  constexpr int R = 1 << LOG2R, Ns = 1 << LOG2NS;
  constexpr int LOG2GPC = LOG2F - LOG2R;            // groups per column
  constexpr int PER = (C::W >> LOG2R) / C::T;       // groups per thread
  float2 v[PER][R];
  if constexpr (IN_G) {
    // every load first, then the windows
    constexpr int Fh = 1 << (LOG2F - 1);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int g = threadIdx.x + u * C::T;
      const int c = g >> LOG2GPC, gl = g & ((1 << LOG2GPC) - 1);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int gq = (gl + (k << LOG2GPC)) >> src.log2S;
        v[u][k] = gq < src.A ? src.X[(size_t)(src.qb0 + c) * Fh + gq] : make_float2(0.f, 0.f);
      }
    }
    auto window = [&](int c, int gq) {
      float w = 0.f;
      if (gq < src.A) {
        const int idx = window_index(src.qb0 + c + (gq << LOG2B), src.nseg, src.sseg, src.maxidx);
        if (idx >= 0) w = mother_value(idx, src.fb, src.toradians, src.mnorm);
      }
      return w;
    };
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int g = threadIdx.x + u * C::T;
      const int c = g >> LOG2GPC, gl = g & ((1 << LOG2GPC) - 1);
      if constexpr (LOG2NS == 1) {
        // S = 2: lanes 2i and 2i + 1 hold the same values, so each evaluates
        // the windows of half the k and takes the other half from its partner
        const int odd = gl & 1;
#pragma unroll
        for (int kk = 0; kk < R / 2; ++kk) {
          const float wo = window(c, (gl + ((2 * kk + odd) << LOG2GPC)) >> 1);
          const float wp = lane_xor1(wo);
          const float w0 = odd ? wp : wo, w1 = odd ? wo : wp;  // k = 2 kk, 2 kk + 1
          v[u][2 * kk] = make_float2(v[u][2 * kk].x * w0, v[u][2 * kk].y * w0);
          v[u][2 * kk + 1] = make_float2(v[u][2 * kk + 1].x * w1, v[u][2 * kk + 1].y * w1);
        }
      } else if constexpr (LOG2NS == 2) {
        // S = 4: the four lanes of a quad hold the same values; lane t
        // evaluates the windows of k = 4 kk + t and reads the others from
        // the quad's other lanes
        const int t = gl & 3;
#pragma unroll
        for (int kk = 0; kk < R / 4; ++kk) {
          const float wo = window(c, (gl + ((4 * kk + t) << LOG2GPC)) >> 2);
          const float w[4] = {lane_quad<0>(wo), lane_quad<1>(wo), lane_quad<2>(wo), lane_quad<3>(wo)};
#pragma unroll
          for (int q = 0; q < 4; ++q)
            v[u][4 * kk + q] = make_float2(v[u][4 * kk + q].x * w[q], v[u][4 * kk + q].y * w[q]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const float w = window(c, (gl + (k << LOG2GPC)) >> src.log2S);
          v[u][k] = make_float2(v[u][k].x * w, v[u][k].y * w);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int g = threadIdx.x + u * C::T;
    const int c0 = (g >> LOG2GPC) << LOG2F, gl = g & ((1 << LOG2GPC) - 1);
    if constexpr (!IN_G) {
#pragma unroll
      for (int k = 0; k < R; ++k) v[u][k] = d[cpad(c0 + gl + (k << LOG2GPC))];
    }
    if constexpr (LOG2NS > 0) {
      const int jj = gl & (Ns - 1);
      constexpr int SH = LOG2F - LOG2NS - LOG2R;
      if constexpr (R == 16) {
        // W^k for k < 16 from W^1 and W^4 (tables), each at most three
        // products deep
        float2 w[16];
        w[1] = cols_tw(t1, t2, jj << SH);
        w[4] = cols_tw(t1, t2, (4 * jj) << SH);
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
        for (int k = 1; k < R; ++k) v[u][k] = cmul(v[u][k], cols_tw(t1, t2, (jj * k) << SH));
      }
    }
    dft_reg<R, 1>(v[u]);
  }
  if constexpr (OUT_G) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int g = threadIdx.x + u * C::T;
      const int c0 = (g >> LOG2GPC) << LOG2F, gl = g & ((1 << LOG2GPC) - 1);
      const int dst = c0 + ((gl >> LOG2NS) << (LOG2NS + LOG2R)) + (gl & (Ns - 1));
#pragma unroll
      for (int k = 0; k < R; ++k) gstore(U + dst + (k << LOG2NS), v[u][k]);
    }
  } else {
    if constexpr (!IN_G) __syncthreads();  // every read of this pass is done
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int g = threadIdx.x + u * C::T;
      const int c0 = (g >> LOG2GPC) << LOG2F, gl = g & ((1 << LOG2GPC) - 1);
      const int dst = c0 + ((gl >> LOG2NS) << (LOG2NS + LOG2R)) + (gl & (Ns - 1));
#pragma unroll
      for (int k = 0; k < R; ++k) d[cpad(dst + (k << LOG2NS))] = v[u][k];
    }
    __syncthreads();
  }
}

// Pass J of the inverse F-point DFTs of the work-group's columns (radix
// 2^(LOG2F mod 4) first, radix 16 after) and the passes after it.  Passes
// before `skip` only replicate the band (cols_skip) and are not run.  With
// S <= 4 (the radix product of the skipped passes) pass `skip` reads the
// band itself; otherwise the caller has stored it in LDS, each value at its
// S places, and pass `skip` reads it there (reading it from the band would
// evaluate each window S times).  The last pass stores to U; the passes
// between exchange through LDS.
template <int LOG2F, int LOG2B, int J>
__device__ __forceinline__ void cols_fft(float2* d, const float2* t1, const float2* t2, int skip,
                                         const BandSrc& src, gfloat2* U) {
  constexpr int FIRST = LOG2F % 4;
  constexpr int NPASS = (FIRST > 0) + (LOG2F - FIRST) / 4;
  static_assert(LOG2F >= kColsMinLog2F && LOG2F <= kColsMaxLog2F, "F out of range");
  if constexpr (J < NPASS) {
    constexpr int LOG2R = (J == 0 && FIRST > 0) ? FIRST : 4;
    constexpr int LOG2NS = J == 0 ? 0 : (FIRST > 0 ? FIRST + 4 * (J - 1) : 4 * J);
    constexpr bool LAST = J == NPASS - 1;
    if constexpr (LOG2NS <= 2) {
      if (skip == J)
        cols_pass<LOG2F, LOG2R, LOG2NS, true, LAST, LOG2B>(d, t1, t2, src, U);
      else if (skip < J)
        cols_pass<LOG2F, LOG2R, LOG2NS, false, LAST, LOG2B>(d, t1, t2, src, U);
    } else if (skip <= J) {
      cols_pass<LOG2F, LOG2R, LOG2NS, false, LAST, LOG2B>(d, t1, t2, src, U);
    }
    cols_fft<LOG2F, LOG2B, J + 1>(d, t1, t2, skip, src, U);
  }
}

// Passes of cols_fft that only replicate for a band of A values, and the
// log2 of their radix product.
template <int LOG2F>
__device__ __forceinline__ int cols_skip(int A, int* log2S) {
  constexpr int FIRST = LOG2F % 4;
  constexpr int NPASS = (FIRST > 0) + (LOG2F - FIRST) / 4;
  int skip = 0, l2 = 0, step = FIRST > 0 ? FIRST : 4;
  while (skip < NPASS && (A << (l2 + step)) <= (1 << LOG2F)) {
    l2 += step;
    ++skip;
    step = 4;
  }
  *log2S = l2;
  return skip;
}

// LDS of the column kernel (complex64 values): pass 1 holds the columns
// (PADW), its twiddle tables (128 + T2N) and the row's segments; pass 2
// holds, per 256-thread item at B = 256, the exchange/tile (16 x 272) and
// two factor tables of 256; at B = 512, per 512-thread item, the exchange
// (16 x 16 x 33) and factor tables of 512 and 256.
template <int LOG2F, int LOG2B = 8>
struct ColsLds {
  using C = ColsCfg<LOG2F>;
  static constexpr int P1 = C::PADW + 128 + ColsTw<LOG2F>::T2N + 2 * kTeamSegLds;
  static constexpr int P2_ITEM = LOG2B == 9 ? 16 * 16 * 33 + 512 + 256 : 16 * 272 + 512;
  static constexpr int P2 = LOG2B == 9 ? P2_ITEM : (C::T / 256) * P2_ITEM;
  static constexpr int SIZE = P1 > P2 ? P1 : P2;
};

// Work-groups per row of the batch in one column-path launch: X1 of pass 1
// (Qc columns each), X2 of pass 2 (SUB items of 16 f each).
template <int LOG2F, int LOG2B>
struct ColsGrid {
  using C = ColsCfg<LOG2F>;
  static constexpr int X1 = (1 << LOG2B) >> C::LOG2QC;
  static constexpr int SUB = LOG2B == 9 ? 1 : C::T / 256;
  static constexpr int X2 = ((1 << LOG2F) / 16) / SUB;
  static_assert(LOG2B == 8 || (LOG2B == 9 && C::T == 512), "B = 512 takes 512-thread work-groups (F <= 2^13)");
  static_assert(((1 << LOG2F) / 16) % SUB == 0, "pass-2 items per row must fill whole work-groups");
};

// Pass 1 for row y of the batch, column block xblk (Qc consecutive qb).
template <int LOG2F, int LOG2B>
__device__ __forceinline__ void cols_pass1(const BigArgs& a, int y, int xblk, float2* lds) {
  using C = ColsCfg<LOG2F>;
  constexpr int F = 1 << LOG2F;
  constexpr int PT = C::W / C::T;  // points per thread: 16
  float2* const d = lds;
  float2* const t1 = lds + C::PADW;
  float2* const t2 = t1 + 128;
  SegD* const sseg = reinterpret_cast<SegD*>(t2 + ColsTw<LOG2F>::T2N);
  const RowD r = a.k.rows[a.rows[y]];
  const int qb0 = xblk << C::LOG2QC;
  for (int e = threadIdx.x; e < ColsTw<LOG2F>::T1N + ColsTw<LOG2F>::T2N; e += C::T) {  // W_F^r = W_N^{B r}
    if (e < ColsTw<LOG2F>::T1N)
      t1[e] = a.k.tw((unsigned)e << LOG2B);
    else
      t2[e - ColsTw<LOG2F>::T1N] = a.k.tw((unsigned)(e - ColsTw<LOG2F>::T1N) << (LOG2B + 7));
  }
  const int nseg = r.seg_count;  // <= kTeamSegLds: the schedule routes rows with more elsewhere
  if ((int)threadIdx.x < nseg) sseg[threadIdx.x] = a.k.segs[r.seg_begin + threadIdx.x];
  __syncthreads();
  BandSrc src;
  src.X = a.xt + (size_t)r.sig * (size_t)(a.k.N >> 1);
  src.sseg = sseg;
  src.qb0 = qb0;
  src.A = min(1 << r.log2A, F >> 1);  // the band lies below N/2
  src.nseg = nseg;
  src.maxidx = (float)a.k.N - 1.0f;
  src.fb = a.fb;
  src.toradians = a.toradians;
  src.mnorm = a.mnorm;
  const int skip = cols_skip<LOG2F>(src.A, &src.log2S);
  gfloat2* U = (gfloat2*)(a.scratch + (size_t)y * a.k.N) + ((size_t)qb0 << LOG2F);
  constexpr int FIRST = LOG2F % 4;
  constexpr int NPASS = (FIRST > 0) + (LOG2F - FIRST) / 4;
  if (skip < NPASS) {
    if (src.log2S >= 3) {
      // the columns' windowed bands into LDS, each value at its S places;
      // all loads are issued before the first is used
      const int log2G = LOG2F - src.log2S;  // values per column
      constexpr int Fh = 1 << (LOG2F - 1);
      float2 xs[PT];
#pragma unroll
      for (int u = 0; u < PT; ++u) {
        const int e = u * C::T + threadIdx.x;  // value g of column c
        const int c = e >> log2G, g = e & ((1 << log2G) - 1);
        xs[u] = (g < src.A && c < (1 << C::LOG2QC)) ? src.X[(size_t)(qb0 + c) * Fh + g] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < PT; ++u) {
        const int e = u * C::T + threadIdx.x;
        const int c = e >> log2G, g = e & ((1 << log2G) - 1);
        if (c < (1 << C::LOG2QC)) {
          float2 z = make_float2(0.f, 0.f);
          if (g < src.A) {
            const int idx = window_index(qb0 + c + (g << LOG2B), nseg, sseg, src.maxidx);
            if (idx >= 0) z = cscale(xs[u], mother_value(idx, a.fb, a.toradians, a.mnorm));
          }
          const int d0 = (c << LOG2F) + (g << src.log2S);
          for (int t = 0; t < (1 << src.log2S); ++t) d[cpad(d0 + t)] = z;
        }
      }
      __syncthreads();
    }
    cols_fft<LOG2F, LOG2B, 0>(d, t1, t2, skip, src, U);
  } else {
    // every pass only replicates: column c is z_c[p >> log2S] (bands of one bin per column)
    constexpr int Fh = 1 << (LOG2F - 1);
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = u * C::T + threadIdx.x;
      const int c = e >> LOG2F, gq = (e & (F - 1)) >> src.log2S;
      float2 z = make_float2(0.f, 0.f);
      if (gq < src.A) {
        const int idx = window_index(qb0 + c + (gq << LOG2B), nseg, sseg, src.maxidx);
        if (idx >= 0) z = cscale(src.X[(size_t)(qb0 + c) * Fh + gq], mother_value(idx, a.fb, a.toradians, a.mnorm));
      }
      gstore(U + e, z);
    }
  }
}

// Pass 2 for row y of the batch, f block fblk (16 consecutive f), by the 256
// threads tid of one item (every thread of the work-group reaches its
// barriers).
//   thread (g, j): f = f0 + g, qb = j + 16 k (k < 16) in registers
//   wb[k1][j] = W_256^{j k1}   t2s[k][g] = W_N^{16 k f}   base = c W_N^{j f}
template <int LOG2F>
__device__ __forceinline__ void cols_pass2(const BigArgs& a, int y, int fblk, int tid, float2* lds) {
  constexpr int XS = 16 * 17;  // exchange [k1][g][j]: k1 * XS + g * 17 + j
  float2* const xb = lds;      // exchange, then the tile [m][g] (m * 17 + g)
  float2* const wb = lds + 16 * XS;
  float2* const t2s = wb + 256;
  const RowD r = a.k.rows[a.rows[y]];
  const int N = a.k.N;
  const int f0 = fblk << 4;
  const int g = tid & 15, j = tid >> 4;
  {
    const int k1 = tid >> 4, jj = tid & 15;
    wb[tid] = a.k.tw((unsigned)(k1 * jj) << LOG2F);
    const int k = tid >> 4, gg = tid & 15;  // [k][g]: the 16 lanes of one k read consecutive entries
    t2s[tid] = a.k.tw((unsigned)((16 * k * (f0 + gg)) & (N - 1)));
  }
  const float2 base = cscale(a.k.tw((unsigned)(j * (f0 + g))), a.k.norm);
  const gfloat2* U = (const gfloat2*)(a.scratch + (size_t)y * N);
  float2 v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = gload(U + ((size_t)(j + 16 * k) << LOG2F) + f0 + g);
  __syncthreads();  // tables
#pragma unroll
  for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], t2s[k * 16 + g]);
  dft_reg<16, 1>(v);
  v[0] = cmul(v[0], base);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul(v[k1], cmul(base, wb[k1 * 16 + j]));
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) xb[k1 * XS + g * 17 + j] = v[k1];
  __syncthreads();
  const int k1 = tid >> 4;
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) v[jj] = xb[k1 * XS + g * 17 + jj];
  dft_reg<16, 1>(v);  // over j: output k2, m = k1 + 16 k2
  // lanes g = 0..15 of one k1 hold 16 consecutive f: each store writes
  // 128-byte runs of the row straight from the registers (no transpose
  // through LDS), clipped at t >= n by the row's buffer range
  const __amdgpu_buffer_rsrc_t rr = row_rsrc(row_out(a.k, r), r.n);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) bstore_nt(rr, (unsigned)(f0 + g) + ((unsigned)(k1 + 16 * k2) << LOG2F), v[k2]);
}

// Pass 2 at B = 512 for row y of the batch, f block fblk (16 consecutive f),
// by the work-group's 512 threads.  With qb = j + 32 k and m = m1 + 16 m2:
//   y[f + F m] = c sum_j W_32^{j m2} W_512^{j m1} W_N^{j f}
//                      sum_k W_16^{k m1} W_N^{32 k f} U[qb][f]
// Stage 1, thread (g, j) = (tid & 15, tid >> 4): the 16-point DFT over k.
// Stage 2, thread (g, e, m1) = (tid & 15, (tid >> 4) & 1, tid >> 5): the
// 16-point DFT over j' of the values j = 2 j' + e, E_e[m2'], then the
// radix-2 step between the lane pair e = 0, 1 (lanes 16 apart):
//   y[f + F (m1 + 16 m2' + 256 h)] = E_0[m2'] + (-1)^h W_32^{m2'} E_1[m2'],
// lane e writing h = e.  Lanes g = 0..15 of one (e, m1) write one 128-byte
// run of an output row, so no LDS transpose precedes the stores.
template <int LOG2F>
__device__ __forceinline__ void cols_pass2_b512(const BigArgs& a, int y, int fblk, float2* lds) {
  constexpr int XS = 16 * 33;         // exchange [m1][g][j]: m1 * XS + g * 33 + j
  float2* const xb = lds;
  float2* const wb = lds + 16 * XS;   // [m1][j] = W_512^{j m1}
  float2* const t2s = wb + 512;       // [k][g] = W_N^{32 k f}
  const RowD r = a.k.rows[a.rows[y]];
  const int N = a.k.N;
  const int f0 = fblk << 4;
  const int tid = threadIdx.x;
  const int g = tid & 15;
  {
    const int m1 = tid >> 5, jj = tid & 31;
    wb[tid] = a.k.tw((unsigned)(m1 * jj) << LOG2F);  // W_N^{(N / 512) j m1}
    if (tid < 256) t2s[tid] = a.k.tw((unsigned)((32 * (tid >> 4) * (f0 + g)) & (N - 1)));
  }
  const int j = tid >> 4;
  const float2 base = cscale(a.k.tw((unsigned)(j * (f0 + g))), a.k.norm);
  const gfloat2* U = (const gfloat2*)(a.scratch + (size_t)y * N);
  float2 v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = gload(U + ((size_t)(j + 32 * k) << LOG2F) + f0 + g);
  __syncthreads();  // tables
#pragma unroll
  for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], t2s[k * 16 + g]);
  dft_reg<16, 1>(v);
  v[0] = cmul(v[0], base);
#pragma unroll
  for (int m1 = 1; m1 < 16; ++m1) v[m1] = cmul(v[m1], cmul(base, wb[m1 * 32 + j]));
#pragma unroll
  for (int m1 = 0; m1 < 16; ++m1) xb[m1 * XS + g * 33 + j] = v[m1];
  __syncthreads();
  const int e = (tid >> 4) & 1, m1 = tid >> 5;
#pragma unroll
  for (int jp = 0; jp < 16; ++jp) v[jp] = xb[m1 * XS + g * 33 + 2 * jp + e];
  dft_reg<16, 1>(v);  // E_e[m2']
  const __amdgpu_buffer_rsrc_t rr = row_rsrc(row_out(a.k, r), r.n);
  const bool odd = e != 0;
#pragma unroll
  for (int m2 = 0; m2 < 16; ++m2) {
    float2 x = v[m2];
    if (m2 > 0) {
      const float2 wx = cmul(x, wb[m2 * 32 + 16]);  // W_32^{m2'} = W_512^{16 m2'}
      x = odd ? wx : x;
    }
    const float2 pr = make_float2(lane_xor16(x.x, odd), lane_xor16(x.y, odd));
    const float2 o = odd ? csub(pr, x) : cadd(x, pr);
    bstore_nt(rr, (unsigned)(f0 + g) + ((unsigned)(m1 + 16 * m2 + 256 * e) << LOG2F), o);
  }
}

// Pass 2 alone: grid (F / 16 f blocks, rows in the batch), 256 threads.
template <int LOG2F>
__global__ __launch_bounds__(256) void k_cols_out(BigArgs a) {
  __shared__ float2 lds[ColsLds<LOG2F>::P2_ITEM];
  cols_pass2<LOG2F>(a, blockIdx.y, blockIdx.x, threadIdx.x, lds);
}

// Pass 2 alone at B = 512: grid (F / 16 f blocks, rows in the batch), 512 threads.
template <int LOG2F>
__global__ __launch_bounds__(512) void k_cols_out512(BigArgs a) {
  __shared__ float2 lds[ColsLds<LOG2F, 9>::P2_ITEM];
  cols_pass2_b512<LOG2F>(a, blockIdx.y, blockIdx.x, lds);
}

// One launch of the column path: pass 1 of a batch of rows (p1, n1 rows)
// and pass 2 of the previous batch (p2, n2 rows, in the other scratch set).
// Their work-groups alternate while both kinds remain, so that the
// arithmetic of pass 1 and the memory traffic of pass 2 share the CUs.
// Grid: X1 n1 + X2 n2 work-groups (ColsGrid).
template <int LOG2F, int LOG2B>
__global__ __launch_bounds__(ColsCfg<LOG2F>::T) void k_cols_step(BigArgs p1, int n1, BigArgs p2, int n2) {
  using G = ColsGrid<LOG2F, LOG2B>;
  constexpr int F = 1 << LOG2F;
  __shared__ float2 lds[ColsLds<LOG2F, LOG2B>::SIZE];
  const int w1 = G::X1 * n1, w2 = G::X2 * n2, lo = min(w1, w2);
  const int w = blockIdx.x;
  bool pass1;
  int idx;
  if (w < 2 * lo) {
    pass1 = (w & 1) != 0;
    idx = w >> 1;
  } else {
    pass1 = w1 > w2;
    idx = lo + (w - 2 * lo);
  }
  if (pass1) {
    cols_pass1<LOG2F, LOG2B>(p1, idx / G::X1, idx % G::X1, lds);
  } else if constexpr (LOG2B == 9) {
    cols_pass2_b512<LOG2F>(p2, idx / G::X2, idx % G::X2, lds);
  } else {
    const int sub = threadIdx.x >> 8;
    const int item = idx * G::SUB + sub;
    cols_pass2<LOG2F>(p2, item / (F / 16), item % (F / 16), threadIdx.x & 255,
                      lds + sub * ColsLds<LOG2F, LOG2B>::P2_ITEM);
  }
}

}  // namespace fcwt_hip
