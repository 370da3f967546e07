// rows_direct.h -- the fused window-multiply + band-limited inverse FFT for
// scales whose band fits one work-group (M = 2^LOG2M <= 1024 bins).
// Device code, included by fcwt_hip.hip.
//
// Row y[t] = c * sum_{q<M} Z[q] e^{2 pi i q t / N} is split into L = N/M
// phases l:  y[l + L*m] = c * IDFT_M(Z[q] * W_N^{q*l})[m].  A work-group takes
// G consecutive phases, so for every m it owns G consecutive outputs and the
// row is written in runs of G complex64 (G >= 16: 128-byte runs).
//
// Each phase's M-point transform is M = R1 * P2 (R1 <= 32): P2 threads each
// hold R1 points q = j + P2*k in registers.  With W_N^{q l} =
// W_N^{j l} * W_{R1 L}^{k l}, the second factor is applied before the R1-point
// DFT (a per-block table of R1 x G values) and the first, being constant in k,
// after it, together with the inter-pass twiddle W_M^{j k1}.  One LDS
// exchange, the P2-point DFTs, then an LDS transpose into [m][g] order for the
// stores.
//
// LDS holds, besides the exchange/transpose buffer: the row's windowed band
// Z[0, M) (shared by all G phases), and the tables W_M^{j k1} as [k1][j] and
// W_{R1 L}^{k l} as [k][g], so that the lanes of one instruction read
// consecutive or identical entries.  The exchange is [g][j][k1] with stride
// R1 + 1 per j, which keeps both its writes and reads on distinct banks.
#pragma once
#include "fft_lds.h"
#include "fft_reg.h"

namespace fcwt_hip {

// TEAM: the configuration of a block run inside the wide-band team kernel
// (rows_team.h), whose work-groups have 256 threads and 4768 complex64 values
// of LDS for a block: at M = 1024 a block takes 8 phases there (64-byte
// output runs) instead of 16, and at M = 16 and 32 (one thread per phase)
// 128 and 64 phases, so that the output transpose fits, with the other
// threads idle until the stores.
template <int LOG2M, bool TEAM = false>
struct DirectCfg {
  static constexpr int M = 1 << LOG2M;
  static constexpr int R1 = M < 32 ? M : 32;
  static constexpr int P2 = M / R1;
  // phases per work-group (upper bound; fewer when the row has fewer phases)
  static constexpr int G = P2 >= 16   ? (TEAM && P2 > 16 ? 256 / P2 : 16)
                           : P2 > 1 || !TEAM ? 256 / P2
                                             : (M >= 32 ? 64 : M >= 16 ? 128 : 256);
  static constexpr int NT = TEAM ? 256 : G * P2;  // threads
  static constexpr int XS = M + P2;                // LDS stride between phases in the exchange
  static constexpr int OS = G + 1;                 // LDS stride between m rows in the transpose
  // Tables: Z[0, M), W_M^{j k1}, the pre-twiddles.  With P2 > 1 the exchange
  // (two rounds over halves of the phases) and the output transpose (two
  // rounds over halves of m) reuse their space once they are read.
  static constexpr int TABLES = M + M + R1 * G;
  static constexpr int BUFH = (G / 2) * XS > (M / 2) * OS ? (G / 2) * XS : (M / 2) * OS;
  static constexpr int BUF = P2 > 1 ? 0 : M * OS;  // offset of the tables
  // at M = 256 the exchange holds 16 k1 x G phases x (P2 + 1) values per round
  static constexpr int EXCH = LOG2M == 8 ? 16 * G * (P2 + 1) : 0;
  static constexpr int LDS0 = P2 > 1 ? (BUFH > TABLES ? BUFH : TABLES) : BUF + TABLES;
  static constexpr int LDS = LDS0 > EXCH ? LDS0 : EXCH;
};

// One block of a row: G phases of item `it`, by C::NT threads (the whole
// work-group), in `lds` (C::LDS complex64 values).
template <int LOG2M, bool TEAM>
__device__ __forceinline__ void direct_block(const KArgs& a, const Item it, float2* lds) {
  using C = DirectCfg<LOG2M, TEAM>;
  constexpr int M = C::M, R1 = C::R1, P2 = C::P2, NT = C::NT;
  float2* const buf = lds;
  float2* const zs = lds + C::BUF;     // Z[0, M)
  float2* const wm = zs + M;           // [k1][j] = W_M^{j k1}
  float2* const pre = wm + M;          // [k][g] = W_{R1 L}^{k l}
  const RowD r = a.rows[it.row];
  const int log2L = a.log2N - LOG2M;
  const int L = 1 << log2L;
  const int G = min(C::G, L);  // phases in this block
  const int l0 = it.block * G;
  const int tid = threadIdx.x;
  {
    // Z[q] for this thread's bins: the segment scan runs once for all of
    // them, then every gather is issued before the first one is used
    constexpr int kZ = (M + NT - 1) / NT;
    const float maxidx = (float)a.N - 1.0f;
    int idx[kZ];
#pragma unroll
    for (int u = 0; u < kZ; ++u) idx[u] = -1;
    for (int s = 0; s < r.seg_count; ++s) {
#pragma clang fp contract(off)
      const SegD sg = a.segs[r.seg_begin + s];
#pragma unroll
      for (int u = 0; u < kZ; ++u) {
        const int q = tid + u * NT;
        if (idx[u] < 0 && q >= sg.lo && q < sg.hi) {  // the first segment holding q
          const float v = sg.step * (float)q;
          idx[u] = (int)(v < maxidx ? v : maxidx);
        }
      }
    }
    const float2* const rspec = row_spec(a, r);
    float2 xs[kZ];
    float ws[kZ];
#pragma unroll
    for (int u = 0; u < kZ; ++u) {
      const int q = tid + u * NT;
      const bool in = q < M && idx[u] >= 0;
      xs[u] = in ? rspec[q] : make_float2(0.f, 0.f);
      ws[u] = in ? a.mother[idx[u]] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kZ; ++u) {
      const int q = tid + u * NT;
      if (q < M) {
        zs[q] = make_float2(xs[u].x * ws[u], xs[u].y * ws[u]);
        if (P2 > 1) {
          const int j = q % P2, k1 = q / P2;  // j k1 < M
          wm[q] = a.tw((unsigned)(j * k1) << log2L);
        }
      }
    }
    // W_{R1 L}^{k l} = W_N^{k l P2}
    for (int e = tid; e < R1 * C::G; e += NT) {
      int k = e / C::G, g = e - k * C::G;
      if (g < G) pre[e] = a.tw(((unsigned)(k * (l0 + g)) * P2) & (unsigned)(a.N - 1));
    }
  }
  __syncthreads();
  const int g = tid / P2, j = tid % P2;
  const bool active = g < G;
  float2 v[R1];
  if (active) {
#pragma unroll
    for (int k = 0; k < R1; ++k) v[k] = cmul(zs[j + P2 * k], pre[k * C::G + g]);
    dft_reg<R1, 1>(v);
  }
  if constexpr (P2 > 1) {
    if (active) {
      // W_N^{j l} (constant in k1) times W_M^{j k1}
      const unsigned l = (unsigned)(l0 + g);
      const float2 base = cscale(a.tw((unsigned)(j * l) & (unsigned)(a.N - 1)), a.norm);  // with the normalisation
      v[0] = cmul(v[0], base);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], cmul(base, wm[k1 * P2 + j]));
    }
    __syncthreads();  // the tables are read: the exchange reuses their space
    if constexpr (LOG2M == 9) {
      // M = 512 (R1 = 32, P2 = 16, 16 phases, 256 threads): the exchange in
      // two rounds over the halves of k1, as [k1'][g][j] (stride 17 per g),
      // every thread a writer and a reader in both; the 16-point DFT over j
      // then runs on thread (g2, kr) = (tid & 15, tid >> 4), so that lanes
      // g2 = 0..15 hold 16 consecutive phases: each store writes 128-byte
      // runs of the row straight from the registers (m = 16 h + kr + 32 k2),
      // with no output transpose through LDS
      static_assert(R1 == 32 && P2 == 16 && C::G == 16 && NT == 256, "M = 512 layout");
      constexpr int KS = C::G * 17;
      const int g2 = tid & 15, kr = tid >> 4;
      float2* out = row_out(a, r);
      const __amdgpu_buffer_rsrc_t rr = row_rsrc(out, r.n);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (active) {
#pragma unroll
          for (int k = 0; k < 16; ++k) buf[k * KS + g * 17 + j] = v[16 * h + k];
        }
        __syncthreads();
        float2 w[16];
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) w[jj] = buf[kr * KS + g2 * 17 + jj];
        dft_reg<16, 1>(w);
        if (g2 < G) {
#pragma unroll
          for (int k2 = 0; k2 < 16; ++k2)
            bstore_nt(rr, (unsigned)(l0 + g2) + ((unsigned)(16 * h + kr + 32 * k2) << log2L), w[k2]);
        }
        if (h == 0) __syncthreads();  // round 0's reads are done before round 1's writes
      }
    } else if constexpr (LOG2M == 8) {
      // M = 256 (R1 = 32, P2 = 8, 32 phases, 256 threads, in either kind of
      // launch): the two rounds over the halves of k1, as [k1'][g][j]
      // (stride 9 per g); thread (g2, gb, kr) = (tid & 15, (tid >> 4) & 1,
      // tid >> 5) runs the 8-point DFTs over j of phase 16 gb + g2 for
      // k1' = kr and kr + 8, and its lanes g2 = 0..15 store 128-byte runs
      // from the registers (m = 16 h + k1' + 32 k2)
      static_assert(R1 == 32 && P2 == 8 && C::G == 32 && NT == 256, "M = 256 layout");
      constexpr int GS = P2 + 1, KS = C::G * GS;
      static_assert(16 * KS <= C::LDS, "exchange fits the block's LDS");
      const int g2 = tid & 15, gb = (tid >> 4) & 1, kr = tid >> 5;
      const int gp = 16 * gb + g2;
      float2* out = row_out(a, r);
      const __amdgpu_buffer_rsrc_t rr = row_rsrc(out, r.n);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (active) {
#pragma unroll
          for (int k = 0; k < 16; ++k) buf[k * KS + g * GS + j] = v[16 * h + k];
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          const int kk = kr + 8 * d;
          float2 w[8];
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) w[jj] = buf[kk * KS + gp * GS + jj];
          dft_reg<8, 1>(w);
          if (gp < G) {
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2)
              bstore_nt(rr, (unsigned)(l0 + gp) + ((unsigned)(16 * h + kk + 32 * k2) << log2L), w[k2]);
          }
        }
        if (h == 0) __syncthreads();  // round 0's reads are done before round 1's writes
      }
    } else if constexpr (LOG2M == 10) {
      // M = 1024 (R1 = 32, P2 = 32; 16 phases and 512 threads, or 8 phases
      // and 256 threads in the team launch): the same two rounds over the
      // halves of k1, as [k1'][g][j] (stride 33 per g).  The 32-point DFT
      // over j of one (g, k1) runs on the lane pair e = 0, 1 (lanes 16 apart)
      // of thread (g2, e, kr), g2 the low lane bits: a 16-point DFT over the
      // j = 2 j' + e, E_e[m2'], then the radix-2 step through
      // v_permlane16_swap,
      //   Y[m2' + 16 e] = E_0[m2'] + (-1)^e W_32^{m2'} E_1[m2'],
      // lane e writing its own e; m = k1 + 32 (m2' + 16 e), and lanes
      // g2 = 0 .. G-1 store runs of G phases from the registers (128 bytes,
      // or 64 in the team launch).  Both layouts do the same arithmetic per
      // phase, so their outputs are identical.
      constexpr int GS = C::G;
      static_assert(R1 == 32 && P2 == 32 && (GS == 16 || (TEAM && GS == 8)) && NT == 32 * GS, "M = 1024 layout");
      constexpr int JS = 33, KS = GS * JS;
      static_assert(16 * KS + 16 <= C::LDS, "exchange and W_32 table fit the block's LDS");
      float2* const w32 = buf + 16 * KS;  // W_32^m, m < 16 (= W_N^{32 m L})
      if (tid < 16) w32[tid] = a.tw((unsigned)(32 * tid) << log2L);
      const int g2 = tid & (GS - 1), e = (tid >> 4) & 1;
      const int kr = GS == 16 ? tid >> 5 : ((tid >> 3) & 1) + 2 * (tid >> 5);
      const bool odd = e != 0;
      float2* out = row_out(a, r);
      const __amdgpu_buffer_rsrc_t rr = row_rsrc(out, r.n);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (active) {
#pragma unroll
          for (int k = 0; k < 16; ++k) buf[k * KS + g * JS + j] = v[16 * h + k];
        }
        __syncthreads();
        float2 w[16];
#pragma unroll
        for (int jp = 0; jp < 16; ++jp) w[jp] = buf[kr * KS + g2 * JS + 2 * jp + e];
        dft_reg<16, 1>(w);  // E_e[m2']
#pragma unroll
        for (int m2 = 0; m2 < 16; ++m2) {
          float2 x = w[m2];
          if (m2 > 0) {
            const float2 wx = cmul(x, w32[m2]);
            x = odd ? wx : x;
          }
          const float2 pr = make_float2(lane_xor16(x.x, odd), lane_xor16(x.y, odd));
          w[m2] = odd ? csub(pr, x) : cadd(x, pr);
        }
        if (g2 < G) {
#pragma unroll
          for (int m2 = 0; m2 < 16; ++m2) {
            const unsigned t = (unsigned)(l0 + g2) + ((unsigned)(16 * h + kr + 32 * (m2 + 16 * e)) << log2L);
            // 64-byte runs (team launch) without the non-temporal hint, so
            // that the L2 joins the runs of neighbouring blocks
            if constexpr (GS < 16)
              bstore(rr, t, w[m2]);
            else
              bstore_nt(rr, t, w[m2]);
          }
        }
        if (h == 0) __syncthreads();  // round 0's reads are done before round 1's writes
      }
    } else {
      // exchange in two rounds, phases [0, G/2) then [G/2, G): g = tid / P2,
      // so each round is a whole set of wavefronts
      constexpr int GH = C::G / 2;
  #pragma unroll
      for (int h = 0; h < 2; ++h) {
        const bool mine = active && g >= h * GH && g < (h + 1) * GH;
        const int gl = g - h * GH;
        if (mine) {
  #pragma unroll
          for (int k1 = 0; k1 < R1; ++k1) buf[gl * C::XS + j * (R1 + 1) + k1] = v[k1];
        }
        __syncthreads();
        if (mine) {
          constexpr int U = R1 / P2;
  #pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k1 = u * P2 + j;
            float2 w[P2];
  #pragma unroll
            for (int jj = 0; jj < P2; ++jj) w[jj] = buf[gl * C::XS + jj * (R1 + 1) + k1];
            dft_reg<P2, 1>(w);
  #pragma unroll
            for (int k2 = 0; k2 < P2; ++k2) v[u * P2 + k2] = w[k2];
          }
        }
        __syncthreads();
      }
      // transpose to [m][g] in two rounds over m = k1 + R1 k2 (k2 < P2/2, then
      // the rest), each followed by its 128-byte output runs
      float2* out = row_out(a, r);
      constexpr int MH = M / 2, KH = P2 / 2;
  #pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (active) {
          constexpr int U = R1 / P2;
  #pragma unroll
          for (int u = 0; u < U; ++u)
  #pragma unroll
            for (int k2 = 0; k2 < KH; ++k2) buf[((u * P2 + j) + R1 * k2) * C::OS + g] = v[u * P2 + h * KH + k2];
        }
        __syncthreads();
        if (G == C::G) {
          // every LDS read first, then the stores, clipped at t >= n by the
          // row's buffer range
          constexpr int S = MH * C::G / NT;
          float2 o[S];
  #pragma unroll
          for (int s = 0; s < S; ++s) {
            const int e = s * NT + tid;
            o[s] = buf[(e / C::G) * C::OS + e % C::G];
          }
          const __amdgpu_buffer_rsrc_t rr = row_rsrc(out, r.n);
  #pragma unroll
          for (int s = 0; s < S; ++s) {
            const int e = s * NT + tid;
            // runs shorter than 128 bytes (M = 1024 in the team launch) go
            // without the non-temporal hint, so that the L2 joins the runs of
            // neighbouring blocks (1.3% of the 384-channel batch, measured)
            if constexpr (TEAM && C::G < 16)
              bstore(rr, (unsigned)(l0 + e % C::G) + ((unsigned)(e / C::G + h * MH) << log2L), o[s]);
            else
              bstore_nt(rr, (unsigned)(l0 + e % C::G) + ((unsigned)(e / C::G + h * MH) << log2L), o[s]);
          }
        } else {
          for (int e = tid; e < MH * G; e += NT) {
            const int mm = e / G, gg = e - mm * G;
            const long long t = (long long)(l0 + gg) + ((long long)(mm + h * MH) << log2L);
            if (t < r.n) gstore_nt((gfloat2*)(out + t), buf[mm * C::OS + gg]);
          }
        }
        if (h == 0) __syncthreads();
      }
    }
  } else {
    // single thread per phase: only the phase factor W_N^{q l} = W_{R1 L}^{k l}
    // applies (j = 0); transpose through LDS to [m][g]
    if (active) {
#pragma unroll
      for (int m = 0; m < R1; ++m) buf[m * C::OS + g] = cscale(v[m], a.norm);
    }
    __syncthreads();
    float2* out = row_out(a, r);
    const int tot = M * G;
    for (int e = tid; e < tot; e += NT) {
      int m = e / G, gg = e - m * G;
      long long t = (long long)(l0 + gg) + ((long long)m << log2L);
      if (t < r.n) gstore_nt((gfloat2*)(out + t), buf[m * C::OS + gg]);
    }
  }
}

template <int LOG2M>
__global__ __launch_bounds__(DirectCfg<LOG2M>::NT) void k_rows_direct(KArgs a) {
  __shared__ float2 lds[DirectCfg<LOG2M>::LDS];
  direct_block<LOG2M, false>(a, a.items[blockIdx.x], lds);
}

}  // namespace fcwt_hip
