// rows_team.h -- band-limited inverse FFT for scales whose band is wider than
// one work-group can transform (M > 1024 bins), by teams of work-groups on
// one XCD exchanging the intermediate through that XCD's L2.
// Device code, included by fcwt_hip.hip.
//
// With N = B * F (B = 256) and the band M = A * B:
//   y[f + F*m] = c * sum_{qb<B} W_B^{qb m} U[qb][f]
//   U[qb][f]   = W_N^{qb f} * sum_{qf<A} Z[qb + B qf] W_F^{qf f}
// Pass 1 (per team member: Qb consecutive qb): the F-point sums, done as F/A
// phases l' of A-point inverse FFTs (f = l' + (F/A) m'), into a scratch
// buffer U[f][qb] in global memory.  Pass 2 (per member: 16 consecutive f):
// B-point inverse FFTs over qb, written as 16-element (128-byte) runs.
//
// All members of a team run on one XCD (chosen by the hardware XCC id), so
// the scratch never leaves that XCD's L2: pass 1 stores normally, waits for
// its stores (s_waitcnt), then bumps a per-buffer counter; pass 2 waits for
// the counter and reads with device-scope loads, which bypass the CU's L1 and
// are served by the shared L2.  Two scratch buffers per team let pass 1 of
// one row overlap pass 2 of the previous row.  Every wait is bounded: on
// timeout the kernel records an error and every work-group exits.
#pragma once
#include "fft_lds.h"
#include "fft_reg.h"

namespace fcwt_hip {

constexpr int kTeamLog2B = 8;                  // B = 256
constexpr int kTeamB = 1 << kTeamLog2B;
constexpr int kTeamThreads = 256;
constexpr int kTeamPoints = 4096;              // points per member per pass
constexpr int kTeamBuffers = 2;
constexpr int kTeamMaxLog2F = 9;               // F <= 512, N <= 2^17
constexpr int kTeamMinLog2F = 7;               // F >= 128, N >= 2^15
constexpr int kTeamSegLds = 32;                // row segments staged in LDS
constexpr long kTeamSpin = 1L << 22;           // polls (with s_sleep) before a wait is declared failed

struct TeamArgs {
  KArgs k;
  const int* team_rows;    // rows of team t: team_rows[t + nteams * i], i < team_count[t]
  const int* team_count;
  int nteams;              // teams on the whole device
  int team_size;           // work-groups per team
  int teams_per_xcd;
  int log2F;
  float2* scratch;         // [team][buffer][N]
  unsigned* sync;          // zeroed before launch; see below
  unsigned* error;         // set to a nonzero code on a failed wait or team formation
};
// sync layout (unsigned, 32 per 128-byte line):
//   [0..7] work-groups per XCD, [8] grid arrival counter,
//   team t: arrive[b] at 64 * (1 + t * kTeamBuffers + b), done[b] 32 later.

__device__ __forceinline__ unsigned ld_agent(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_agent(unsigned* p, unsigned v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Device-scope load of one complex64 (global_load_dwordx2 ... sc1): misses the
// CU's L1, served by the XCD's L2.
__device__ __forceinline__ float2 ld_agent2(const gfloat2* p) {
  typedef const __attribute__((address_space(1))) unsigned long long gu64;
  unsigned long long u = __hip_atomic_load(reinterpret_cast<gu64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float2(__uint_as_float((unsigned)u), __uint_as_float((unsigned)(u >> 32)));
}

// Thread 0 waits until *p >= target (bounded); the result is broadcast.
__device__ __forceinline__ bool team_wait(const unsigned* p, unsigned target, int* s_flag, unsigned* error) {
  if (threadIdx.x == 0) {
    long s = 0;
    while (ld_agent(p) < target && ld_agent(error) == 0 && s < kTeamSpin) {
      __builtin_amdgcn_s_sleep(1);
      ++s;
    }
    bool ok = ld_agent(p) >= target;
    if (!ok) atomicOr(error, 1u);
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

// Window factor mother[idx(q)] of the segment holding q (0 when none), with
// the row's segments staged in LDS (fcwt.cpp:224-248).
__device__ __forceinline__ float window_at(int q, int nseg, const SegD* sseg, const float* __restrict__ mother,
                                           float maxidx) {
#pragma clang fp contract(off)
  int hit = -1;
  for (int s = 0; s < nseg; ++s)
    if (q >= sseg[s].lo && q < sseg[s].hi) hit = s;
  if (hit < 0) return 0.0f;
  float v = sseg[hit].step * (float)q;
  int idx = (int)(v < maxidx ? v : maxidx);
  return mother[idx];
}

// Pass 1 for one row and the member's Qb = 4096/F qb values, A = 2^LOG2A.
// zs holds the row's Z[qb0 + ql + B qf] at [ql * A + qf].
// The passes are separate (not inlined) functions: inlined into the row loop
// together they exceed the register file and spill.
template <int LOG2A>
__device__ __noinline__ void team_pass1(int log2F, int qb0, gfloat2* __restrict__ U, const float2* zs,
                                        const float2* wf, const float2* wqf, float2* buf) {
  constexpr int A = 1 << LOG2A;
  constexpr int R1 = A < 16 ? A : 16;
  constexpr int P2 = A / R1;
  static_assert(P2 <= R1, "A too large for two passes");
  const int F = 1 << log2F;
  const int log2Ph = log2F - LOG2A;  // phases l' per qb
  const int log2Qb = 12 - log2F;
  const int ntr = (1 << log2Qb) << log2Ph;  // transforms
  const int rounds = (ntr * P2 + kTeamThreads - 1) / kTeamThreads;
  for (int rd = 0; rd < rounds; ++rd) {
    const int e = rd * kTeamThreads + threadIdx.x;
    const int tr = e / P2, j = e % P2;
    const bool active = tr < ntr;
    const int ql = tr >> log2Ph, lp = tr & ((1 << log2Ph) - 1);
    float2 v[R1];
    if (active) {
#pragma unroll
      for (int k = 0; k < R1; ++k) {
        const int qf = j + P2 * k;
        v[k] = cmul(zs[ql * A + qf], wf[(qf * lp) & (F - 1)]);
      }
      dft_reg<R1, 1>(v);
    }
    if constexpr (P2 > 1) {
      // inter-pass twiddle W_A^{j k1} = W_F^{j k1 F/A}, exchange, P2-point DFTs
      constexpr int XS = R1 * (P2 + 1);
      if (active) {
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], wf[(j * k1) << log2Ph]);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) buf[tr * XS + k1 * (P2 + 1) + j] = v[k1];
      }
      __syncthreads();
      constexpr int Uu = R1 / P2;
      if (active) {
#pragma unroll
        for (int u = 0; u < Uu; ++u) {
          const int k1 = j * Uu + u;
          float2 w2[P2];
#pragma unroll
          for (int jj = 0; jj < P2; ++jj) w2[jj] = buf[tr * XS + k1 * (P2 + 1) + jj];
          dft_reg<P2, 1>(w2);
#pragma unroll
          for (int k2 = 0; k2 < P2; ++k2) v[u * P2 + k2] = w2[k2];
        }
      }
      __syncthreads();
      // results to the LDS tile [f][ql] (output m = k1 + R1 k2, f = lp + m F/A)
      if (active) {
#pragma unroll
        for (int u = 0; u < Uu; ++u)
#pragma unroll
          for (int k2 = 0; k2 < P2; ++k2) {
            const int f = lp + (((j * Uu + u) + R1 * k2) << log2Ph);
            buf[(f << log2Qb) + ql] = cmul(v[u * P2 + k2], wqf[ql * F + f]);
          }
      }
    } else {
      if (active) {
#pragma unroll
        for (int m = 0; m < R1; ++m) {
          const int f = lp + (m << log2Ph);
          buf[(f << log2Qb) + ql] = cmul(v[m], wqf[ql * F + f]);
        }
      }
    }
  }
  __syncthreads();
  // coalesced stores of U[f][qb0 .. qb0 + Qb)
  for (int e = threadIdx.x; e < kTeamPoints; e += kTeamThreads) {
    const int f = e >> log2Qb, ql = e & ((1 << log2Qb) - 1);
    gstore(U + f * kTeamB + qb0 + ql, buf[e]);
  }
}

// Pass 2 for one row: 16 consecutive f starting at f0, 256-point inverse FFTs
// over qb, written to out (the row) scaled by norm.
__device__ __noinline__ void team_pass2(gfloat2* __restrict__ out, int n, float norm, int log2F, int f0,
                                        const gfloat2* U, const float2* wb, float2* buf) {
  constexpr int R1 = 16, P2 = 16, XS = R1 * (P2 + 1), OS = 17;
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  float2 v[R1];
#pragma unroll
  for (int k = 0; k < R1; ++k) v[k] = ld_agent2(U + (size_t)(f0 + g) * kTeamB + j + P2 * k);
  dft_reg<R1, 1>(v);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], wb[j * k1]);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) buf[g * XS + k1 * (P2 + 1) + j] = v[k1];
  __syncthreads();
  float2 w[P2];
#pragma unroll
  for (int jj = 0; jj < P2; ++jj) w[jj] = buf[g * XS + j * (P2 + 1) + jj];
  dft_reg<P2, 1>(w);
  __syncthreads();
  // m = j + 16 k2; transpose to [m][g]
#pragma unroll
  for (int k2 = 0; k2 < P2; ++k2) buf[(j + R1 * k2) * OS + g] = cscale(w[k2], norm);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < (kTeamB * 16) / kTeamThreads; ++s) {
    const int e = s * kTeamThreads + threadIdx.x;
    const int m = e >> 4, gg = e & 15;
    const long long t = (long long)(f0 + gg) + ((long long)m << log2F);
    if (t < n) gstore(out + t, buf[m * OS + gg]);
  }
}

__global__ __launch_bounds__(kTeamThreads) void k_rows_team(TeamArgs ta) {
  constexpr int BUF = kTeamB * 17;  // pass-2 exchange / transpose
  // pass-1 exchange: (4096/A) transforms x 16 * (A/16 + 1) entries, largest at A = 32
  constexpr int BUF1 = kTeamPoints + kTeamPoints / 2;
  constexpr int BUFMAX = BUF > BUF1 ? BUF : BUF1;
  __shared__ float2 buf[BUFMAX];
  __shared__ float2 wf[1 << kTeamMaxLog2F];  // W_F^r
  __shared__ float2 wqf[kTeamPoints];        // W_N^{qb f} for the member's qb
  __shared__ float2 xs[kTeamPoints / 2];     // X[qb + B qf], qf < F/2
  __shared__ float2 zs[kTeamPoints / 2];     // the row's Z on the member's bins
  __shared__ float2 wb[kTeamB];              // W_B^r
  __shared__ SegD sseg[kTeamSegLds];
  __shared__ int s_rank, s_flag;
  // plain copies of the arguments (never take the address of the kernel argument)
  const float2* __restrict__ spec = ta.k.spec;
  const float* __restrict__ mother = ta.k.mother;
  const SegD* __restrict__ segs = ta.k.segs;
  const RowD* __restrict__ rows = ta.k.rows;
  float2* __restrict__ out = ta.k.out;
  const Twiddles tw = ta.k.tw;
  const int n = ta.k.n, N = ta.k.N;
  const float norm = ta.k.norm;
  const int* __restrict__ team_rows = ta.team_rows;
  const int nteams = ta.nteams, team_size = ta.team_size;
  const int log2F = ta.log2F;
  unsigned* const sync = ta.sync;
  unsigned* const error = ta.error;

  // team formation: rank among the work-groups of this XCD
  const int xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7;
  if (threadIdx.x == 0) {
    s_rank = (int)atomicAdd(&sync[xcc], 1u);
    add_agent(&sync[8], 1u);
  }
  __syncthreads();
  if (!team_wait(&sync[8], gridDim.x, &s_flag, error)) return;
  if (threadIdx.x == 0) {
    bool ok = true;
    for (int x = 0; x < 8; ++x) ok &= ld_agent(&sync[x]) == (unsigned)(team_size * ta.teams_per_xcd);
    if (!ok) atomicOr(error, 2u);
    s_flag = ok;
  }
  __syncthreads();
  if (!s_flag) return;
  const int team = xcc * ta.teams_per_xcd + s_rank / team_size;
  const int member = s_rank % team_size;
  const int F = 1 << log2F;
  const int log2Qb = 12 - log2F;
  const int items = F >> 4;  // pass-1 qb blocks == pass-2 f blocks per row
  // Each member owns exactly one pass-1 qb block and one pass-2 f block
  // (items == team_size for F <= 512), so its tables are filled once.
  const int item = member;
  const bool has_item = item < items;
  const int qb0 = item << log2Qb;
  for (int e = threadIdx.x; e < F; e += kTeamThreads) wf[e] = tw((unsigned)e << kTeamLog2B);
  for (int e = threadIdx.x; e < kTeamB; e += kTeamThreads) wb[e] = tw((unsigned)e << log2F);
  if (has_item) {
    for (int e = threadIdx.x; e < kTeamPoints; e += kTeamThreads) {
      const int ql = e >> log2F, f = e & (F - 1);
      wqf[e] = tw((unsigned)((qb0 + ql) * f));
    }
    for (int e = threadIdx.x; e < kTeamPoints / 2; e += kTeamThreads) {
      const int ql = e >> (log2F - 1), qf = e & ((F >> 1) - 1);
      xs[e] = spec[qb0 + ql + (qf << kTeamLog2B)];
    }
  }
  const int nrows = ta.team_count[team];
  float2* const scratch = ta.scratch + (size_t)team * kTeamBuffers * N;
  unsigned* const arrive = sync + 64 * (1 + team * kTeamBuffers);
  for (int it = 0; it <= nrows; ++it) {
    if (it < nrows) {
      const RowD r = rows[team_rows[team + nteams * it]];
      const int b = it % kTeamBuffers, gen = it / kTeamBuffers;
      __syncthreads();  // previous users of sseg and buf are done
      const int nseg = r.seg_count;  // <= kTeamSegLds: the schedule routes rows with more elsewhere
      if ((int)threadIdx.x < nseg) sseg[threadIdx.x] = segs[r.seg_begin + threadIdx.x];
      if (gen > 0) {
        if (!team_wait(arrive + 64 * b + 32, (unsigned)(team_size * gen), &s_flag, error)) return;
      } else {
        __syncthreads();
      }
      gfloat2* U = (gfloat2*)(scratch + (size_t)b * N);
      if (has_item) {
        // the row's windowed band on this member's bins, into LDS
        const int log2A = r.log2A, A = 1 << log2A, half = F >> 1;
        const float maxidx = (float)N - 1.0f;
        for (int e = threadIdx.x; e < (1 << log2Qb) * A; e += kTeamThreads) {
          const int ql = e >> log2A, qf = e & (A - 1);
          const float w = window_at(qb0 + ql + (qf << kTeamLog2B), nseg, sseg, mother, maxidx);
          const float2 x = xs[ql * half + qf];
          zs[e] = make_float2(x.x * w, x.y * w);
        }
        __syncthreads();
        switch (log2A) {
          case 1: team_pass1<1>(log2F, qb0, U, zs, wf, wqf, buf); break;
          case 2: team_pass1<2>(log2F, qb0, U, zs, wf, wqf, buf); break;
          case 3: team_pass1<3>(log2F, qb0, U, zs, wf, wqf, buf); break;
          case 4: team_pass1<4>(log2F, qb0, U, zs, wf, wqf, buf); break;
          case 5: team_pass1<5>(log2F, qb0, U, zs, wf, wqf, buf); break;
          case 6: team_pass1<6>(log2F, qb0, U, zs, wf, wqf, buf); break;
          case 7: team_pass1<7>(log2F, qb0, U, zs, wf, wqf, buf); break;
          default: team_pass1<8>(log2F, qb0, U, zs, wf, wqf, buf); break;
        }
      }
      // the pass-1 stores are complete (s_waitcnt) before the arrival is counted
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __syncthreads();
      if (threadIdx.x == 0) add_agent(arrive + 64 * b, 1u);
    }
    if (it >= 1) {
      const int prev = it - 1;
      const RowD r = rows[team_rows[team + nteams * prev]];
      const int b = prev % kTeamBuffers, gen = prev / kTeamBuffers;
      if (!team_wait(arrive + 64 * b, (unsigned)(team_size * (gen + 1)), &s_flag, error)) return;
      if (has_item)
        team_pass2((gfloat2*)(out + r.out_off), n, norm, log2F, item << 4, (const gfloat2*)(scratch + (size_t)b * N),
                   wb, buf);
      // every read of this buffer has landed in registers (they fed the LDS
      // exchange), so it may be reused; the output stores need not complete
      __syncthreads();
      if (threadIdx.x == 0) add_agent(arrive + 64 * b + 32, 1u);
    }
  }
}

}  // namespace fcwt_hip
