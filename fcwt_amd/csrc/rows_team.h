// rows_team.h -- band-limited inverse FFT for scales whose band is wider than
// one work-group can transform (M > 1024 bins), by teams of work-groups on
// one XCD exchanging the intermediate through that XCD's L2.
// Device code, included by fcwt_hip.hip.
//
// With N = B * F (B = 256 for N = 2^15 .. 2^17, B = 512 and F = 512 for
// N = 2^18) and the band M = A * B:
//   y[f + F*m] = c * sum_{qb<B} W_B^{qb m} W_N^{qb f} U[f][qb]
//   U[f][qb]   = sum_{qf<A} Z[qb + B qf] W_F^{qf f}
// A pass-1 item (Qb = 4096/F consecutive qb): the F-point sums, done as F/A
// phases l' of A-point inverse FFTs (f = l' + (F/A) m'), into a scratch
// buffer U[f][qb] in global memory.  A pass-2 item (4096/B consecutive f):
// the factor W_N^{qb f} and the B-point inverse FFTs over qb, written as
// runs of 16 (B = 256, 128 bytes) or 8 (B = 512, 64 bytes) elements of the
// output row.  A row is N/4096 items of each pass.
//
// One team per XCD: all work-groups the hardware places on that XCD (found
// from the XCC id), so the scratch stays in the XCD's L2.  The team walks a
// fixed sequence of item blocks -- step s holds pass 1 of row s and pass 2 of
// row s - lag.  Members claim items in sequence order from the team's
// counter (an item claims its member's next one when it starts, so that the
// claim's latency is hidden), which balances uneven items across members.
// Every item depends only on earlier ones, and the smallest unfinished item
// is always some running member's current item (items are claimed in order,
// and only by running work-groups), so the walk progresses with however many
// members are resident: the launch is an ordinary one, with no grid-wide
// barrier and no cooperative-launch resource.  A work-group's team is the XCD
// it runs on; when the last work-group exits it checks that every team's
// sequence was claimed to its end (an XCD that received no work-group would
// leave its rows unwritten; that is reported as an error).  A pass-1 item stores normally, waits
// for its stores (s_waitcnt), then bumps its buffer's arrival counter; a
// pass-2 item waits for the row's count and reads with device-scope loads,
// which bypass the CU's L1 and are served by the shared L2, then bumps the
// buffer's done counter, which the next pass 1 into that buffer waits for
// (after staging its band in LDS, so that those loads overlap the wait).
// nbuf buffers per team give the sequence slack on both sides.  Every wait is
// bounded: on timeout the kernel records an error in the launch's abort word
// (every other wait of the launch then gives up too) and in the plan's
// report word, every work-group exits, and the last one to exit fills the
// launch's team rows with NaN.  The next launch starts with a clear abort word.
//
// Twiddle tables in LDS are laid out so that the lanes of one instruction
// read consecutive or identical entries (no bank conflicts): pass-2 factors
// W_B^{j k1} as [k1][j]; pass-1 factors as [k][l'] and [k1][l'][j]; padded
// strides on the staged band, the exchanges and the tile.
#pragma once
#include "fft_lds.h"
#include "fft_reg.h"

namespace fcwt_hip {

constexpr int kTeamLog2B = 8;                  // B = 256
constexpr int kTeamB = 1 << kTeamLog2B;
constexpr int kTeamThreads = 256;
constexpr int kTeamWgPerCu = 3;                // work-groups of the team on every CU
constexpr int kTeamWgPerXcd = 32 * kTeamWgPerCu;
constexpr int kTeamPoints = 4096;              // points per item
constexpr int kTeamBuffers = 12;               // most scratch buffers per team
constexpr size_t kTeamScratchPerXcd = 6u << 20;  // scratch bytes per team (default buffer count)
constexpr int kTeamMaxLog2F = 9;               // F <= 512
constexpr int kTeamMinLog2F = 7;               // F >= 128, N >= 2^15
constexpr int kTeamMaxLog2N = 18;              // N = 2^18 splits as B = 512, F = 512
// log2 B of the team split of N = 2^log2N: 256 up to 2^17, 512 at 2^18 (so
// that F, and with it pass 1's tables and tile, stays within 512)
__host__ __device__ constexpr int team_log2b(int log2N) { return log2N > 17 ? 9 : kTeamLog2B; }
constexpr int kTeamSegLds = 32;                // most band segments of a row on the team and column paths
constexpr long kTeamSpin = 1L << 22;           // default polls (with s_sleep) before a wait is declared failed
constexpr int kFillMinLog2M = 0;               // smallest band of a direct row the team kernel runs
constexpr int kFillGroups = 10 - kFillMinLog2M + 1;  // bands 2^0 .. 2^10

// One block of a team's item sequence with what its items read about their
// row, so that an item starts after one burst of scalar loads.
struct TeamBlk {
  int sq;                    // 2 i + pass: row i of the team, pass 0 or 1
  int seg_begin, seg_count;  // the row's band segments in KArgs::segs
  int log2A;                 // band M = B A
  int sig, n;                // signal of the row and its length
  long long out_off;         // offset of the row in its signal's output
  SegD seg[2];               // the first two segments (any further ones stay in segs)
};
static_assert(sizeof(TeamBlk) == 64, "TeamBlk is read as 16 dwords");

struct TeamArgs {
  KArgs k;
  const TeamBlk* team_desc;  // item blocks of team t: team_desc[t * seq_stride + s]
  const int* team_blocks;  // blocks in team t's sequence
  int seq_stride;
  int nteams;              // teams on the whole device (one per XCD)
  int log2F;               // F = N / B (the kernel's template parameter)
  int nbuf;                // scratch buffers per team (2..kTeamBuffers)
  long spin;               // polls before a wait is declared failed (kTeamSpin unless a test lowers it)
  float2* scratch;         // [team][nbuf][N]
  unsigned* sync;          // zeroed before launch; see below
  unsigned* error;         // sticky report: ORed with the launch's abort code, cleared by the host when read
  // Blocks of the rows whose band fits one work-group (2^kFillMinLog2M <=
  // M <= 2^kMaxLog2B), run by direct_block once a member's team sequence is
  // claimed to its end: fill_items[0, nfill), those with band 2^(kFillMinLog2M
  // + g) before fill_end[g].
  const Item* fill_items;
  int nfill;
  int fill_end[kFillGroups];
};
// sync layout (unsigned, 32 per 128-byte line):
//   [0] the next fill block,
//   [32] this launch's abort code, [33] work-groups that have finished,
//   team t: arrive[b] at 64 * (1 + t * kTeamBuffers + b), done[b] 32 later;
//   team t's next-item counter at 64 * (1 + nteams * kTeamBuffers + t).
constexpr int kSyncFill = 0;
constexpr int kSyncAbort = 32;
constexpr int kSyncExited = 33;

// Load through the constant address space. The schedule arrays are written
// before the launch and never during it, so a uniform address becomes a
// scalar load instead of a vector load and a readfirstlane.
__device__ __forceinline__ int ld_const(const int* p) {
  return *(const __attribute__((address_space(4))) int*)(p);
}

// Waits until this thread's global stores are complete, i.e. written to the
// L2 of its XCD, which the other members of the team read through.  (A
// workgroup-scope release fence does not wait for global stores here: the
// work-group's waves share one CU.)
__device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ unsigned ld_agent(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_agent(unsigned* p, unsigned v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Device-scope load of one complex64 (global_load_dwordx2 ... sc1): misses the
// CU's L1, served by the XCD's L2.
__device__ __forceinline__ float2 ld_agent2(const gfloat2* p) {
  unsigned long long u = __hip_atomic_load(reinterpret_cast<const gu64*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float2(__uint_as_float((unsigned)u), __uint_as_float((unsigned)(u >> 32)));
}

// Orders this wavefront's LDS writes before its later LDS reads of other
// lanes' values: the LDS executes one wavefront's instructions in order, so
// only the compiler must be kept from moving memory operations across.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  __builtin_amdgcn_sched_barrier(0);  // and keeps the register DFTs on either side apart
}

// Records a failure: the launch's abort word (every wait of this launch
// gives up on it) and the sticky report word the host reads.
__device__ __forceinline__ void team_fail(unsigned* abort, unsigned* error, unsigned code) {
  atomicOr(abort, code);
  atomicOr(error, code);
}

// Thread 0 waits until *p >= target (at most `spin` polls, and not after the
// launch has aborted); the result is broadcast.  `first` is a value of *p
// thread 0 read earlier in the item (its load overlapped the item's own
// loads); the counter only grows, so when it already meets the target no
// further poll is made.  spin == 0 fails every wait at once: the host sets it
// only to test the failure path.
__device__ __forceinline__ bool team_wait(const unsigned* p, unsigned target, unsigned first, int* s_flag,
                                          unsigned* abort, unsigned* error, long spin) {
  if (threadIdx.x == 0) {
    long s = 0;
    unsigned v = first;
    while (v < target && ld_agent(abort) == 0 && s < spin) {
      __builtin_amdgcn_s_sleep(1);
      v = ld_agent(p);
      ++s;
    }
    bool ok = spin > 0 && v >= target;
    if (!ok) team_fail(abort, error, 1u);
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

// A work-group's exit.  Its stores are complete before it counts itself out
// of the launch (ctl[kSyncExited]), so the last work-group to finish knows
// that no other writes any more: true, for every thread, in that one.  It
// then checks that every team claimed its whole sequence
// (team_check_claims), and when the launch has aborted (or a team's sequence
// was left unclaimed) it fills every team row of the launch with NaN
// (team_nan_rows): a failed launch never leaves rows that look valid.
__device__ __noinline__ bool team_exit_last(unsigned* ctl, int* s_flag) {
  wait_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    __atomic_thread_fence(__ATOMIC_RELEASE);
    const unsigned done = __hip_atomic_fetch_add(ctl + kSyncExited, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    *s_flag = done + 1 == gridDim.x;
  }
  __syncthreads();
  return *s_flag != 0;
}

// Thread 0 of the last work-group: every team of one plan's sequences was
// claimed to its end (an XCD that received no work-group would leave its
// rows unwritten); a shortfall aborts the launch with code 2.
__device__ __noinline__ void team_check_claims(const unsigned* sync, unsigned* abort, unsigned* error,
                                               const int* team_blocks, int nteams, int log2I) {
  for (int t = 0; t < nteams && ld_agent(abort) == 0; ++t) {
    const unsigned claimed = ld_agent(sync + 64 * (1 + nteams * kTeamBuffers + t));
    if (claimed < ((unsigned)team_blocks[t] << log2I)) team_fail(abort, error, 2u);
  }
}

// Every thread of the last work-group of an aborted launch: one plan's team
// rows filled with NaN.
__device__ __noinline__ void team_nan_rows(const TeamBlk* team_desc, const int* team_blocks, int seq_stride,
                                           int nteams, float2* out1, float2* const* outs) {
  const float qnan = __int_as_float(0x7fc00000);
  for (int t = 0; t < nteams; ++t) {
    const int nb = team_blocks[t];
    for (int s = 0; s < nb; ++s) {
      const TeamBlk d = team_desc[(size_t)t * seq_stride + s];
      if ((d.sq & 1) == 0) continue;  // each row once: its pass-2 block
      float2* out = (outs ? outs[d.sig] : out1) + d.out_off;
      for (int e = threadIdx.x; e < d.n; e += blockDim.x) out[e] = make_float2(qnan, qnan);
    }
  }
}

// Index (int)fminf(N-1, step*q) into the mother table for the segment holding
// q, or -1 when no segment holds it (fcwt.cpp:224-248); the row's segments
// are staged in LDS.
__device__ __forceinline__ int window_index(int q, int nseg, const SegD* sseg, float maxidx) {
#pragma clang fp contract(off)
  int hit = -1;
  for (int s = 0; s < nseg; ++s)
    if (q >= sseg[s].lo && q < sseg[s].hi) hit = s;
  if (hit < 0) return -1;
  float v = sseg[hit].step * (float)q;
  return (int)(v < maxidx ? v : maxidx);
}

// The same with the row's first two segments in registers (from TeamBlk)
// and any further ones read by scalar loads.
__device__ __forceinline__ int window_index_d(int q, int nseg, const SegD& s0, const SegD& s1,
                                              const SegD* __restrict__ gseg, float maxidx) {
#pragma clang fp contract(off)
  bool hit = false;
  float step = 0.f;
  if (nseg > 0 && q >= s0.lo && q < s0.hi) {
    hit = true;
    step = s0.step;
  }
  if (nseg > 1 && q >= s1.lo && q < s1.hi) {
    hit = true;
    step = s1.step;
  }
  for (int s = 2; s < nseg; ++s) {  // a later segment overrides an earlier one
    const int* w = (const int*)(gseg + s);
    if (q >= ld_const(w) && q < ld_const(w + 1)) {
      hit = true;
      step = __int_as_float(ld_const(w + 2));
    }
  }
  if (!hit) return -1;
  float v = step * (float)q;
  return (int)(v < maxidx ? v : maxidx);
}

// Pass-1 register split of the A-point transform: R1 = min(A, 16) points per
// thread, P2 = A / R1 threads per transform.
__host__ __device__ constexpr int team_log2r1(int log2A) { return log2A < 4 ? log2A : 4; }
// Row stride of the staged band in LDS (and, with P2 > 1, of the pass-1
// exchange rows, which reuse it): at least A + P2, rounded so that the
// staging writes of one 16-lane LDS cycle fall on distinct banks -- odd when
// those lanes hold 16 rows ql (Qb >= 16), 2 mod 4 when they hold 8 rows and
// 2 columns (Qb = 8).  The pass-1 reads of a band row are consecutive or
// broadcast whatever the stride.
__host__ __device__ constexpr int team_zs_stride(int log2A, int log2F) {
  return log2A <= 4 ? (1 << log2A) + 1
         : 12 - log2F == 3 ? (((1 << log2A) + (1 << (log2A - 4))) | 2) & ~1
                           : ((1 << log2A) + (1 << (log2A - 4))) | 1;
}
// The staged band shares the work-group's LDS buffer with the exchanges and
// the tile.  With P2 > 1 it sits at the start (pass 1 reads it into registers
// and waits at a barrier before the exchange overwrites it); with P2 = 1 the
// transforms run in several rounds that read it and write the tile, so it
// sits after the tile of F (Qb + 1) values.
__host__ __device__ constexpr int team_zs_offset(int log2A, int log2F) {
  return log2A > 4 ? 0 : (1 << log2F) * ((1 << (12 - log2F)) + 1);
}
// LDS buffer the pass functions need (complex64 values) for F = 2^log2F:
// tile F (Qb + 1) <= 4608, exchanges 4352, band after the tile when A <= 16
__host__ __device__ constexpr int team_buf_need(int log2F) {
  return team_zs_offset(4, log2F) + (1 << (12 - log2F)) * 17;
}
constexpr int kTeamBuf = 4768;
static_assert(kTeamThreads == kTeamB, "one thread per pass-2 table entry");
static_assert(team_buf_need(7) <= kTeamBuf && team_buf_need(8) <= kTeamBuf && team_buf_need(9) <= kTeamBuf &&
                  (1 << 9) * ((1 << 3) + 1) <= kTeamBuf && 4352 <= kTeamBuf,
              "team LDS buffer too small");

// The team kernel's LDS, at namespace scope so that the pass functions (not
// inlined) address it as LDS rather than through generic pointers.
//   tm_buf: the staged band, pass-1 exchange and tile, pass-2 exchange
//           (team_zs_offset, team_buf_need)
__shared__ float2 tm_buf[kTeamBuf];
__shared__ float2 tm_pre[1 << kTeamMaxLog2F];   // pass-1 [k][l']
__shared__ float2 tm_post[1 << kTeamMaxLog2F];  // pass-1 [k1][l'][j]
__shared__ float2 tm_t2[kTeamB];                // pass-2 [g][k]
__shared__ float2 tm_wb[512];                   // pass-2 [k1][j] (B = 256) or [m1][j] (B = 512)

// Pass 1 for one row and the member's Qb = 4096/F qb values, A = 2^LOG2A,
// B = 2^LOG2B.
//   band (tm_buf + zs_off)  the row's Z[qb0 + ql + B qf] at [ql * (A + P2) + qf]
//   tm_pre[k][l']           = W_F^{P2 k l'}             (input factor)
//   tm_post[k1][l'][j]      = W_F^{j (l' + (F/A) k1)}  (between the register passes, P2 > 1)
// Result U[f][qb0 .. qb0+Qb), stored through an LDS tile in Qb-element runs.
// The passes are separate (not inlined) functions: inlined into the row loop
// together they exceed the register file and spill.
template <int LOG2A, int LOG2F, int LOG2B>
__device__ __noinline__ void team_pass1(int qb0, gfloat2* __restrict__ U) {
  constexpr int log2F = LOG2F;
  constexpr int zs_off = team_zs_offset(LOG2A, LOG2F);
  float2* const buf = tm_buf;
  const float2* const zs = tm_buf + zs_off;
  const float2* const pre = tm_pre;
  const float2* const post = tm_post;
  constexpr int A = 1 << LOG2A;
  constexpr int LOG2R1 = team_log2r1(LOG2A);
  constexpr int R1 = 1 << LOG2R1;
  constexpr int LOG2P2 = LOG2A - LOG2R1;
  constexpr int P2 = 1 << LOG2P2;
  constexpr int ZS = team_zs_stride(LOG2A, LOG2F);
  constexpr int log2Ph = log2F - LOG2A;  // phases l' per qb
  constexpr int log2Qb = 12 - log2F;
  constexpr int TS = (1 << log2Qb) + 1;  // padded tile row
  if constexpr (P2 == 1) {
    // 4096/A transforms of A <= 16 points: 16/A rounds of 256
#pragma unroll 1
    for (int rd = 0; rd < 16 / A; ++rd) {
      const int tr = rd * kTeamThreads + threadIdx.x;
      const int ql = tr >> log2Ph, lp = tr & ((1 << log2Ph) - 1);
      float2 v[R1];
#pragma unroll
      for (int k = 0; k < R1; ++k) v[k] = cmul(zs[ql * ZS + k], pre[(k << log2Ph) + lp]);
      dft_reg<R1, 1>(v);
#pragma unroll
      for (int m = 0; m < R1; ++m) buf[(lp + (m << log2Ph)) * TS + ql] = v[m];
    }
  } else {
    // 4096/A transforms of A = 16 P2 points, P2 threads each: one round
    static_assert(R1 == 16 && P2 <= 16, "A out of range");
    constexpr int XS = ZS;  // exchange [tr][j][k1], j stride R1 + 1 (band row ql is transform (ql, 0)'s row)
    const int tr = threadIdx.x >> LOG2P2, j = threadIdx.x & (P2 - 1);
    const int ql = tr >> log2Ph, lp = tr & ((1 << log2Ph) - 1);
    float2 v[R1];
#pragma unroll
    for (int k = 0; k < R1; ++k) v[k] = cmul(zs[ql * ZS + j + P2 * k], pre[(k << log2Ph) + lp]);
    dft_reg<R1, 1>(v);
    // The table reads below must not be hoisted above the DFT (the compiler
    // would hold 16 more values in registers and lose an occupancy step):
    // their index depends on an opaque zero produced after the DFT results.
    int after = 0;
#pragma unroll
    for (int k = 0; k < R1; ++k) asm volatile("" : "+v"(v[k].x), "+v"(v[k].y));
    asm volatile("" : "+v"(after));
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[k1] = cmul(v[k1], post[(((k1 << log2Ph) + lp) << LOG2P2) + j + after]);
    // Exchange rows: transform (ql, 0) reuses band row ql, the others sit
    // after the Qb band rows (16 values further on, against bank conflicts).
    // Every reader of band row ql is in this wavefront (Ph P2 = F/16 <= 64
    // consecutive threads) and has read it above, so only the wavefront's own
    // LDS order is needed between those reads, these writes and the reads below.
    const int xr = lp == 0 ? ql * XS : ((1 << log2Qb) + ql * ((1 << log2Ph) - 1) + lp - 1) * XS + 16;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) buf[xr + j * (R1 + 1) + k1] = v[k1];
    wave_lds_order();
    constexpr int Uu = R1 / P2;
#pragma unroll
    for (int u = 0; u < Uu; ++u) {
      const int k1 = u * P2 + j;
      float2 w2[P2];
#pragma unroll
      for (int jj = 0; jj < P2; ++jj) w2[jj] = buf[xr + jj * (R1 + 1) + k1];
      dft_reg<P2, 1>(w2);
#pragma unroll
      for (int k2 = 0; k2 < P2; ++k2) v[u * P2 + k2] = w2[k2];
    }
    __syncthreads();
    // output m' = k1 + R1 k2 -> f = l' + (F/A) m', into the tile [f][ql]
#pragma unroll
    for (int u = 0; u < Uu; ++u)
#pragma unroll
      for (int k2 = 0; k2 < P2; ++k2) {
        const int f = lp + ((u * P2 + j + R1 * k2) << log2Ph);
        buf[f * TS + ql] = v[u * P2 + k2];
      }
  }
  __syncthreads();
  // every LDS read first, then the stores: pairs of consecutive qb (Qb >= 8
  // is even, so a pair never straddles two f), 16 bytes per store
  constexpr int S = kTeamPoints / kTeamThreads;
  float2 o[S];
#pragma unroll
  for (int s = 0; s < S / 2; ++s) {
    const int e = 2 * (s * kTeamThreads + threadIdx.x);
    o[2 * s] = buf[(e >> log2Qb) * TS + (e & ((1 << log2Qb) - 1))];
    o[2 * s + 1] = buf[(e >> log2Qb) * TS + (e & ((1 << log2Qb) - 1)) + 1];
  }
#pragma unroll
  for (int s = 0; s < S / 2; ++s) {
    const int e = 2 * (s * kTeamThreads + threadIdx.x);
    gstore2(U + ((e >> log2Qb) << LOG2B) + qb0 + (e & ((1 << log2Qb) - 1)), o[2 * s], o[2 * s + 1]);
  }
}

// Pass 2 for one row: 16 consecutive f starting at f0, the factor W_N^{qb f}
// and 256-point inverse FFTs over qb = j + 16 k, written to out (the row)
// with non-temporal stores (written once, never re-read here).  Inlined into
// the kernel: a function's return waits for its stores, and these need not
// complete before the member starts its next item.
//   tm_t2[g][k]  = W_N^{16 k (f0 + g)}   tm_wb[k1][j] = W_B^{j k1}
//   base      = c W_N^{j (f0 + g)} for this thread, c the row normalisation
template <int LOG2F>
__device__ __forceinline__ void team_pass2(gfloat2* __restrict__ out, int n, int f0,
                                        const gfloat2* U, float2 base, unsigned* done) {
  float2* const buf = tm_buf;
  const float2* const t2 = tm_t2;
  const float2* const wb = tm_wb;
  constexpr int R1 = 16, P2 = 16, XS = 16 * (P2 + 1);
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  float2 v[R1];
#pragma unroll
  for (int k = 0; k < R1; ++k) v[k] = ld_agent2(U + (size_t)(f0 + g) * kTeamB + j + P2 * k);
#pragma unroll
  for (int k = 1; k < R1; ++k) v[k] = cmul(v[k], t2[g * R1 + k]);
  dft_reg<R1, 1>(v);
  v[0] = cmul(v[0], base);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], cmul(base, wb[k1 * P2 + j]));
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) buf[k1 * XS + g * (P2 + 1) + j] = v[k1];  // [k1][g][j]
  __syncthreads();
  // every thread's reads of U have landed (they fed the exchange above), so
  // the buffer may be reused; the output stores need not complete
  if (threadIdx.x == 0) add_agent(done, 1u);
  // The second stage by thread (g2, k1) = (tid & 15, tid >> 4): the DFT over
  // j gives m = k1 + 16 k2 at f = f0 + g2, so the 16 lanes of one k1 store a
  // 128-byte run of the row straight from the registers, with no output
  // transpose through LDS (the exchange above is the only one; against a
  // wave-local exchange, an LDS transpose and 16-byte pair stores this is
  // 1.7% of the headline step, measured).  Stores at t >= n are clipped by
  // the row's buffer range; the rows m >= mlim of every f of the item are
  // not stored.
  const int g2 = threadIdx.x & 15, k1s = threadIdx.x >> 4;
  float2 w[P2];
#pragma unroll
  for (int jj = 0; jj < P2; ++jj) w[jj] = buf[k1s * XS + g2 * (P2 + 1) + jj];
  dft_reg<P2, 1>(w);
  const __amdgpu_buffer_rsrc_t rows = row_rsrc((float2*)out, n);
  const int mlim = __builtin_amdgcn_readfirstlane((n - f0 + (1 << LOG2F) - 1) >> LOG2F);
  const int slim = (mlim + 15) >> 4;
#pragma unroll
  for (int k2 = 0; k2 < P2; ++k2)
    if (k2 < slim) bstore_nt(rows, (unsigned)(f0 + g2 + ((k1s + 16 * k2) << LOG2F)), w[k2]);
}

// Pass 2 at B = 512 for one row: 8 consecutive f starting at f0.  With
// qb = j + 32 k and m = m1 + 16 m2:
//   y[f + F m] = c sum_j W_32^{j m2} W_512^{j m1} W_N^{j f}
//                      sum_k W_16^{k m1} W_N^{32 k f} U[f][qb]
// Stage 1, thread (g, j) = (tid >> 5, tid & 31): the 16-point DFT over k
// (lanes of one g read 32 consecutive qb of its scratch row).  Stage 2,
// thread (g, m1, e) with g = tid & 7, e = (tid >> 4) & 1 and
// m1 = ((tid >> 3) & 1) + 2 (tid >> 5): the 16-point DFT over j' of the
// values j = 2 j' + e, E_e[m2'], then the radix-2 step between the lane pair
// e = 0, 1 (lanes 16 apart):
//   y[f + F (m1 + 16 m2' + 256 h)] = E_0[m2'] + (-1)^h W_32^{m2'} E_1[m2'],
// lane e writing h = e.  Lanes g = 0..7 of one (m1, e) write 64-byte runs of
// the output row straight from the registers.
//   tm_t2[g][k] = W_N^{32 k (f0 + g)}   tm_wb[m1][j] = W_512^{j m1}
//   base = c W_N^{j (f0 + g)} for this thread's stage-1 (g, j)
template <int LOG2F>
__device__ __forceinline__ void team_pass2_b512(gfloat2* __restrict__ out, int n, int f0, const gfloat2* U,
                                                float2 base, unsigned* done) {
  // exchange [g][m1][j]: strides 34 per m1 and 548 per g, so that the 32
  // lanes of one stage-2 read (8 g, m1 & 1, e) fall on distinct banks
  constexpr int JS = 34, XS = 16 * JS + 4;
  float2* const xb = tm_buf;
  const float2* const t2 = tm_t2;
  const float2* const wb = tm_wb;
  const int tid = threadIdx.x;
  {
    const int g = tid >> 5, j = tid & 31;
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = ld_agent2(U + (size_t)(f0 + g) * 512 + j + 32 * k);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], t2[g * 16 + k]);
    dft_reg<16, 1>(v);
    v[0] = cmul(v[0], base);
#pragma unroll
    for (int m1 = 1; m1 < 16; ++m1) v[m1] = cmul(v[m1], cmul(base, wb[m1 * 32 + j]));
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) xb[g * XS + m1 * JS + j] = v[m1];
  }
  __syncthreads();
  // every thread's reads of U have landed (they fed the exchange above), so
  // the buffer may be reused; the output stores need not complete
  if (tid == 0) add_agent(done, 1u);
  const int g = tid & 7, e = (tid >> 4) & 1, m1 = ((tid >> 3) & 1) + 2 * (tid >> 5);
  float2 v[16];
#pragma unroll
  for (int jp = 0; jp < 16; ++jp) v[jp] = xb[g * XS + m1 * JS + 2 * jp + e];
  dft_reg<16, 1>(v);  // E_e[m2']
  const bool odd = e != 0;
  float2 o[16];
#pragma unroll
  for (int m2 = 0; m2 < 16; ++m2) {
    float2 x = v[m2];
    if (m2 > 0) {
      const float2 wx = cmul(x, wb[m2 * 32 + 16]);  // W_32^{m2'} = W_512^{16 m2'}
      x = odd ? wx : x;
    }
    const float2 pr = make_float2(lane_xor16(x.x, odd), lane_xor16(x.y, odd));
    o[m2] = odd ? csub(pr, x) : cadd(x, pr);
  }
  // stores clipped at t >= n by the row's buffer range; 64-byte runs, whose
  // other halves the neighbouring item writes, so without the non-temporal
  // hint (the L2 joins the halves: 6.5% of the 384-channel batch's N = 2^18
  // channels, measured)
  const __amdgpu_buffer_rsrc_t rows = row_rsrc((float2*)out, n);
#pragma unroll
  for (int m2 = 0; m2 < 16; ++m2)
    bstore(rows, (unsigned)(f0 + g) + ((unsigned)(m1 + 16 * m2 + 256 * e) << LOG2F), o[m2]);
}

// Calls pass 1 for the row's A = 2^log2A (A <= F/2).
template <int LOG2F, int LOG2B>
__device__ __forceinline__ void team_pass1_for(int log2A, int qb0, gfloat2* U) {
  switch (log2A) {
    case 1: team_pass1<1, LOG2F, LOG2B>(qb0, U); break;
    case 2: team_pass1<2, LOG2F, LOG2B>(qb0, U); break;
    case 3: team_pass1<3, LOG2F, LOG2B>(qb0, U); break;
    case 4: team_pass1<4, LOG2F, LOG2B>(qb0, U); break;
    case 5: team_pass1<5, LOG2F, LOG2B>(qb0, U); break;
    case 6: team_pass1<6, LOG2F, LOG2B>(qb0, U); break;
    case 7:
      if constexpr (LOG2F > 7) team_pass1<7, LOG2F, LOG2B>(qb0, U);
      break;
    default:
      if constexpr (LOG2F > 8) team_pass1<8, LOG2F, LOG2B>(qb0, U);
      break;
  }
}

// One block of a direct row (rows_direct.h) in the team kernel's LDS buffer.
template <int LOG2M>
__device__ __noinline__ void team_direct(const float2* spec, const float* mother, const SegD* segs, const RowD* rows,
                                         float2* out, float2* const* outs, const float2* t1, const float2* t2,
                                         int s1bits, int log2N, float norm, int row, int block) {
  static_assert(DirectCfg<LOG2M, true>::NT == kTeamThreads && DirectCfg<LOG2M, true>::LDS <= kTeamBuf,
                "a direct block must fit the team kernel's work-group");
  KArgs a;
  a.spec = spec;
  a.mother = mother;
  a.segs = segs;
  a.rows = rows;
  a.items = nullptr;
  a.scratch = nullptr;
  a.out = out;
  a.outs = outs;
  a.N = 1 << log2N;
  a.log2N = log2N;
  a.norm = norm;
  a.tw.t1 = t1;
  a.tw.t2 = t2;
  a.tw.s1bits = s1bits;
  a.tw.log2N = log2N;
  direct_block<LOG2M, true>(a, Item{row, block}, tm_buf);
}

// Direct-row blocks claimed from the launch-wide counter until none is left.
// The team's last rows leave members idle (pass 2 of the last rows waits for
// their pass 1, and the teams finish at different times); these blocks,
// which depend on nothing, fill that time.
__device__ __forceinline__ void team_fill(const TeamArgs& ta, unsigned* next, int* s_next) {
  for (;;) {
    __syncthreads();  // every thread has read s_next
    if (threadIdx.x == 0) *s_next = (int)atomicAdd(next, 1u);
    __syncthreads();
    const int d = __builtin_amdgcn_readfirstlane(*s_next);
    if (d >= ta.nfill) break;
    const int* const w = (const int*)(ta.fill_items + d);
    const Item it = {ld_const(w), ld_const(w + 1)};
    int lm = kFillMinLog2M;
#pragma unroll
    for (int g = 0; g < kFillGroups - 1; ++g) lm += d >= ta.fill_end[g];
    const KArgs& k = ta.k;
#define FCWT_TEAM_DIRECT(L) \
  team_direct<L>(k.spec, k.mother, k.segs, k.rows, k.out, k.outs, k.tw.t1, k.tw.t2, k.tw.s1bits, k.log2N, k.norm, it.row, it.block)
    switch (lm) {
      case 0: FCWT_TEAM_DIRECT(0); break;
      case 1: FCWT_TEAM_DIRECT(1); break;
      case 2: FCWT_TEAM_DIRECT(2); break;
      case 3: FCWT_TEAM_DIRECT(3); break;
      case 4: FCWT_TEAM_DIRECT(4); break;
      case 5: FCWT_TEAM_DIRECT(5); break;
      case 6: FCWT_TEAM_DIRECT(6); break;
      case 7: FCWT_TEAM_DIRECT(7); break;
      case 8: FCWT_TEAM_DIRECT(8); break;
      case 9: FCWT_TEAM_DIRECT(9); break;
      default: FCWT_TEAM_DIRECT(10); break;
    }
#undef FCWT_TEAM_DIRECT
  }
}

// Items per pass of one row of N = 2^log2N points: N / 4096.
__host__ __device__ constexpr int team_log2i(int log2N) { return log2N - 12; }

// This work-group's walk of its team's sequence in one plan's launch, until
// the sequence is claimed to its end or the launch aborts (a failed wait).
// F = 2^LOG2F = N / B and B = 2^LOG2B are template parameters, so every
// stride and table index in the pass functions is a compile-time constant.
// s_flag and s_next are the work-group's LDS words for broadcasts.
template <int LOG2F, int LOG2B>
__device__ __forceinline__ void team_walk(const TeamArgs& ta, unsigned* const abort, const int team, int& s_flag,
                                          int& s_next) {
  static_assert(LOG2B == kTeamLog2B || (LOG2B == 9 && LOG2F == 9), "team split out of range");
  float2* const buf = tm_buf;
  float2* const pre = tm_pre;
  float2* const post = tm_post;
  float2* const t2s = tm_t2;
  // plain copies of the arguments
  const float2* __restrict__ spec = ta.k.spec;
  const float* __restrict__ mother = ta.k.mother;
  const SegD* __restrict__ segs = ta.k.segs;
  const Twiddles tw = ta.k.tw;
  const int N = ta.k.N;
  const float norm = ta.k.norm;
  constexpr int log2F = LOG2F;
  unsigned* const sync = ta.sync;
  unsigned* const error = ta.error;
  const long spin = ta.spin;
  constexpr int F = 1 << log2F;
  constexpr int log2Qb = 12 - log2F;
  constexpr int log2I = team_log2i(log2F + LOG2B);  // items per pass per row
  constexpr int log2Fi = 12 - LOG2B;         // f per pass-2 item: 4096/B
  if constexpr (LOG2B == 9) {
    for (int e = threadIdx.x; e < 512; e += kTeamThreads) {  // [m1][j] = W_512^{j m1}
      const int m1 = e >> 5, j = e & 31;
      tm_wb[e] = tw((unsigned)(m1 * j) << log2F);
    }
  } else {
    for (int e = threadIdx.x; e < kTeamB; e += kTeamThreads) {  // [k1][j] = W_256^{j k1}
      const int k1 = e >> 4, j = e & 15;
      tm_wb[e] = tw((unsigned)(k1 * j) << log2F);
    }
  }
  const int nbuf = ta.nbuf;
  float2* const scratch = ta.scratch + (size_t)team * nbuf * N;
  unsigned* const arrive = sync + 64 * (1 + team * kTeamBuffers);
  const TeamBlk* const desc = ta.team_desc + (size_t)team * ta.seq_stride;
  const int items = ta.team_blocks[team] << log2I;
  int tab_log2A = -1;  // pass-1 tables currently built for this A
  // items are claimed in sequence order from the team's counter; an item
  // claims its successor when it starts, so the claim's latency is hidden
  unsigned* const next_item = sync + 64 * (1 + ta.nteams * kTeamBuffers + team);
  // every wave has read the broadcast words (s_next, s_flag) of what ran
  // before this walk (another plan's walk in k_rows_team_multi) before
  // thread 0 writes them again
  __syncthreads();
  if (threadIdx.x == 0) s_next = (int)atomicAdd(next_item, 1u);
  __syncthreads();
  int it = __builtin_amdgcn_readfirstlane(s_next);
  while (it < items) {
    unsigned claim = 0;
    if (threadIdx.x == 0) claim = atomicAdd(next_item, 1u);
    const int* const dw = (const int*)(desc + (it >> log2I));
    const int sq = ld_const(dw), blk = it & ((1 << log2I) - 1);
    const int i = sq >> 1;
    // every field is read here, in one burst, before any wait
    const int seg_begin = ld_const(dw + 1);
    const int nseg = ld_const(dw + 2);  // <= kTeamSegLds: the schedule routes rows with more elsewhere
    const int log2A = ld_const(dw + 3);
    const int sig = ld_const(dw + 4), n = ld_const(dw + 5);
    const long long out_off = (long long)(unsigned)ld_const(dw + 6) | ((long long)ld_const(dw + 7) << 32);
    SegD s0, s1;
    s0.lo = ld_const(dw + 8);
    s0.hi = ld_const(dw + 9);
    s0.step = __int_as_float(ld_const(dw + 10));
    s1.lo = ld_const(dw + 12);
    s1.hi = ld_const(dw + 13);
    s1.step = __int_as_float(ld_const(dw + 14));
    float2* out = ta.k.out;  // the row's output (pass-2 items)
    if (ta.k.outs) {
      const int* const pw = (const int*)(ta.k.outs + sig);
      out = (float2*)((unsigned long long)(unsigned)ld_const(pw) | ((unsigned long long)(unsigned)ld_const(pw + 1) << 32));
    }
    out += out_off;
    const int b = i % nbuf, gen = i / nbuf;
    gfloat2* const U = (gfloat2*)(scratch + (size_t)b * N);
    // the counter this item waits on (pass 1: the buffer's pass-2 readers of
    // the row nbuf earlier; pass 2: the row's pass-1 arrivals), read now so
    // that the read overlaps the item's own loads
    const unsigned* const dep = arrive + 64 * b + ((sq & 1) ? 0 : 32);
    unsigned first = 0;
    if (threadIdx.x == 0 && ((sq & 1) || gen > 0)) first = ld_agent(dep);
    if ((sq & 1) == 0) {
      // ---- pass 1 of row i, qb block blk
      const int qb0 = blk << log2Qb;
      const int A = 1 << log2A;
      if (log2A != tab_log2A) {
        tab_log2A = log2A;
        const int log2R1 = team_log2r1(log2A), R1 = 1 << log2R1;
        const int log2P2 = log2A - log2R1, P2 = 1 << log2P2;
        const int log2Ph = log2F - log2A;
        for (int e = threadIdx.x; e < (R1 << log2Ph); e += kTeamThreads) {
          // [k][l'] = W_F^{P2 k l'} = W_N^{B P2 k l'}; P2 k l' < F
          const int lp = e & ((1 << log2Ph) - 1), k = e >> log2Ph;
          pre[e] = tw((unsigned)(k * lp) << (log2P2 + LOG2B));
        }
        if (P2 > 1)
          for (int e = threadIdx.x; e < F; e += kTeamThreads) {  // [k1][l'][j] = W_F^{j (l' + (F/A) k1)}
            const int j = e & (P2 - 1), lp = (e >> log2P2) & ((1 << log2Ph) - 1), k1 = e >> (log2P2 + log2Ph);
            post[e] = tw((unsigned)((j * (lp + (k1 << log2Ph))) & (F - 1)) << LOG2B);
          }
      }
      // the row's windowed band on this item's bins, into LDS: at most
      // A * Qb = 2048 values, 8 per thread, all loads issued before the
      // first is used
      const float maxidx = (float)N - 1.0f;
      const float2* __restrict__ rspec = spec + (size_t)sig * (size_t)(N >> 1);
      const int zstride = team_zs_stride(log2A, log2F);
      float2* const zs = buf + team_zs_offset(log2A, log2F);
      const int total = A << log2Qb;
      constexpr int kPer = kTeamPoints / 2 / kTeamThreads;
      int qs[kPer], idx[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = u * kTeamThreads + threadIdx.x;
        const int ql = e & ((1 << log2Qb) - 1), qf = e >> log2Qb;  // consecutive lanes, consecutive q
        qs[u] = qb0 + ql + (qf << LOG2B);  // qf < A <= F/2: q < N/2
        idx[u] = e < total ? window_index_d(qs[u], nseg, s0, s1, segs + seg_begin, maxidx) : -1;
      }
      float2 xs[kPer];
      float ws[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        xs[u] = idx[u] >= 0 ? rspec[qs[u]] : make_float2(0.f, 0.f);
        ws[u] = idx[u] >= 0 ? mother[idx[u]] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = u * kTeamThreads + threadIdx.x;
        const int ql = e & ((1 << log2Qb) - 1), qf = e >> log2Qb;
        if (e < total) zs[ql * zstride + qf] = make_float2(xs[u].x * ws[u], xs[u].y * ws[u]);
      }
      // the buffer's previous row has been read by every pass-2 item (the
      // wait ends in a barrier, which also publishes the staged band)
      if (gen > 0) {
        if (!team_wait(dep, (unsigned)(gen << log2I), first, &s_flag, abort, error, spin)) break;
      } else {
        __syncthreads();
      }
      team_pass1_for<LOG2F, LOG2B>(log2A, qb0, U);
      // every thread's pass-1 stores are complete before the arrival is counted
      wait_stores();
      __syncthreads();
      if (threadIdx.x == 0) add_agent(arrive + 64 * b, 1u);
    } else {
      // ---- pass 2 of row i, f block blk
      // issued ahead of the other work-groups' pass-1 work on the SIMD: a
      // pass-2 item's end releases its buffer and finishes rows (0.5-1% of
      // the step, measured)
      __builtin_amdgcn_s_setprio(2);
      const int f0 = blk << log2Fi;
      float2 base;  // W_N^{j f} times the row normalisation, for this thread's (g, j)
      if constexpr (LOG2B == 9) {
        if (threadIdx.x < 128) {
          const int g = threadIdx.x >> 4, k = threadIdx.x & 15;  // t2s[g][k] = W_N^{32 k (f0 + g)}
          t2s[threadIdx.x] = tw((unsigned)((32 * k * (f0 + g)) & (N - 1)));
        }
        base = cscale(tw((unsigned)((threadIdx.x & 31) * (f0 + (threadIdx.x >> 5)))), norm);
      } else {
        const int g = threadIdx.x >> 4, k = threadIdx.x & 15;  // t2s[g][k] = W_N^{16 k (f0 + g)}
        t2s[threadIdx.x] = tw((unsigned)((16 * k * (f0 + g)) & (N - 1)));
        base = cscale(tw((unsigned)((threadIdx.x & 15) * (f0 + (threadIdx.x >> 4)))), norm);
      }
      // every pass-1 item of the row has stored its part (the wait ends in a barrier)
      if (!team_wait(dep, (unsigned)((gen + 1) << log2I), first, &s_flag, abort, error, spin)) break;
      if constexpr (LOG2B == 9)
        team_pass2_b512<LOG2F>((gfloat2*)out, n, f0, U, base, arrive + 64 * b + 32);
      else
        team_pass2<LOG2F>((gfloat2*)out, n, f0, U, base, arrive + 64 * b + 32);
      __builtin_amdgcn_s_setprio(0);
    }
    // publishes the next item, and every thread is done with this item's LDS
    if (threadIdx.x == 0) s_next = (int)claim;
    __syncthreads();
    it = __builtin_amdgcn_readfirstlane(s_next);
  }
}

// The team: the XCD this work-group runs on (its XCC id).
__device__ __forceinline__ int team_of_workgroup() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7; }

// One plan's wide rows (and, once a member's sequence is claimed to its
// end, its direct-row blocks).  The launch's abort and exit words are the
// plan's own sync words.
template <int LOG2F, int LOG2B>
__global__ __launch_bounds__(kTeamThreads, kTeamWgPerCu) void k_rows_team(TeamArgs ta) {
  __shared__ int s_flag, s_next;
  unsigned* const abort = ta.sync + kSyncAbort;
  team_walk<LOG2F, LOG2B>(ta, abort, team_of_workgroup(), s_flag, s_next);
  if (ta.nfill > 0) team_fill(ta, ta.sync + kSyncFill, &s_next);
  // (s_flag holds the exit's broadcast until every wave has read it; the
  // abort broadcast goes through s_next, which every wave read before the
  // exit's barrier)
  if (team_exit_last(ta.sync, &s_flag)) {
    if (threadIdx.x == 0) {
      team_check_claims(ta.sync, abort, ta.error, ta.team_blocks, ta.nteams, team_log2i(LOG2F + LOG2B));
      s_next = ld_agent(abort) != 0;
    }
    __syncthreads();
    if (s_next) team_nan_rows(ta.team_desc, ta.team_blocks, ta.seq_stride, ta.nteams, ta.k.out, ta.k.outs);
  }
}

// One launch for the wide rows of several plans of a batch, one plan per
// padded size: seg[0..3] hold the plans of N = 2^18, 2^17, 2^16, 2^15 (bit s
// of `used` set when seg[s] holds one).  Every work-group walks its team's
// sequence of each plan in turn, largest N first: once a plan's sequence is
// claimed to its end the member goes on to the next plan's, so the plans'
// sequences overlap at their ends instead of each launch draining alone.
// Each plan keeps its own scratch, item counters and buffer counters, so a
// plan's items depend only on earlier items of the same plan and every
// plan's walk progresses as in its own launch.  After the last plan, the
// members take the direct-row blocks of every plan.  The abort and exit words
// are the launch's own (ctl): a failed wait in any plan ends the whole
// launch, every plan reports it, and every plan's team rows are filled with
// NaN.
constexpr int kTeamMaxSegs = 4;
struct TeamMulti {
  TeamArgs seg[kTeamMaxSegs];
  unsigned* ctl;  // [kSyncAbort] abort code, [kSyncExited] finished work-groups; zeroed before launch
  int used;
};
static_assert(sizeof(TeamMulti) <= 4096, "kernel arguments are limited to 4 KiB");

__global__ __launch_bounds__(kTeamThreads, kTeamWgPerCu) void k_rows_team_multi(TeamMulti m) {
  __shared__ int s_flag, s_next;
  unsigned* const abort = m.ctl + kSyncAbort;
  const int team = team_of_workgroup();
  const int used = m.used;
  if (used & 1) team_walk<9, 9>(m.seg[0], abort, team, s_flag, s_next);
  if (used & 2) team_walk<9, 8>(m.seg[1], abort, team, s_flag, s_next);
  if (used & 4) team_walk<8, 8>(m.seg[2], abort, team, s_flag, s_next);
  if (used & 8) team_walk<7, 8>(m.seg[3], abort, team, s_flag, s_next);
#pragma unroll
  for (int s = 0; s < kTeamMaxSegs; ++s)
    if ((used >> s & 1) && m.seg[s].nfill > 0) team_fill(m.seg[s], m.seg[s].sync + kSyncFill, &s_next);
  if (team_exit_last(m.ctl, &s_flag)) {
    if (threadIdx.x == 0) {
#pragma unroll
      for (int s = 0; s < kTeamMaxSegs; ++s)
        if (used >> s & 1)
          team_check_claims(m.seg[s].sync, abort, m.seg[s].error, m.seg[s].team_blocks, m.seg[s].nteams,
                            team_log2i(18 - s));
      const unsigned code = ld_agent(abort);
      if (code != 0)
#pragma unroll
        for (int s = 0; s < kTeamMaxSegs; ++s)
          if (used >> s & 1) atomicOr(m.seg[s].error, code);
      s_next = code != 0;  // as in k_rows_team: not s_flag, which a wave may still be reading
    }
    __syncthreads();
    if (s_next)
#pragma unroll
      for (int s = 0; s < kTeamMaxSegs; ++s)
        if (used >> s & 1)
          team_nan_rows(m.seg[s].team_desc, m.seg[s].team_blocks, m.seg[s].seq_stride, m.seg[s].nteams,
                        m.seg[s].k.out, m.seg[s].k.outs);
  }
}

}  // namespace fcwt_hip
