// dev_solver.hip -- device-resident LP solver for one large LP: the ops of
// solver_ops.hpp run as a wide "step" kernel (one op per launch, each
// workgroup leaving its partial reductions in memory) followed by a
// one-workgroup "control" kernel that combines the partials in block
// order and advances the state machine.  The kernel boundary is the
// global barrier.  Pairs of (step, control) launches are captured once per
// problem in a hipGraph and replayed in chunks; between chunks the host
// reads the small state record, replays log events and stops when the
// solve is done.  Only the control kernel takes decisions, so the host
// never participates in an iteration.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <algorithm>
#include "acd_dev.h"
#include "solver_ops.hpp"

using namespace acd;

namespace {

inline hipStream_t S() { return (hipStream_t)acd_dev_stream(); }

constexpr int kWPB = 4;  // wavefronts per workgroup of the step kernel
constexpr uint32_t kEventCap = 4096;
constexpr int kPairs = 32;  // ops per graph (fused launches or step/control pairs)
constexpr uint32_t kMaxBlocks = 2048;  // step-kernel grid cap (control loads)
// Precision keys: iterate vectors / constant vectors / matrix values.
enum {
  K_F64 = 0,    // double / double / double
  K_MIX_F32,    // double / float / float
  K_MIX_BF16,   // double / float / bf16
  K_F32,        // float / float / float
  K_F32_BF16,   // float / float / bf16
  K_MIX_SGN,    // double / float / value-free (0/+-1 operators)
  K_MIX_C16,    // double / float / float, 16-bit columns (batched kernel)
  kKeys
};
uint64_t g_prof_count[32];
double g_prof_us[32];
double g_prof_step_us[32];
double g_prof_bytes[32];
double g_ctl_us[4];

struct GridG {
  // Row loop with the next chunk's entry loads waited for before the
  // gathers (measured faster on the large LPs' HBM-streamed operators:
  // set cover 51.7 s against 53.5 s per solve).
  static constexpr bool kPipelinedRows = false;
  uint32_t lane, gwave, nwaves;
};

// Slice s goes to wave s (grid-stride beyond the grid).  Workgroups are
// dispatched round-robin over the 8 XCDs, so consecutive slices land on
// different XCDs and an operator with fewer slices than the grid has waves
// (A of a set-cover LP: 3,125 slices for 8,192 waves) still spreads over
// every XCD.  Renumbering blocks into contiguous per-XCD ranges put all of
// such an operator on the first three XCDs and halved its throughput.
template <class VI, class VC, class MT>
__global__ void __launch_bounds__(64 * kWPB) k_step(Args a) {
  const State& s = *a.st;
  if (s.op == OP_NONE) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->t_step = wall_clock64();
  GridG g;
  g.lane = threadIdx.x & 63;
  g.gwave = blockIdx.x * kWPB + (threadIdx.x >> 6);
  g.nwaves = gridDim.x * kWPB;
  Ops<GridG, VI, VC, MT> ops(g, a, s);
  ops.run();
  const uint32_t mm = s.op_maxmask;
  __shared__ double lds[kWPB][kRes];
#pragma unroll
  for (int k = 0; k < kRes; k++) {
    const double v = wave_reduce(ops.p[k], (mm >> k) & 1);
    if (g.lane == 0) lds[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < kRes) {
    const int k = threadIdx.x;
    const bool mx = (mm >> k) & 1;
    double acc = lds[0][k];
    for (int w = 1; w < kWPB; w++) acc = combine(acc, lds[w][k], mx);
    a.partials[(size_t)k * gridDim.x + blockIdx.x] = acc;
  }
}

// Combines the step kernel's per-block partials (slot-major: slot k of
// block b at k * nblocks + b; wave w owns slots 2w and 2w+1; lane l sums
// blocks l, l+64, ... in that order, then the lanes combine by butterfly),
// then advances the state machine on one thread with the state record held
// in LDS.  Called by a workgroup of 256 threads: by the control kernel, and
// by the last workgroup of the fused step kernel.  The partials are loaded
// kChunk per lane and slot at a time; the combining order does not depend
// on kChunk.  With kChunk = kMaxBlocks / 64 every load (state words and all
// partials) is issued before the first wait, so the control kernel pays one
// memory latency instead of one per batch; the fused kernel takes smaller
// chunks so that its register count stays that of the step.
template <int kChunk, bool OutOfLine>
__device__ __forceinline__ void control_block(const Args& a, uint32_t nblocks,
                                              State& ls, double* res,
                                              uint64_t t_entry) {
  State* st = a.st;
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int kPer = kMaxBlocks / 64;
  static_assert(kPer % kChunk == 0, "chunks tile the partials");
  // The per-op counters (thread t holds word t of the four arrays) and
  // the phase timers (thread 0) are read in the same round as the
  // state, so updating them costs no extra memory round trip.
  uint64_t* pw = reinterpret_cast<uint64_t*>(a.prof);
  uint64_t pf = 0, cp[4] = {0, 0, 0, 0};
  if (threadIdx.x < 128) pf = pw[threadIdx.x];
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) cp[k] = a.prof->ctl_phase[k];
  const double* p0 = a.partials + (size_t)(2 * w) * nblocks;
  const double* p1 = p0 + nblocks;
  double v0[kChunk], v1[kChunk];
  auto load = [&](int j0) {
#pragma unroll
    for (int j = 0; j < kChunk; j++) {
      const uint32_t b = lane + 64 * (j0 + j);
      v0[j] = b < nblocks ? p0[b] : 0.0;
      v1[j] = b < nblocks ? p1[b] : 0.0;
    }
  };
  load(0);  // issued together with the state loads
  {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(st);
    uint64_t* dst = reinterpret_cast<uint64_t*>(&ls);
    for (uint32_t i = threadIdx.x; i < sizeof(State) / 8; i += 256)
      dst[i] = src[i];
  }
  __syncthreads();
  const uint32_t mm = ls.op_maxmask;
  const int32_t op = ls.op;
  const bool ran = op != OP_NONE;
  const bool live = ls.pc != PC_DONE && !ls.error;
  const bool mx0 = (mm >> (2 * w)) & 1, mx1 = (mm >> (2 * w + 1)) & 1;
  double acc0 = mx0 ? -HUGE_VAL : 0.0, acc1 = mx1 ? -HUGE_VAL : 0.0;
  if (ran && lane < nblocks) {
#pragma unroll 1
    for (int j0 = 0; j0 < kPer; j0 += kChunk) {
      if (j0 > 0) {
        if (lane + 64u * j0 >= nblocks) break;
        load(j0);
      }
#pragma unroll
      for (int j = 0; j < kChunk; j++) {
        if (j0 + j == 0) {
          acc0 = v0[0];
          acc1 = v1[0];
        } else if (lane + 64u * (j0 + j) < nblocks) {
          acc0 = combine(acc0, v0[j], mx0);
          acc1 = combine(acc1, v1[j], mx1);
        }
      }
    }
  }
  const uint64_t t_loaded = wall_clock64();
  acc0 = wave_reduce(acc0, mx0);
  acc1 = wave_reduce(acc1, mx1);
  if (lane == 0) {
    res[2 * w] = acc0;
    res[2 * w + 1] = acc1;
  }
  // Per-op counters: the four threads holding this op's words.
  if (threadIdx.x < 128 && (threadIdx.x & 31) == (uint32_t)op && live &&
      ran && op >= 0 && op < 32) {
    uint64_t d;
    switch (threadIdx.x >> 5) {
      case 0: d = 1; break;
      case 1: d = t_entry - ls.t_last; break;
      case 2: d = t_entry - ls.t_step; break;
      default: d = (uint64_t)alg_bytes(op, ls.op_flags, a); break;
    }
    pw[threadIdx.x] = pf + d;
  }
  __syncthreads();
  if (threadIdx.x == 0 && live) {
    const uint64_t t_comb = wall_clock64();
    Ctl c{*(LdsState*)&ls, a, (const LdsDouble*)res};
    c.advance<OutOfLine>();
    ls.t_last = wall_clock64();
    if (ran) {
      a.prof->ctl_phase[0] = cp[0] + (t_loaded - t_entry);
      a.prof->ctl_phase[1] = cp[1] + (t_comb - t_loaded);
      a.prof->ctl_phase[2] = cp[2] + (ls.t_last - t_comb);
      a.prof->ctl_phase[3] = cp[3] + 1;
    }
  }
  __syncthreads();
  if (live) {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&ls);
    uint64_t* dst = reinterpret_cast<uint64_t*>(st);
    for (uint32_t i = threadIdx.x; i < sizeof(State) / 8; i += 256)
      dst[i] = src[i];
  }
}

__global__ void __launch_bounds__(256) k_ctl(Args a, uint32_t nblocks) {
  __shared__ double res[kRes];
  __shared__ State ls;
  const uint64_t t_entry = wall_clock64();
  control_block<kMaxBlocks / 64, false>(a, nblocks, ls, res, t_entry);
}

// The step and the control step in one launch: every workgroup runs its
// share of the op and writes its partials, then takes a ticket; the last
// workgroup to arrive (its ticket equals the grid size less one) sees
// every partial (release/acquire at device scope around the ticket),
// resets the ticket and runs the control step exactly as the control
// kernel does, so the results are the same bit for bit and one launch and
// one kernel boundary per op disappear.  Only wave 0 wrote what the last
// workgroup reads (the partials, and block 0's start time), so only it
// releases.  Measured on the set-cover LP it is still slower than the
// (step, control) pair: a device-scope release writes back the whole L2
// of the XCD, once per workgroup (off by default; ACD_SOLVER_FUSED=1).
// Relaxed device-scope atomics alone (no fences) do not make the partials
// visible across XCDs here: the control step read stale values.  An op of
// OP_NONE runs nothing here (the host issues a control kernel if a live
// solve is ever left without an op).
template <class VI, class VC, class MT>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(4)))
k_fused(Args a, uint32_t* ticket) {
  static_assert(64 * kWPB == 256, "the control step runs on 256 threads");
  const State& s = *a.st;
  if (s.op == OP_NONE) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st->t_step = wall_clock64();
  __shared__ double lds[kWPB][kRes];
  __shared__ double res[kRes];
  __shared__ State ls;
  __shared__ uint32_t last;
  {
    GridG g;
    g.lane = threadIdx.x & 63;
    g.gwave = blockIdx.x * kWPB + (threadIdx.x >> 6);
    g.nwaves = gridDim.x * kWPB;
    Ops<GridG, VI, VC, MT> ops(g, a, s);
    ops.run();
    const uint32_t mm = s.op_maxmask;
#pragma unroll
    for (int k = 0; k < kRes; k++) {
      const double v = wave_reduce(ops.p[k], (mm >> k) & 1);
      if (g.lane == 0) lds[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kRes) {
      const int k = threadIdx.x;
      const bool mx = (mm >> k) & 1;
      double acc = lds[0][k];
      for (int w = 1; w < kWPB; w++) acc = combine(acc, lds[w][k], mx);
      a.partials[(size_t)k * gridDim.x + blockIdx.x] = acc;
    }
  }
  if (threadIdx.x < 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (threadIdx.x == 0)
      last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1
                 ? 1u
                 : 0u;
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (threadIdx.x == 0) *ticket = 0u;
  control_block<8, true>(a, gridDim.x, ls, res, wall_clock64());
}

// ---------------------------------------------------------------------------
// Batched solves: one workgroup per LP runs the same ops and control
// continuation with the workgroup as the group and __syncthreads as the
// barrier, with the LP's argument and state records in LDS.  A launch
// advances every unfinished LP by up to `budget` inner iterations (and at
// most kBatchOps ops), then saves its state, so launches can be repeated,
// and LPs added between them, until every LP is done.
//
// Every op is computed as V virtual waves (Args::vwaves: 8 for the LPs of
// the largest size class, whose operators have more than 64 slices, 4 for
// the others; a function of the LP alone): virtual wave v takes slices v,
// v + V, ... and reduces its lanes' partials on its own; the V results are
// combined in order.  A workgroup of WPB physical waves (1, 2, 4, or 8 for
// V = 8) runs virtual waves w, w + WPB, ... one after the other, so the
// sums, and the LP's whole trajectory, are the same for every WPB: the
// workgroup size is a pure speed choice.
constexpr uint32_t kBatchOps = 1u << 22;
constexpr uint32_t kVWaves = 4;     // virtual waves of most LPs
constexpr uint32_t kMaxVWaves = 8;  // and of the largest size class
constexpr uint32_t kBigSlices = 64; // more slices than this: kMaxVWaves

struct BlockG {
  // Row loop with the gathers and the next chunk's entry loads in flight
  // together (measured faster on LPs whose operators the last-level cache
  // holds: 799 against 760 solves/s on 4,096 LPs).
  static constexpr bool kPipelinedRows = true;
  uint32_t lane, gwave, nwaves;
};

// The control continuation is inlined, with the kernel held to 3 waves per
// SIMD (amdgpu_waves_per_eu(3): 168 VGPRs, a few spilled), i.e. 3 LP
// workgroups of 4 waves per CU.  Called out of line (Ctl::advance<true>)
// the kernel needs ~150 VGPRs without spills, but every control step pays
// the call, its register saves in scratch and the reloads of the Ctl
// members: measured on 4,096 LPs run to completion, 621 inlined against
// 604 solves/s called.  (With 32-iteration rounds and no register bound,
// inlining took ~240 VGPRs, 2 waves per SIMD: 336 against 344.)
constexpr bool kBatchCtlCall = false;

// Vector pairs that a batched LP may keep in LDS, in the order they are
// given LDS (first the candidate point z / zp that the A pass gathers,
// then the violation pairs the A' pass gathers and the momentum update
// reads and writes): {side (-1: n-side, +1: m-side), role, partner role};
// each pair is closed under the solver's pointer swaps (approx.c
// :461-477 and the speculative commit).
constexpr int kLdsPairs = 4;
__device__ constexpr int kLdsPair[kLdsPairs][3] = {
    {-1, NZ, NZP}, {1, MYV, MSY}, {1, MZV, MZPV}, {1, MXV, MSX}};
// Per-workgroup LDS, static records plus resident vectors: 3 LP
// workgroups per CU (the kernel's register-bound occupancy) fit in the
// 160 KB of LDS when each takes at most 53,760 bytes (52.5 KB, a whole
// number of allocation granules); the resident vectors get what the
// static records (~1.5 KB) leave.  One static record 16 bytes larger
// (53,776 bytes in all) dropped the kernel to 2 workgroups per CU: 350
// against 388 solves/s on 4,096 LPs.  Measured with a 51 KB vector
// budget: 3 per CU 387 solves/s, 76 KB (2 per CU) 365, 150 KB (1 per CU)
// 281; without residency 345.
constexpr uint32_t kLdsPerWorkgroup = 53760;

extern __shared__ __align__(16) unsigned char lds_vecs[];

// Copies the resident slots between their global homes and LDS
// (to_lds) or back; slot order as assigned by k_batch.  Each slot is
// copied in 16-byte words: a slot's LDS size is its length rounded up to
// 16 bytes, which the global vector's spare element covers (vectors hold
// length + 1 doubles), and the padding word goes back unchanged (ops
// write only indices below the length).  The loop is kept plain: every
// byte of the kernel's code competes for the instruction cache that two
// CUs share (unrolled 4 ways it added 3.9 KB), and with LPs run to
// completion the copies happen once per LP.
template <int WPB>
__device__ void lds_copy(const Args& la, const State& ls, uint32_t resident,
                         void* const* home, bool to_lds) {
  size_t off = 0;
  uint32_t k = 0;
  constexpr uint32_t kT = 64 * WPB;
  for (int q = 0; q < kLdsPairs; q++) {
    if (!((resident >> q) & 1)) continue;
    const bool nside = kLdsPair[q][0] < 0;
    const uint32_t len = nside ? la.n : la.m;
    const size_t bytes = (len * (size_t)la.si + 15) & ~(size_t)15;
    const uint32_t words = (uint32_t)(bytes / 16);
    for (int e = 0; e < 2; e++) {
      uint4* g = reinterpret_cast<uint4*>(home[k]);
      uint4* l = reinterpret_cast<uint4*>(lds_vecs + off);
      for (uint32_t i = threadIdx.x; i < words; i += kT) {
        if (to_lds)
          l[i] = g[i];
        else
          g[i] = l[i];
      }
      off += bytes;
      k++;
    }
  }
}

// LDS bytes and pair mask for an LP of m rows and n columns with
// si-byte iterates under the budget: pairs are taken in order, each one
// that still fits.
uint32_t lds_plan(uint32_t m, uint32_t n, uint32_t si, uint32_t budget,
                  uint32_t* bytes) {
  const size_t bn = 2 * (((size_t)n * si + 15) & ~(size_t)15);
  const size_t bm = 2 * (((size_t)m * si + 15) & ~(size_t)15);
  const size_t need[kLdsPairs] = {bn, bm, bm, bm};
  size_t used = 0;
  uint32_t mask = 0;
  for (int q = 0; q < kLdsPairs; q++) {
    if (used + need[q] > budget) continue;
    used += need[q];
    mask |= 1u << q;
  }
  *bytes = (uint32_t)used;
  return mask;
}
template <class VI, class VC, class MT, int WPB>
__global__ void __launch_bounds__(64 * WPB)
__attribute__((amdgpu_waves_per_eu(3)))
k_batch(const Args* __restrict__ args, State* states, const uint32_t* slots,
        uint64_t budget, uint32_t* done_count, uint64_t* t_done,
        uint64_t* bytes, uint32_t* flags, uint64_t* ticks) {
  static_assert(kMaxVWaves % WPB == 0, "virtual waves per physical wave");
  __shared__ Args la;
  __shared__ State ls;
  __shared__ double red[kMaxVWaves][kRes];
  __shared__ double res[kRes];
  const uint32_t lp = slots[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  {
    const uint64_t* sa = reinterpret_cast<const uint64_t*>(args + lp);
    uint64_t* da = reinterpret_cast<uint64_t*>(&la);
    for (uint32_t i = tid; i < sizeof(Args) / 8; i += 64 * WPB) da[i] = sa[i];
    const uint64_t* ss = reinterpret_cast<const uint64_t*>(states + lp);
    uint64_t* ds = reinterpret_cast<uint64_t*>(&ls);
    for (uint32_t i = tid; i < sizeof(State) / 8; i += 64 * WPB) ds[i] = ss[i];
    if (tid < kRes) res[tid] = 0.0;
  }
  __syncthreads();
  if (ls.pc == PC_DONE || ls.error) return;
  // Vector pairs kept in LDS for this launch (Args::resident): their
  // roles' pointers are redirected to LDS copies, so gathers and epilogue
  // accesses of those vectors cost LDS latency instead of L2 / last-level
  // cache latency; the ops are unchanged (they address vectors through
  // generic pointers).  Each pair is closed under the solver's pointer
  // swaps, so at the end every role again points into one of the pair's
  // LDS slots and is written back to that slot's global home.
  __shared__ void* home[2 * kLdsPairs];
  __shared__ uint32_t nslot;
  const uint32_t resident = la.resident;
  if (resident) {
    if (tid == 0) {
      uint32_t k = 0;
      size_t off = 0;
      for (int q = 0; q < kLdsPairs; q++) {
        if (!((resident >> q) & 1)) continue;
        const bool nside = kLdsPair[q][0] < 0;
        const uint32_t len = nside ? la.n : la.m;
        const size_t bytes = (len * (size_t)la.si + 15) & ~(size_t)15;
        for (int e = 0; e < 2; e++) {
          const int r = kLdsPair[q][1 + e];
          double*& ptr = nside ? ls.pn[r] : ls.pm[r];
          home[k] = ptr;
          ptr = reinterpret_cast<double*>(lds_vecs + off);
          off += bytes;
          k++;
        }
      }
      nslot = k;
    }
    __syncthreads();
    lds_copy<WPB>(la, ls, resident, home, true);
    __syncthreads();
  }
  const uint64_t stop_at = ls.inner_total + budget;
  uint64_t moved = 0;
  // Wall-clock ticks of this launch spent in ops (passes + reductions)
  // and in the control continuation, kept by thread 0.
  uint64_t t_ops = 0, t_ctl = 0, t_mark = wall_clock64();
  for (uint32_t guard = 0; guard < kBatchOps; guard++) {
    if (ls.op != OP_NONE) {
      const uint32_t mm = ls.op_maxmask;
      const uint32_t nv = la.vwaves;
      for (uint32_t v = tid >> 6; v < nv; v += WPB) {
        // Reload the LDS records in each pass rather than holding them in
        // registers across passes (which would spill).
        asm volatile("" ::: "memory");
        BlockG g;
        g.lane = tid & 63;
        g.gwave = v;
        g.nwaves = nv;
        Ops<BlockG, VI, VC, MT> ops(g, la, ls);
        ops.template run<true>();
#pragma unroll
        for (int k = 0; k < kRes; k++) {
          const double r = wave_reduce(ops.p[k], (mm >> k) & 1);
          if (g.lane == 0) red[v][k] = r;
        }
      }
      __syncthreads();
      if (tid < kRes) {
        const bool mx = (mm >> tid) & 1;
        double acc = red[0][tid];
        for (uint32_t w = 1; w < nv; w++)
          acc = combine(acc, red[w][tid], mx);
        res[tid] = acc;
      }
      __syncthreads();
    }
    if (tid == 0) {
      const uint64_t t1 = wall_clock64();
      t_ops += t1 - t_mark;
      if (ls.op != OP_NONE) moved += (uint64_t)alg_bytes(ls.op, ls.op_flags, la);
      Ctl c{*(LdsState*)&ls, la, (const LdsDouble*)res};
      c.advance<kBatchCtlCall, true>();
      t_mark = wall_clock64();
      t_ctl += t_mark - t1;
    }
    __syncthreads();
    if (ls.pc == PC_DONE || ls.error || ls.op == OP_NONE ||
        ls.inner_total >= stop_at)
      break;
  }
  if (resident) {
    __syncthreads();
    lds_copy<WPB>(la, ls, resident, home, false);
    if (tid == 0) {  // roles back to the global homes of their slots
      size_t off = 0;
      uint32_t k = 0;
      for (int q = 0; q < kLdsPairs; q++) {
        if (!((resident >> q) & 1)) continue;
        const bool nside = kLdsPair[q][0] < 0;
        const uint32_t len = nside ? la.n : la.m;
        const size_t bytes = (len * (size_t)la.si + 15) & ~(size_t)15;
        double* s0 = reinterpret_cast<double*>(lds_vecs + off);
        double* s1 = reinterpret_cast<double*>(lds_vecs + off + bytes);
        for (int e = 0; e < 2; e++) {
          const int r = kLdsPair[q][1 + e];
          double*& ptr = nside ? ls.pn[r] : ls.pm[r];
          ptr = (double*)(ptr == s0 ? home[k] : ptr == s1 ? home[k + 1] : ptr);
        }
        off += 2 * bytes;
        k += 2;
      }
    }
    __syncthreads();
  }
  {
    const uint64_t* ss = reinterpret_cast<const uint64_t*>(&ls);
    uint64_t* ds = reinterpret_cast<uint64_t*>(states + lp);
    for (uint32_t i = tid; i < sizeof(State) / 8; i += 64 * WPB) ds[i] = ss[i];
  }
  if (tid == 0) {
    bytes[lp] += moved;
    ticks[2 * lp] += t_ops;
    ticks[2 * lp + 1] += t_ctl;
    if (ls.pc == PC_DONE || ls.error) {
      t_done[lp] = wall_clock64();
      flags[lp] = 1u;
      atomicAdd(done_count, 1u);
    }
  }
}

__global__ void k_stamp(uint64_t* t) { *t = wall_clock64(); }

// LDS left for resident vectors: the per-workgroup total less the batched
// kernel's static LDS (its records do not depend on the instantiation;
// the largest of two is taken), in 16-byte units; 0 turns residency off.
uint32_t lds_vector_budget() {
  static const uint32_t budget = [] {
    const void* fns[2] = {
        reinterpret_cast<const void*>(&k_batch<double, double, double, 4>),
        reinterpret_cast<const void*>(&k_batch<double, float, uint16_t, 1>)};
    size_t st = 0;
    for (const void* f : fns) {
      hipFuncAttributes fa;
      if (hipFuncGetAttributes(&fa, f) != hipSuccess) return 0u;
      if (fa.sharedSizeBytes > st) st = fa.sharedSizeBytes;
    }
    return st + 1024 < kLdsPerWorkgroup
               ? (uint32_t)((kLdsPerWorkgroup - st) & ~(size_t)15)
               : 0u;
  }();
  return budget;
}

Sell view(const acd_sell& op) {
  Sell v;
  v.off = op.off;
  v.col = op.col;
  v.val = op.val;
  v.val32 = op.val32;
  v.val16 = op.val16;
  v.colsgn = op.colsgn;
  v.col16 = op.col16;
  v.nrows = op.nrows;
  v.nslices = op.nslices;
  return v;
}

}  // namespace

struct acd_dev_lp {
  const acd_dev_matrix* orig;
  const acd_dev_matrix* perm;
  uint32_t m, n;
  double* vec[ACD_V_COUNT];
  double* r;
  double *p_rhs, *p_w, *p_c, *p_l, *p_u, *p_v, *p_iv;
  double* vn[kNumVN];
  double* vm[kNumVM];
  State* st;
  Prof* prof;
  double* partials;
  uint32_t* ticket;  // arrivals of the fused step kernel's workgroups
  acd_event* ev;
  int blocks;
  hipGraphExec_t graph[kKeys];
  int have_graph[kKeys];
  int pkey;  // precision key the permuted problem was last written in
  // non-blocking solve (acd_dev_lp_async_*)
  hipStream_t as;
  hipEvent_t aev;
  State* ahs;  // pinned host copy of the state
  int akey, achunk, abusy;
  uint64_t alaunches;
};

namespace {

Args make_args(acd_dev_lp* lp, int key) {
  static const uint32_t kSizes[kKeys][3] = {
      {8, 8, 8}, {8, 4, 4}, {8, 4, 2}, {4, 4, 4}, {4, 4, 2}, {8, 4, 0},
      {8, 4, 4}};
  Args a;
  memset(&a, 0, sizeof a);
  a.m = lp->m;
  a.n = lp->n;
  a.A = view(lp->perm->a);
  a.AT = view(lp->perm->at);
  a.A0 = view(lp->orig->a);
  a.AT0 = view(lp->orig->at);
  a.row_idx = lp->perm->row_idx;
  a.col_idx = lp->perm->col_idx;
  a.col_ridx = lp->perm->col_ridx;
  a.beta = lp->perm->a.rowlen;
  a.rhs = lp->vec[ACD_V_RHS];
  a.weight = lp->vec[ACD_V_WEIGHT];
  a.linc = lp->vec[ACD_V_LINEAR];
  a.lower = lp->vec[ACD_V_LOWER];
  a.upper = lp->vec[ACD_V_UPPER];
  a.alm_c = lp->vec[ACD_V_ALM_LINEAR];
  a.lam_lo = lp->vec[ACD_V_LAM_LO];
  a.lam_hi = lp->vec[ACD_V_LAM_HI];
  a.x_io = lp->vec[ACD_V_X];
  a.lam = lp->vec[ACD_V_LAMBDA];
  a.r = lp->r;
  a.p_rhs = lp->p_rhs;
  a.p_w = lp->p_w;
  a.p_c = lp->p_c;
  a.p_l = lp->p_l;
  a.p_u = lp->p_u;
  a.p_v = lp->p_v;
  a.p_iv = lp->p_iv;
  a.st = lp->st;
  a.prof = lp->prof;
  a.partials = lp->partials;
  a.ev = lp->ev;
  a.ev_cap = kEventCap;
  a.nnz = lp->orig->nnz;
  a.vwaves = std::max(lp->perm->a.nslices, lp->perm->at.nslices) > kBigSlices
                 ? kMaxVWaves
                 : kVWaves;
  a.si = kSizes[key][0];
  a.sc = kSizes[key][1];
  a.sv = kSizes[key][2];
  a.sx = key == K_MIX_C16 ? 2 : 4;
  return a;
}

// Enough wavefronts for every slice of the largest operator, capped at
// what the chip keeps resident at once (grid-stride beyond that).
int pick_blocks(const acd_dev_matrix* orig, const acd_dev_matrix* perm) {
  uint32_t slices = perm->at.nslices;
  if (perm->a.nslices > slices) slices = perm->a.nslices;
  if (orig->at.nslices > slices) slices = orig->at.nslices;
  if (orig->a.nslices > slices) slices = orig->a.nslices;
  int want = (int)((slices + kWPB - 1) / kWPB);
  // One residency round: the step kernel runs 4 waves per SIMD (its
  // register count), i.e. 4 workgroups of 4 waves per CU; measured on the
  // set-cover LP, 4 per CU beat 2, 6 and 8 per CU.
  int cap = 4 * acd_dev_cu_count();
  if (const char* e = getenv("ACD_STEP_BLOCKS")) {  // experiments
    const int v = atoi(e);
    if (v >= 64 && v <= (int)kMaxBlocks) cap = v;
  }
  if (want > 64) want = (want + 7) / 8 * 8;  // whole XCD rounds
  if (cap < 64) cap = 64;
  if (cap > (int)kMaxBlocks) cap = (int)kMaxBlocks;
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  return want;
}

void launch_step(int key, int blocks, hipStream_t st, const Args& a) {
  const dim3 grid(blocks), block(64 * kWPB);
  switch (key) {
    case K_MIX_F32: k_step<double, float, float><<<grid, block, 0, st>>>(a); break;
    case K_MIX_BF16:
      k_step<double, float, uint16_t><<<grid, block, 0, st>>>(a);
      break;
    case K_F32: k_step<float, float, float><<<grid, block, 0, st>>>(a); break;
    case K_F32_BF16:
      k_step<float, float, uint16_t><<<grid, block, 0, st>>>(a);
      break;
    case K_MIX_SGN: k_step<double, float, SgnT><<<grid, block, 0, st>>>(a); break;
    default: k_step<double, double, double><<<grid, block, 0, st>>>(a); break;
  }
}

void launch_fused(int key, int blocks, hipStream_t st, const Args& a,
                  uint32_t* ticket) {
  const dim3 grid(blocks), block(64 * kWPB);
  switch (key) {
    case K_MIX_F32:
      k_fused<double, float, float><<<grid, block, 0, st>>>(a, ticket);
      break;
    case K_MIX_BF16:
      k_fused<double, float, uint16_t><<<grid, block, 0, st>>>(a, ticket);
      break;
    case K_F32:
      k_fused<float, float, float><<<grid, block, 0, st>>>(a, ticket);
      break;
    case K_F32_BF16:
      k_fused<float, float, uint16_t><<<grid, block, 0, st>>>(a, ticket);
      break;
    case K_MIX_SGN:
      k_fused<double, float, SgnT><<<grid, block, 0, st>>>(a, ticket);
      break;
    default:
      k_fused<double, double, double><<<grid, block, 0, st>>>(a, ticket);
      break;
  }
}

// ACD_SOLVER_FUSED=1 runs each op as one fused launch instead of a step
// kernel followed by a control kernel (same results bit for bit; slower
// here, see k_fused).
bool use_fused() {
  static int on = -1;
  if (on < 0) {
    const char* e = getenv("ACD_SOLVER_FUSED");
    on = e && e[0] == '1';
  }
  return on;
}

// One op: a fused launch, or a (step, control) pair.
void launch_op(acd_dev_lp* lp, int key, hipStream_t st, const Args& a) {
  if (use_fused()) {
    launch_fused(key, lp->blocks, st, a, lp->ticket);
  } else {
    launch_step(key, lp->blocks, st, a);
    k_ctl<<<1, 256, 0, st>>>(a, (uint32_t)lp->blocks);
  }
}

inline uint64_t launches_per_op() { return use_fused() ? 1 : 2; }

// The fused kernel runs nothing for OP_NONE; a live solve whose control
// step issued no op (its continuation guard ran out) is advanced by a
// control kernel before the next chunk.
void kick_if_idle(acd_dev_lp* lp, const State* hs, hipStream_t st,
                  const Args& a) {
  if (use_fused() && hs->op == OP_NONE && hs->pc != PC_DONE && !hs->error)
    k_ctl<<<1, 256, 0, st>>>(a, (uint32_t)lp->blocks);
}

int build_graph(acd_dev_lp* lp, int key) {
  if (lp->have_graph[key]) return 0;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return -1;
  Args a = make_args(lp, key);
  hipGraph_t g;
  int rc = 0;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) !=
      hipSuccess) {
    rc = -1;
  } else {
    for (int i = 0; i < kPairs; i++) launch_op(lp, key, cs, a);
    if (hipStreamEndCapture(cs, &g) != hipSuccess) rc = -1;
    if (!rc && hipGraphInstantiate(&lp->graph[key], g, nullptr, nullptr,
                                   0) != hipSuccess)
      rc = -1;
    if (!rc) hipGraphDestroy(g);
  }
  hipStreamDestroy(cs);
  if (rc) {
    acd_dev_record_error(rc, "solver graph capture");
    return rc;
  }
  lp->have_graph[key] = 1;
  return 0;
}

void init_state(State* s, const acd_dev_lp* lp, const acd_solve_request* q) {
  memset(s, 0, sizeof *s);
  for (int i = 0; i < kNumVN; i++) s->pn[i] = lp->vn[i];
  for (int i = 0; i < kNumVM; i++) s->pm[i] = lp->vm[i];
  s->mode = q->mode;
  s->log = q->log;
  s->niter = q->niter;
  s->max_pg = q->max_pg;
  s->max_value = q->max_value;
  s->min_delta = q->min_delta;
  s->offset = q->offset;
  s->period = q->period;
  s->theta = 1;
  s->step_length = 1;
  s->stop_violation = q->stop_violation;
  s->stop_pg = q->stop_pg;
  s->precision_floor = q->precision_floor;
  s->inner_min_delta = q->inner_min_delta;
  s->mixed = q->precision == ACD_DEV_FP32;
  s->stop_kkt = q->stop_kkt;
  s->kkt = NAN;
  s->op = OP_NONE;
  switch (q->mode) {
    case ACD_MODE_APPROX: s->pc = PC_A_INIT; break;
    case ACD_MODE_MPG: s->pc = PC_M0; break;
    case ACD_MODE_PCD: {
      s->pc = PC_P0;
      const double pr = q->pcd_p > 1 ? 1.0 : q->pcd_p;
      s->pcd_thresh = pr >= 1 ? 0xffffffffu : (uint32_t)(pr * 4294967296.0);
      s->pcd_steps = q->pcd_steps ? q->pcd_steps : 1;
      s->pcd_seed = q->pcd_seed;
      s->pcd_p = pr;
      break;
    }
    case ACD_MODE_ALM:
      s->pc = PC_L0;
      s->outer_niter = q->niter;
      break;
    default: s->pc = PC_U0; break;
  }
}

// ACD_SOLVER_GRAPHS=0 launches the same (step, control) pairs directly on
// the stream instead of replaying captured graphs (for tools that do not
// trace graph kernel nodes).
bool use_graphs() {
  static int on = -1;
  if (on < 0) {
    const char* e = getenv("ACD_SOLVER_GRAPHS");
    on = !(e && e[0] == '0');
  }
  return on;
}

int run_to_completion(acd_dev_lp* lp, int key, State* hs,
                      acd_solve_result* res, acd_event_sink sink, void* ctx,
                      bool graphs = true) {
  graphs = graphs && use_graphs();
  if (graphs && build_graph(lp, key)) return -1;
  Args a = make_args(lp, key);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  acd_event* evbuf = (acd_event*)malloc(sizeof(acd_event) * kEventCap);
  int rc = 0;
  uint64_t launches = 0;
  if (hipMemcpyAsync(lp->st, hs, sizeof(State), hipMemcpyHostToDevice, S()) !=
          hipSuccess ||
      hipMemsetAsync(lp->prof, 0, sizeof(Prof), S()) != hipSuccess)
    rc = -1;
  hipEventRecord(e0, S());
  if (!rc) k_ctl<<<1, 256, 0, S()>>>(a, (uint32_t)lp->blocks);
  int chunk = 1;
  while (!rc) {
    for (int c = 0; c < chunk && !rc; c++) {
      if (graphs) {
        if (hipGraphLaunch(lp->graph[key], S()) != hipSuccess) rc = -1;
      } else {
        for (int i = 0; i < kPairs; i++) launch_op(lp, key, S(), a);
        if (hipGetLastError() != hipSuccess) rc = -1;
      }
      launches += launches_per_op() * kPairs;
    }
    if (rc) {
      acd_dev_record_error(rc, "solver graph launch");
      break;
    }
    if (hipMemcpyAsync(hs, lp->st, sizeof(State), hipMemcpyDeviceToHost,
                       S()) != hipSuccess ||
        hipStreamSynchronize(S()) != hipSuccess) {
      acd_dev_record_error(-1, "solver chunk");
      rc = -1;
      break;
    }
    if (hs->events) {
      const uint64_t ne = hs->events < kEventCap ? hs->events : kEventCap;
      hipMemcpy(evbuf, lp->ev, sizeof(acd_event) * ne, hipMemcpyDeviceToHost);
      if (sink)
        for (uint64_t i = 0; i < ne; i++) sink(ctx, &evbuf[i]);
      hs->events = 0;
      hs->events_full = 0;
      hipMemcpyAsync(&lp->st->events, &hs->events, sizeof(uint64_t),
                     hipMemcpyHostToDevice, S());
      hipMemcpyAsync(&lp->st->events_full, &hs->events_full, sizeof(int32_t),
                     hipMemcpyHostToDevice, S());
    }
    if (hs->error || hs->pc == PC_DONE) break;
    kick_if_idle(lp, hs, S(), a);
    if (chunk < 64) chunk *= 2;
  }
  hipEventRecord(e1, S());
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  free(evbuf);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  {
    Prof pf;
    memset(&pf, 0, sizeof pf);
    hipMemcpy(&pf, lp->prof, sizeof pf, hipMemcpyDeviceToHost);
    int khz = 100000;
    hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                          acd_dev_device());
    for (int i = 0; i < 32; i++) {
      g_prof_count[i] = pf.op_count[i];
      g_prof_us[i] = khz > 0 ? 1e3 * (double)pf.op_ticks[i] / khz : 0.0;
      g_prof_step_us[i] =
          khz > 0 ? 1e3 * (double)pf.step_ticks[i] / khz : 0.0;
      g_prof_bytes[i] = (double)pf.op_bytes[i];
    }
    for (int i = 0; i < 3; i++)
      g_ctl_us[i] = khz > 0 ? 1e3 * (double)pf.ctl_phase[i] / khz : 0.0;
    g_ctl_us[3] = (double)pf.ctl_phase[3];
  }
  if (res) {
    res->error = hs->error ? hs->error : (rc ? ACD_ERR_LAUNCH : 0);
    res->launches = launches;
    res->device_ms = ms;
    res->inner_iterations = hs->inner_total;
    res->outer_iterations = hs->outer;
    res->kkt = hs->kkt;
    if (hs->mode == ACD_MODE_ALM) {
      res->ret = hs->alm_ret;
      for (int i = 0; i < 3; i++) res->diag[i] = hs->alm_diag[i];
      res->diag[3] = res->diag[4] = 0;
    } else {
      res->ret = hs->ret;
      for (int i = 0; i < 5; i++) res->diag[i] = hs->diag[i];
    }
  }
  return rc;
}

// ACD_DEV_MIXED keeps the iterates in FP64 and the working copy's constant
// vectors in FP32; ACD_DEV_FP32 also stores the iterates in FP32.  Both
// read matrix values in BF16 when every value is exact there, else FP32.
int precision_key(acd_dev_lp* lp, int precision) {
  if (precision != ACD_DEV_MIXED && precision != ACD_DEV_FP32) return K_F64;
  acd_dev_matrix* pm = const_cast<acd_dev_matrix*>(lp->perm);
  acd_dev_matrix* om = const_cast<acd_dev_matrix*>(lp->orig);
  acd_sell* ops[4] = {&pm->a, &pm->at, &om->a, &om->at};
  bool bf = true;
  for (acd_sell* op : ops) {
    const int r = acd_dev_sell_make_bf16(op);
    if (r < 0) return -1;
    bf = bf && r == 1;
  }
  if (!bf)
    for (acd_sell* op : ops)
      if (acd_dev_sell_make_fp32(op)) return -1;
  if (precision == ACD_DEV_FP32) return bf ? K_F32_BF16 : K_F32;
  // Operators whose values are all -1, 0 or +1 (set cover, network flow)
  // drop the value stream altogether; the products are the BF16 form's
  // bit for bit (ACD_VALUE_FREE=0 keeps the BF16 form).
  const char* vf = getenv("ACD_VALUE_FREE");
  if (bf && !(vf && vf[0] == '0')) {
    bool sg = true;
    for (acd_sell* op : ops) {
      const int r = acd_dev_sell_make_sign(op);
      if (r < 0) return -1;
      sg = sg && r == 1;
    }
    if (sg) return K_MIX_SGN;
  }
  return bf ? K_MIX_BF16 : K_MIX_F32;
}

// The batch kernel's buffers, shared by every launch of one batch or pool.
struct BatchBufs {
  const Args* args;
  State* st;
  uint32_t* done;
  uint64_t* t_done;
  uint64_t* bytes;
  uint32_t* flags;
  uint64_t* ticks;  // [2 per LP]: ops, control
};

template <int WPB>
void launch_batch_w(int key, uint32_t nblocks, const uint32_t* slots,
                    uint64_t budget, const BatchBufs& b, hipStream_t st,
                    uint32_t shmem) {
  const dim3 grid(nblocks), block(64 * WPB);
#define ACD_BATCH_ARGS \
  b.args, b.st, slots, budget, b.done, b.t_done, b.bytes, b.flags, b.ticks
  switch (key) {
    case K_MIX_F32:
      k_batch<double, float, float, WPB><<<grid, block, shmem, st>>>(
          ACD_BATCH_ARGS);
      break;
    case K_MIX_BF16:
      k_batch<double, float, uint16_t, WPB><<<grid, block, shmem, st>>>(
          ACD_BATCH_ARGS);
      break;
    case K_MIX_C16:
      k_batch<double, float, F32C16, WPB><<<grid, block, shmem, st>>>(
          ACD_BATCH_ARGS);
      break;
    case K_F32:
      k_batch<float, float, float, WPB><<<grid, block, shmem, st>>>(
          ACD_BATCH_ARGS);
      break;
    case K_F32_BF16:
      k_batch<float, float, uint16_t, WPB><<<grid, block, shmem, st>>>(
          ACD_BATCH_ARGS);
      break;
    default:
      k_batch<double, double, double, WPB><<<grid, block, shmem, st>>>(
          ACD_BATCH_ARGS);
      break;
  }
#undef ACD_BATCH_ARGS
}

// Waves per LP workgroup: 1, 2 or 4 (the same results for each).
void launch_batch(int key, int wpb, uint32_t nblocks, const uint32_t* slots,
                  uint64_t budget, const BatchBufs& b, hipStream_t st,
                  uint32_t shmem) {
  if (wpb == 1)
    launch_batch_w<1>(key, nblocks, slots, budget, b, st, shmem);
  else if (wpb == 2)
    launch_batch_w<2>(key, nblocks, slots, budget, b, st, shmem);
  else if (wpb == 8)
    launch_batch_w<8>(key, nblocks, slots, budget, b, st, shmem);
  else
    launch_batch_w<4>(key, nblocks, slots, budget, b, st, shmem);
}

// ---- workgroup-size groups ------------------------------------------------
// LPs are put in size classes by their wider operator's slice count
// (class k: at most 2^k slices of 64 rows; the last class holds the rest),
// and each class runs with the workgroup size of the current
// configuration.  LPs of one workgroup size share a launch; the (at most
// four) launches of a round run on their own streams so that they
// overlap.
constexpr int kClasses = ACD_DEV_BATCH_CLASSES;
constexpr int kWpbChoices = 4;
constexpr int kWpbOf[kWpbChoices] = {1, 2, 4, 8};

// The tuner's candidate configurations: workgroup size (index into
// kWpbOf) per size class.  Fewer waves per LP cost an LP more time per
// op (its slices run in more passes) but leave more of the machine to
// the others; 4 waves is the latency-optimal size of the 4-virtual-wave
// classes.  The largest class (8 virtual waves) runs 8 waves in every
// candidate: these LPs need the most iterations and set the batch's
// makespan (measured on 4,096 LPs: 8 virtual waves for them, with twice
// the LDS and launched first, 765 against 619 solves/s).
constexpr int kConfigs = 5;
constexpr uint8_t kConfig[kConfigs][kClasses] = {
    {2, 2, 2, 2, 2, 2, 2, 3},  // 4 waves (8 for the largest class)
    {1, 1, 1, 2, 2, 2, 2, 3},  // 2 waves up to 4 slices, 4 above
    {1, 1, 1, 1, 2, 2, 2, 3},  // 2 waves up to 8 slices, 4 above
    {1, 1, 1, 1, 1, 2, 2, 3},  // 2 waves up to 16 slices, 4 above
    {0, 0, 1, 1, 2, 2, 2, 3},  // 1 wave up to 2 slices, 2 up to 8, then 4
};

int size_class(const acd_dev_lp* lp) {
  const uint32_t sl = std::max(lp->perm->a.nslices, lp->perm->at.nslices);
  int c = 0;
  while (c < kClasses - 1 && (1u << c) < sl) c++;
  return c;
}

// Configuration picked by the last tuned batch in this process (batches
// too small to tune on and the serving pool use it).
int g_config = 0;

int forced_choice() {
  const char* e = getenv("ACD_BATCH_WAVES");
  if (!e) return -1;
  const int w = atoi(e);
  for (int i = 0; i < kWpbChoices; i++)
    if (kWpbOf[i] == w) return i;
  return -1;
}

double now_s() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

// 8 physical waves only for LPs of 8 virtual waves (the largest class).
int lp_choice(int cls, int forced, int cfg) {
  const int c = forced >= 0 ? forced : kConfig[cfg][cls];
  return c == kWpbChoices - 1 && cls < kClasses - 1 ? kWpbChoices - 2 : c;
}

struct GroupStreams {
  hipStream_t s[kWpbChoices];
  hipEvent_t fork, join[kWpbChoices];
  bool ok;
};

GroupStreams* group_streams() {
  static GroupStreams gs = {};
  static bool init = false;
  if (!init) {
    init = true;
    gs.ok = hipEventCreateWithFlags(&gs.fork, hipEventDisableTiming) ==
            hipSuccess;
    for (int i = 0; i < kWpbChoices && gs.ok; i++)
      gs.ok = hipStreamCreateWithFlags(&gs.s[i], hipStreamNonBlocking) ==
                  hipSuccess &&
              hipEventCreateWithFlags(&gs.join[i], hipEventDisableTiming) ==
                  hipSuccess;
  }
  return gs.ok ? &gs : nullptr;
}

// One round: `launches` launches for each non-empty group, the groups on
// their own streams, ordered after and joined back into the library
// stream.  A group with a captured graph replays it instead.
struct Group {
  uint32_t n;          // LPs in the group
  uint32_t shmem;      // dynamic LDS of its launches (resident vectors)
  uint32_t* d_slots;   // its slot list (device)
  hipGraphExec_t exec; // captured `launches` launches, if any
  bool have_exec;
};

int run_round(int key, Group* grp, uint32_t launches, uint64_t budget,
              const BatchBufs& b, bool graphs, bool concurrent,
              hipEvent_t start = nullptr) {
  // One non-empty group (or concurrent = false) runs on the library
  // stream itself: the group streams are created only when sizes are
  // mixed in a batch (extra streams share the device's few hardware
  // queues with the request loop's large-LP streams, and a batch launch
  // queued behind a large LP's graph waits for it).
  int groups = 0;
  for (int i = 0; i < kWpbChoices; i++) groups += grp[i].n != 0;
  GroupStreams* gs = nullptr;
  if (groups > 1 && concurrent) {
    gs = group_streams();
    if (!gs) return -1;
  }
  // Graphs are captured (host work) before `start` is recorded, so the
  // round's event interval covers the launches only.
  for (int i = 0; i < kWpbChoices && graphs; i++) {
    Group& g = grp[i];
    if (!g.n || g.have_exec) continue;
    hipStream_t st = gs ? gs->s[i] : S();
    hipGraph_t graph;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) !=
        hipSuccess)
      return -1;
    for (uint32_t l = 0; l < launches; l++)
      launch_batch(key, kWpbOf[i], g.n, g.d_slots, budget, b, st, g.shmem);
    if (hipStreamEndCapture(st, &graph) != hipSuccess) return -1;
    const hipError_t e =
        hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return -1;
    g.have_exec = true;
  }
  if (start && hipEventRecord(start, S()) != hipSuccess) return -1;
  if (gs && hipEventRecord(gs->fork, S()) != hipSuccess) return -1;
  // Widest workgroups first: their LPs (the largest) start at once
  // instead of waiting for slots the other groups' workgroups fill
  // (measured: 765 against 665 solves/s on 4,096 LPs).
  for (int i = kWpbChoices - 1; i >= 0; i--) {
    Group& g = grp[i];
    if (!g.n) continue;
    hipStream_t st = gs ? gs->s[i] : S();
    if (gs && hipStreamWaitEvent(st, gs->fork, 0) != hipSuccess) return -1;
    if (graphs) {
      if (hipGraphLaunch(g.exec, st) != hipSuccess) return -1;
    } else {
      for (uint32_t l = 0; l < launches; l++)
        launch_batch(key, kWpbOf[i], g.n, g.d_slots, budget, b, st, g.shmem);
    }
    if (gs && (hipEventRecord(gs->join[i], st) != hipSuccess ||
               hipStreamWaitEvent(S(), gs->join[i], 0) != hipSuccess))
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

void drop_graphs(Group* grp) {
  for (int i = 0; i < kWpbChoices; i++)
    if (grp[i].have_exec) {
      hipGraphExecDestroy(grp[i].exec);
      grp[i].have_exec = false;
    }
}

// Split `live` (in launch order) into the four workgroup-size groups by
// choice[i] (index into kWpbOf) and upload the lists, group 0 first, into
// d_slots.  Order inside a group is kept.
int upload_groups(const uint32_t* live, uint32_t n, const uint8_t* choice,
                  uint32_t* h_tmp, uint32_t* d_slots, Group* grp,
                  const uint32_t* lds_bytes = nullptr) {
  drop_graphs(grp);
  uint32_t k = 0;
  for (int i = 0; i < kWpbChoices; i++) {
    grp[i].d_slots = d_slots + k;
    grp[i].shmem = 0;
    uint32_t c = 0;
    for (uint32_t j = 0; j < n; j++)
      if (choice[live[j]] == i) {
        h_tmp[k + c++] = live[j];
        if (lds_bytes && lds_bytes[live[j]] > grp[i].shmem)
          grp[i].shmem = lds_bytes[live[j]];
      }
    grp[i].n = c;
    k += c;
  }
  if (n && hipMemcpyAsync(d_slots, h_tmp, n * sizeof(uint32_t),
                          hipMemcpyHostToDevice, S()) != hipSuccess)
    return -1;
  return 0;
}

// Replays the log events a batch of one has written since the last drain
// (its state record at d_st) and resets the ring, so a solve paused on a
// full ring goes on at the next launch.
int drain_events(acd_dev_lp* lp, State* d_st, acd_event_sink sink,
                 void* ctx) {
  State hs;
  if (hipMemcpy(&hs, d_st, sizeof hs, hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  if (!hs.events) return 0;
  const uint64_t ne = hs.events < kEventCap ? hs.events : kEventCap;
  acd_event* buf = (acd_event*)malloc(sizeof(acd_event) * ne);
  if (!buf || hipMemcpy(buf, lp->ev, sizeof(acd_event) * ne,
                        hipMemcpyDeviceToHost) != hipSuccess) {
    free(buf);
    return -1;
  }
  for (uint64_t i = 0; i < ne; i++) sink(ctx, &buf[i]);
  free(buf);
  const uint64_t zero = 0;
  const int32_t zero32 = 0;
  if (hipMemcpy(&d_st->events, &zero, sizeof zero, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(&d_st->events_full, &zero32, sizeof zero32,
                hipMemcpyHostToDevice) != hipSuccess)
    return -1;
  return 0;
}

// Batches of at least kTuneMinLps LPs tune on their own first rounds
// (each candidate configuration runs one round, the fastest is kept) and
// again whenever the number of unfinished LPs has halved, down to
// kTuneMinLps; smaller batches use the last tuned configuration.
constexpr size_t kTuneMinLps = 192;
constexpr double kTuneMargin = 0.90;
constexpr uint32_t kRoundLaunches = 8;
constexpr uint64_t kBatchBudget = 32;
// Once the workgroup sizes are chosen, a batch runs every LP to
// completion in its workgroup: the hardware deals the workgroups out in
// launch order (largest LPs first) as slots free up, so every slot stays
// busy until the last LPs start, and an LP that needs many iterations
// runs at the speed of a resident workgroup instead of advancing 32
// iterations per round of the whole batch.  Measured on 4,096 LPs: 612
// against 487 solves/s with 32-iteration rounds (4,096 iterations: 505).
// The results do not depend on the budget (the state is saved exactly).
// 65,536 iterations cover every LP of that batch (at most 60,370) and
// bound one workgroup's share of a launch to a few seconds.
constexpr uint64_t kRunBudget = 1ull << 16;

// Iteration budget per launch after tuning: kRunBudget, or a fixed one
// for every round (experiments: ACD_BATCH_BUDGET).
uint64_t run_budget() {
  const char* e = getenv("ACD_BATCH_BUDGET");
  if (!e) return kRunBudget;
  const long v = atol(e);
  return v >= 1 ? (uint64_t)v : kRunBudget;
}

}  // namespace

extern "C" {

void acd_dev_last_profile(uint64_t* counts, double* us, int n) {
  for (int i = 0; i < n && i < 32; i++) {
    counts[i] = g_prof_count[i];
    us[i] = g_prof_us[i];
  }
}

void acd_dev_last_ctl_phases(double* us4) {
  for (int i = 0; i < 4; i++) us4[i] = g_ctl_us[i];
}

void acd_dev_last_profile_ex(double* step_us, double* bytes, int n) {
  for (int i = 0; i < n && i < 32; i++) {
    step_us[i] = g_prof_step_us[i];
    bytes[i] = g_prof_bytes[i];
  }
}

struct acd_dev_lp* acd_dev_lp_create(const struct acd_dev_matrix* orig,
                                     const struct acd_dev_matrix* perm) {
  if (acd_dev_init()) return nullptr;
  acd_dev_lp* lp = (acd_dev_lp*)calloc(1, sizeof(acd_dev_lp));
  if (!lp) return nullptr;
  lp->orig = orig;
  lp->perm = perm;
  lp->m = orig->nrows;
  lp->n = orig->ncols;
  size_t m = lp->m + 1, n = lp->n + 1;
  bool ok = true;
  auto al = [&](size_t count) -> double* {
    double* p = (double*)acd_dev_malloc(count * sizeof(double));
    if (!p) ok = false;
    return p;
  };
  const int is_m[ACD_V_COUNT] = {1, 1, 0, 0, 0, 0, 1, 1, 0, 1};
  for (int v = 0; v < ACD_V_COUNT; v++) lp->vec[v] = al(is_m[v] ? m : n);
  lp->r = al(m);
  lp->p_rhs = al(m);
  lp->p_w = al(m);
  lp->p_c = al(n);
  lp->p_l = al(n);
  lp->p_u = al(n);
  lp->p_v = al(n);
  lp->p_iv = al(n);
  for (int i = 0; i < kNumVN; i++) lp->vn[i] = al(n);
  for (int i = 0; i < kNumVM; i++) lp->vm[i] = al(m);
  lp->blocks = pick_blocks(orig, perm);
  lp->partials = al((size_t)lp->blocks * kRes);
  lp->ticket = (uint32_t*)acd_dev_malloc(sizeof(uint32_t));
  lp->st = (State*)acd_dev_malloc(sizeof(State));
  lp->prof = (Prof*)acd_dev_malloc(sizeof(Prof));
  lp->ev = (acd_event*)acd_dev_malloc(sizeof(acd_event) * kEventCap);
  if (!ok || !lp->st || !lp->prof || !lp->ev || !lp->ticket) {
    acd_dev_lp_free(lp);
    return nullptr;
  }
  return lp;
}

void acd_dev_lp_free(struct acd_dev_lp* lp) {
  if (!lp) return;
  for (int k = 0; k < kKeys; k++)
    if (lp->have_graph[k]) hipGraphExecDestroy(lp->graph[k]);
  if (lp->aev) hipEventDestroy(lp->aev);
  if (lp->ahs) hipHostFree(lp->ahs);
  if (lp->as) hipStreamDestroy(lp->as);
  for (int v = 0; v < ACD_V_COUNT; v++) acd_dev_free(lp->vec[v]);
  acd_dev_free(lp->r);
  acd_dev_free(lp->p_rhs);
  acd_dev_free(lp->p_w);
  acd_dev_free(lp->p_c);
  acd_dev_free(lp->p_l);
  acd_dev_free(lp->p_u);
  acd_dev_free(lp->p_v);
  acd_dev_free(lp->p_iv);
  for (int i = 0; i < kNumVN; i++) acd_dev_free(lp->vn[i]);
  for (int i = 0; i < kNumVM; i++) acd_dev_free(lp->vm[i]);
  acd_dev_free(lp->partials);
  acd_dev_free(lp->ticket);
  acd_dev_free(lp->st);
  acd_dev_free(lp->prof);
  acd_dev_free(lp->ev);
  free(lp);
}

double* acd_dev_lp_vec(struct acd_dev_lp* lp, int which) {
  return (which >= 0 && which < ACD_V_COUNT) ? lp->vec[which] : nullptr;
}

int acd_dev_approx_update(struct acd_dev_lp* lp) {
  acd_solve_request q;
  memset(&q, 0, sizeof q);
  State hs;
  init_state(&hs, lp, &q);
  acd_solve_result res;
  // Two ops: launched directly, no graph to capture.
  const int rc =
      run_to_completion(lp, K_F64, &hs, &res, nullptr, nullptr, false);
  lp->pkey = K_F64;
  return rc;
}

int acd_dev_solve(struct acd_dev_lp* lp, const struct acd_solve_request* q,
                  struct acd_solve_result* res, acd_event_sink sink,
                  void* ctx) {
  const int key = precision_key(lp, q->precision);
  if (key < 0) return -1;
  // APPROX/SPG solve the permuted problem as last written; rewrite it if
  // it was written in another precision.  ALM rewrites it every outer
  // iteration.
  if (q->mode != ACD_MODE_ALM && key != lp->pkey) {
    acd_solve_request u;
    memset(&u, 0, sizeof u);
    State hu;
    init_state(&hu, lp, &u);
    if (run_to_completion(lp, key, &hu, nullptr, nullptr, nullptr, false))
      return -1;
    lp->pkey = key;
  }
  State hs;
  init_state(&hs, lp, q);
  const int rc = run_to_completion(lp, key, &hs, res, sink, ctx);
  if (q->mode == ACD_MODE_ALM) lp->pkey = key;
  return rc;
}

int acd_dev_solve_batch(struct acd_dev_lp* const* lps, size_t count,
                        const struct acd_solve_request* q,
                        struct acd_solve_result* res, double* done_ms,
                        struct acd_batch_stats* stats, acd_event_sink sink,
                        void* ctx) {
  if (count == 0) return 0;
  // A log is replayed for a batch of one only (the LP's event ring).
  const bool logged = sink && q->log && count == 1;
  if (count > 0xffffffffu) return -1;
  // One kernel instantiation per launch: every LP runs in the same key
  // (BF16 values only when all of them are exact there).
  int key = K_F64;
  if (q->precision == ACD_DEV_MIXED || q->precision == ACD_DEV_FP32) {
    bool all_bf = true;
    for (size_t i = 0; i < count && all_bf; i++) {
      const int k = precision_key(lps[i], q->precision);
      if (k < 0) return -1;
      all_bf = k == K_MIX_BF16 || k == K_F32_BF16 || k == K_MIX_SGN;
    }
    if (!all_bf)
      for (size_t i = 0; i < count; i++) {
        acd_dev_matrix* pm = const_cast<acd_dev_matrix*>(lps[i]->perm);
        acd_dev_matrix* om = const_cast<acd_dev_matrix*>(lps[i]->orig);
        if (acd_dev_sell_make_fp32(&pm->a) || acd_dev_sell_make_fp32(&pm->at) ||
            acd_dev_sell_make_fp32(&om->a) || acd_dev_sell_make_fp32(&om->at))
          return -1;
      }
    key = q->precision == ACD_DEV_FP32 ? (all_bf ? K_F32_BF16 : K_F32)
                                       : (all_bf ? K_MIX_BF16 : K_MIX_F32);
    // FP32 values with 16-bit column indices when every operator of every
    // LP has at most 65,536 columns: 6 bytes per entry instead of 8, the
    // same products bit for bit (ACD_BATCH_COL16=0 keeps 32-bit indices).
    const char* c16 = getenv("ACD_BATCH_COL16");
    if (key == K_MIX_F32 && !(c16 && c16[0] == '0')) {
      bool narrow = true;
      for (size_t i = 0; i < count && narrow; i++)
        narrow = lps[i]->perm->a.ncols <= 65536u &&
                 lps[i]->perm->at.ncols <= 65536u &&
                 lps[i]->orig->a.ncols <= 65536u &&
                 lps[i]->orig->at.ncols <= 65536u;
      for (size_t i = 0; i < count && narrow; i++) {
        acd_dev_matrix* pm = const_cast<acd_dev_matrix*>(lps[i]->perm);
        acd_dev_matrix* om = const_cast<acd_dev_matrix*>(lps[i]->orig);
        acd_sell* ops[4] = {&pm->a, &pm->at, &om->a, &om->at};
        for (acd_sell* op : ops)
          if (acd_dev_sell_make_col16(op) != 1) return -1;
      }
      if (narrow) key = K_MIX_C16;
    }
  }
  Args* h_args = (Args*)calloc(count, sizeof(Args));
  State* h_st = (State*)calloc(count, sizeof(State));
  uint32_t* h_slots = (uint32_t*)malloc(count * sizeof(uint32_t));
  uint32_t* h_tmp = (uint32_t*)malloc(count * sizeof(uint32_t));
  uint64_t* h_done = (uint64_t*)calloc(count, sizeof(uint64_t));
  uint64_t* h_bytes = (uint64_t*)calloc(count, sizeof(uint64_t));
  uint8_t* cls = (uint8_t*)malloc(count);
  uint8_t* choice = (uint8_t*)malloc(count);
  uint32_t* lds_bytes = (uint32_t*)calloc(count, sizeof(uint32_t));
  Args* d_args = nullptr;
  State* d_st = nullptr;
  uint32_t *d_slots = nullptr, *d_count = nullptr;
  uint64_t *d_done = nullptr, *d_bytes = nullptr, *d_t0 = nullptr;
  uint32_t* d_flags = nullptr;
  uint64_t* d_ticks = nullptr;
  int rc = 0;
  hipEvent_t e0, e1, r0, r1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventCreate(&r0);
  hipEventCreate(&r1);
  double kernel_ms = 0;
  if (!h_args || !h_st || !h_slots || !h_tmp || !h_done || !h_bytes ||
      !cls || !choice || !lds_bytes ||
      hipMalloc((void**)&d_args, count * sizeof(Args)) != hipSuccess ||
      hipMalloc((void**)&d_st, count * sizeof(State)) != hipSuccess ||
      hipMalloc((void**)&d_slots, count * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&d_count, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&d_done, count * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void**)&d_bytes, count * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void**)&d_t0, sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void**)&d_flags, count * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&d_ticks, 2 * count * sizeof(uint64_t)) !=
          hipSuccess) {
    acd_dev_record_error(-1, "batch allocation");
    rc = -1;
  }
  // Workgroup size per LP: forced (ACD_BATCH_WAVES=1|2|4|8), or the
  // configuration the tuner picks on this batch's own rounds.
  const int forced = forced_choice();
  const bool tune = forced < 0 && count >= kTuneMinLps;
  const char* lds_env = getenv("ACD_BATCH_LDS");
  const bool lds_on = !(lds_env && lds_env[0] == '0');
  const bool tune_log = getenv("ACD_BATCH_TUNE_LOG") != nullptr;
  int cfg = tune ? 0 : g_config, first_cfg = -1;
  int trial = tune ? 0 : -1;  // configuration on trial, -1: none
  double trial_s[kConfigs] = {};
  uint32_t tuned_at = (uint32_t)count;
  uint64_t launches = 0;
  if (!rc) {
    for (size_t i = 0; i < count; i++) {
      h_args[i] = make_args(lps[i], key);
      if (!logged) {
        h_args[i].ev = nullptr;
        h_args[i].ev_cap = 0;
      }
      h_args[i].st = d_st + i;
      h_args[i].prof = nullptr;
      h_args[i].partials = nullptr;
      if (lds_on && q->mode == ACD_MODE_ALM)
        // An 8-wave workgroup takes two thirds of a CU's wave slots; it
        // gets twice the LDS of a 4-wave one (the CU then holds one of
        // each), enough for the largest LPs' candidate-point pair.
        h_args[i].resident = lds_plan(
            lps[i]->m, lps[i]->n, h_args[i].si,
            h_args[i].vwaves == kMaxVWaves
                ? lds_vector_budget() + kLdsPerWorkgroup
                : lds_vector_budget(),
            &lds_bytes[i]);
      acd_solve_request qi = *q;
      if (!logged) qi.log = 0;
      init_state(&h_st[i], lps[i], &qi);
      h_slots[i] = (uint32_t)i;
      cls[i] = (uint8_t)size_class(lps[i]);
      choice[i] = (uint8_t)lp_choice(cls[i], forced, cfg);
    }
    // Largest LPs first: workgroups are dispatched in launch order, so the
    // long solves start at once instead of trailing the batch.
    std::sort(h_slots, h_slots + count, [&](uint32_t x, uint32_t y) {
      const uint64_t nx = lps[x]->orig->nnz, ny = lps[y]->orig->nnz;
      return nx != ny ? nx > ny : x < y;
    });
    if (hipMemcpyAsync(d_args, h_args, count * sizeof(Args),
                       hipMemcpyHostToDevice, S()) != hipSuccess ||
        hipMemcpyAsync(d_st, h_st, count * sizeof(State),
                       hipMemcpyHostToDevice, S()) != hipSuccess ||
        hipMemsetAsync(d_count, 0, sizeof(uint32_t), S()) != hipSuccess ||
        hipMemsetAsync(d_bytes, 0, count * sizeof(uint64_t), S()) !=
            hipSuccess ||
        hipMemsetAsync(d_flags, 0, count * sizeof(uint32_t), S()) !=
            hipSuccess ||
        hipMemsetAsync(d_ticks, 0, 2 * count * sizeof(uint64_t), S()) !=
            hipSuccess ||
        hipStreamSynchronize(S()) != hipSuccess)
      rc = -1;
  }
  const BatchBufs bufs = {d_args,  d_st,    d_count, d_done,
                          d_bytes, d_flags, d_ticks};
  Group grp[kWpbChoices] = {};
  const bool graphs = use_graphs();
  // Relaunch over the unfinished LPs (compacted between rounds) until all
  // are done, 8 launches per group per round (one captured graph per group
  // while its list holds): 32 inner iterations per launch while the tuner
  // compares workgroup sizes, then run to completion.
  uint32_t active = (uint32_t)count;
  uint32_t* live = h_slots;
  uint32_t done = 0;
  const char* fixed_budget = getenv("ACD_BATCH_BUDGET");
  const uint64_t after_tuning = run_budget();
  uint64_t budget =
      trial >= 0 && !fixed_budget ? kBatchBudget : after_tuning;
  if (!rc &&
      upload_groups(live, active, choice, h_tmp, d_slots, grp, lds_bytes))
    rc = -1;
  if (!rc) {
    hipEventRecord(e0, S());
    k_stamp<<<1, 1, 0, S()>>>(d_t0);
    for (int round = 0; !rc && active; round++) {
      const double t = now_s();
      if (run_round(key, grp, kRoundLaunches, budget, bufs, graphs,
                    true, r0) ||
          hipEventRecord(r1, S()) != hipSuccess ||
          hipMemcpyAsync(&done, d_count, sizeof done, hipMemcpyDeviceToHost,
                         S()) != hipSuccess ||
          hipStreamSynchronize(S()) != hipSuccess) {
        acd_dev_record_error(-1, "batch launch");
        rc = -1;
        break;
      }
      {
        float rms = 0;
        if (hipEventElapsedTime(&rms, r0, r1) == hipSuccess) kernel_ms += rms;
      }
      if (logged && drain_events(lps[0], d_st, sink, ctx)) {
        rc = -1;
        break;
      }
      for (int i = 0; i < kWpbChoices; i++)
        if (grp[i].n) launches += kRoundLaunches;
      if (done >= count) break;
      // The tuner: each candidate configuration runs one round (all of
      // them advance every unfinished LP by the same iteration budget, so
      // the round times compare), the fastest is kept; again once half of
      // the LPs it was tuned on have finished.
      bool regroup = false;
      if (trial >= 0) {
        trial_s[trial] = now_s() - t;
        if (++trial < kConfigs) {
          cfg = trial;
        } else {
          // Another configuration replaces 4 waves everywhere only when
          // its round is at least 10% shorter: on 4,096 mixed-size LPs
          // the candidates' rounds differ by a few percent (noise can
          // reach 5%), and the ones that slow the small LPs down delay
          // their completion without speeding up the large ones.
          cfg = 0;
          for (int c = 1; c < kConfigs; c++)
            if (trial_s[c] < kTuneMargin * trial_s[cfg]) cfg = c;
          if (tune_log)
            fprintf(stderr, "batch tune at %u unfinished: %.2f %.2f %.2f "
                    "%.2f %.2f ms -> %d\n", (unsigned)(count - done),
                    1e3 * trial_s[0], 1e3 * trial_s[1], 1e3 * trial_s[2],
                    1e3 * trial_s[3], 1e3 * trial_s[4], cfg);
          if (first_cfg < 0) {
            first_cfg = cfg;
            g_config = cfg;
          }
          trial = -1;
          tuned_at = (uint32_t)(count - done);
        }
        regroup = true;
      } else if (tune && 2 * (count - done) <= tuned_at &&
                 count - done >= kTuneMinLps) {
        trial = 0;
        cfg = 0;
        regroup = true;
      }
      // Every 16 rounds, and when the configuration changes, drop
      // finished LPs from the launch lists.
      if (round % 16 == 15 || regroup) {
        if (hipMemcpy(h_st, d_st, count * sizeof(State),
                      hipMemcpyDeviceToHost) != hipSuccess) {
          rc = -1;
          break;
        }
        uint32_t k = 0;
        for (uint32_t i = 0; i < active; i++) {
          const uint32_t lp = live[i];
          if (h_st[lp].pc != PC_DONE && !h_st[lp].error) live[k++] = lp;
          choice[lp] = (uint8_t)lp_choice(cls[lp], forced, cfg);
        }
        active = k;
        // A new budget is captured with the new lists (upload_groups
        // drops the graphs).
        budget = trial >= 0 && !fixed_budget ? kBatchBudget : after_tuning;
        if (upload_groups(live, active, choice, h_tmp, d_slots, grp,
                          lds_bytes)) {
          rc = -1;
          break;
        }
      }
    }
    hipEventRecord(e1, S());
    hipEventSynchronize(e1);
  }
  drop_graphs(grp);
  float ms = 0;
  if (!rc) hipEventElapsedTime(&ms, e0, e1);
  uint64_t t0 = 0;
  if (!rc &&
      (hipMemcpy(h_st, d_st, count * sizeof(State), hipMemcpyDeviceToHost) !=
           hipSuccess ||
       hipMemcpy(h_done, d_done, count * sizeof(uint64_t),
                 hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(h_bytes, d_bytes, count * sizeof(uint64_t),
                 hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(&t0, d_t0, sizeof t0, hipMemcpyDeviceToHost) != hipSuccess))
    rc = -1;
  uint64_t tick_ops = 0, tick_ctl = 0;
  if (!rc) {
    uint64_t* h_ticks = (uint64_t*)malloc(2 * count * sizeof(uint64_t));
    if (h_ticks && hipMemcpy(h_ticks, d_ticks, 2 * count * sizeof(uint64_t),
                             hipMemcpyDeviceToHost) == hipSuccess)
      for (size_t i = 0; i < count; i++) {
        tick_ops += h_ticks[2 * i];
        tick_ctl += h_ticks[2 * i + 1];
      }
    free(h_ticks);
  }
  if (!rc) {
    int khz = 100000;
    hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                          acd_dev_device());
    uint64_t inner = 0, total_bytes = 0;
    for (size_t i = 0; i < count; i++) {
      const State& hs = h_st[i];
      acd_solve_result& r = res[i];
      memset(&r, 0, sizeof r);
      r.error = hs.error;
      r.ret = hs.alm_ret;
      for (int k = 0; k < 3; k++) r.diag[k] = hs.alm_diag[k];
      r.inner_iterations = hs.inner_total;
      r.outer_iterations = hs.outer;
      r.kkt = hs.kkt;
      r.launches = launches;
      r.device_ms = ms;
      if (done_ms)
        done_ms[i] = h_done[i] >= t0 && khz > 0
                         ? (double)(h_done[i] - t0) / khz
                         : -1.0;
      inner += hs.inner_total;
      total_bytes += h_bytes[i];
    }
    if (stats) {
      // The configuration of the first tuning (the bulk of the batch).
      const int c0 = first_cfg >= 0 ? first_cfg : cfg;
      for (int c = 0; c < kClasses; c++) stats->waves_per_lp[c] = 0;
      for (size_t i = 0; i < count; i++)
        stats->waves_per_lp[cls[i]] = kWpbOf[lp_choice(cls[i], forced, c0)];
      stats->device_ms = ms;
      stats->kernel_ms = kernel_ms;
      stats->lp_ops_ms = khz > 0 ? (double)tick_ops / khz : 0.0;
      stats->lp_ctl_ms = khz > 0 ? (double)tick_ctl / khz : 0.0;
      stats->launches = launches;
      stats->inner_iterations = inner;
      stats->alg_bytes = (double)total_bytes;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(r0);
  hipEventDestroy(r1);
  hipFree(d_args);
  hipFree(d_st);
  hipFree(d_slots);
  hipFree(d_count);
  hipFree(d_done);
  hipFree(d_bytes);
  hipFree(d_t0);
  hipFree(d_flags);
  hipFree(d_ticks);
  free(h_args);
  free(h_st);
  free(h_slots);
  free(h_tmp);
  free(h_done);
  free(h_bytes);
  free(cls);
  free(choice);
  free(lds_bytes);
  return rc;
}

// ---- slot pool for continuous batching ----------------------------------
// Each round splits the active slots by workgroup size (the tuned choice of
// the slot's size class) and launches the groups one after the other on
// the library stream.
struct acd_dev_pool {
  uint32_t cap;
  int key;
  Args* d_args;
  State* d_st;
  uint32_t *d_slots, *d_count, *d_flags;
  uint64_t *d_done, *d_bytes, *d_ticks;
  uint32_t* h_flags;  // pinned
  uint32_t* h_tmp;
  uint8_t* choice;    // per slot
};

struct acd_dev_pool* acd_dev_pool_create(uint32_t cap, int precision) {
  if (acd_dev_init() || cap == 0) return nullptr;
  acd_dev_pool* p = (acd_dev_pool*)calloc(1, sizeof(acd_dev_pool));
  if (!p) return nullptr;
  p->cap = cap;
  // One kernel instantiation for every LP that may arrive: reduced
  // precision uses FP32 matrix values (BF16 would need every future LP to
  // be exact there).
  p->key = precision == ACD_DEV_MIXED ? K_MIX_F32
           : precision == ACD_DEV_FP32 ? K_F32 : K_F64;
  p->h_tmp = (uint32_t*)malloc(cap * sizeof(uint32_t));
  p->choice = (uint8_t*)calloc(cap, 1);
  if (!p->h_tmp || !p->choice ||
      hipMalloc((void**)&p->d_args, cap * sizeof(Args)) != hipSuccess ||
      hipMalloc((void**)&p->d_st, cap * sizeof(State)) != hipSuccess ||
      hipMalloc((void**)&p->d_slots, cap * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&p->d_count, sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&p->d_flags, cap * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void**)&p->d_done, cap * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void**)&p->d_bytes, cap * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void**)&p->d_ticks, 2 * cap * sizeof(uint64_t)) !=
          hipSuccess ||
      hipHostMalloc((void**)&p->h_flags, cap * sizeof(uint32_t),
                    hipHostMallocDefault) != hipSuccess) {
    acd_dev_record_error(-1, "pool allocation");
    acd_dev_pool_free(p);
    return nullptr;
  }
  hipMemsetAsync(p->d_flags, 0, cap * sizeof(uint32_t), S());
  hipMemsetAsync(p->d_count, 0, sizeof(uint32_t), S());
  hipMemsetAsync(p->d_bytes, 0, cap * sizeof(uint64_t), S());
  hipMemsetAsync(p->d_ticks, 0, 2 * cap * sizeof(uint64_t), S());
  hipStreamSynchronize(S());
  return p;
}

void acd_dev_pool_free(struct acd_dev_pool* p) {
  if (!p) return;
  hipStreamSynchronize(S());
  hipFree(p->d_args);
  hipFree(p->d_st);
  hipFree(p->d_slots);
  hipFree(p->d_count);
  hipFree(p->d_flags);
  hipFree(p->d_done);
  hipFree(p->d_bytes);
  hipFree(p->d_ticks);
  if (p->h_flags) hipHostFree(p->h_flags);
  free(p->h_tmp);
  free(p->choice);
  free(p);
}

int acd_dev_pool_admit(struct acd_dev_pool* p, uint32_t slot,
                       struct acd_dev_lp* lp,
                       const struct acd_solve_request* q) {
  if (slot >= p->cap) return -1;
  if (p->key != K_F64) {
    acd_dev_matrix* pm = const_cast<acd_dev_matrix*>(lp->perm);
    acd_dev_matrix* om = const_cast<acd_dev_matrix*>(lp->orig);
    if (acd_dev_sell_make_fp32(&pm->a) || acd_dev_sell_make_fp32(&pm->at) ||
        acd_dev_sell_make_fp32(&om->a) || acd_dev_sell_make_fp32(&om->at))
      return -1;
  }
  const int forced = forced_choice();
  p->choice[slot] =
      (uint8_t)lp_choice(size_class(lp), forced, g_config);
  Args a = make_args(lp, p->key);
  a.ev = nullptr;
  a.ev_cap = 0;
  a.st = p->d_st + slot;
  a.prof = nullptr;
  a.partials = nullptr;
  State st;
  acd_solve_request qi = *q;
  qi.log = 0;
  init_state(&st, lp, &qi);
  const uint32_t zero = 0;
  if (hipMemcpyAsync(p->d_args + slot, &a, sizeof a, hipMemcpyHostToDevice,
                     S()) != hipSuccess ||
      hipMemcpyAsync(p->d_st + slot, &st, sizeof st, hipMemcpyHostToDevice,
                     S()) != hipSuccess ||
      hipMemcpyAsync(p->d_flags + slot, &zero, 4, hipMemcpyHostToDevice,
                     S()) != hipSuccess ||
      hipStreamSynchronize(S()) != hipSuccess) {
    acd_dev_record_error(-1, "pool admit");
    return -1;
  }
  return 0;
}

int acd_dev_pool_round(struct acd_dev_pool* p, const uint32_t* slots,
                       uint32_t n, uint32_t launches, uint64_t budget,
                       uint32_t** flags) {
  if (n > p->cap) return -1;
  for (uint32_t i = 0; i < n; i++)
    if (slots[i] >= p->cap) return -1;
  if (n) {
    Group grp[kWpbChoices] = {};
    const BatchBufs bufs = {p->d_args,  p->d_st,    p->d_count, p->d_done,
                            p->d_bytes, p->d_flags, p->d_ticks};
    if (upload_groups(slots, n, p->choice, p->h_tmp, p->d_slots, grp) ||
        run_round(p->key, grp, launches, budget, bufs, false, false)) {
      acd_dev_record_error(-1, "pool round");
      return -1;
    }
  }
  if (hipMemcpyAsync(p->h_flags, p->d_flags, p->cap * sizeof(uint32_t),
                     hipMemcpyDeviceToHost, S()) != hipSuccess ||
      hipStreamSynchronize(S()) != hipSuccess) {
    acd_dev_record_error(-1, "pool round");
    return -1;
  }
  if (flags) *flags = p->h_flags;
  return 0;
}

int acd_dev_pool_result(struct acd_dev_pool* p, uint32_t slot,
                        struct acd_solve_result* r) {
  State hs;
  if (slot >= p->cap ||
      hipMemcpyAsync(&hs, p->d_st + slot, sizeof hs, hipMemcpyDeviceToHost,
                     S()) != hipSuccess ||
      hipStreamSynchronize(S()) != hipSuccess)
    return -1;
  memset(r, 0, sizeof *r);
  r->error = hs.error;
  r->ret = hs.alm_ret;
  for (int k = 0; k < 3; k++) r->diag[k] = hs.alm_diag[k];
  r->inner_iterations = hs.inner_total;
  r->outer_iterations = hs.outer;
  r->kkt = hs.kkt;
  return 0;
}

// ---- non-blocking single-LP solve on its own stream ------------------------
// begin: state upload, graph, first chunk; poll: if the last chunk has
// finished, either report completion (1) or launch the next chunk (0).
int acd_dev_lp_async_begin(struct acd_dev_lp* lp,
                           const struct acd_solve_request* q) {
  if (!lp->as &&
      hipStreamCreateWithFlags(&lp->as, hipStreamNonBlocking) != hipSuccess)
    return -1;
  if (!lp->aev && hipEventCreateWithFlags(&lp->aev, hipEventDisableTiming) !=
                      hipSuccess)
    return -1;
  if (!lp->ahs && hipHostMalloc((void**)&lp->ahs, sizeof(State),
                                hipHostMallocDefault) != hipSuccess)
    return -1;
  const int key = precision_key(lp, q->precision);
  if (key < 0 || build_graph(lp, key)) return -1;
  acd_solve_request qi = *q;
  qi.log = 0;
  init_state(lp->ahs, lp, &qi);
  lp->akey = key;
  lp->achunk = 1;
  lp->alaunches = 0;
  Args a = make_args(lp, key);
  hipStream_t st = lp->as;
  if (hipMemcpyAsync(lp->st, lp->ahs, sizeof(State), hipMemcpyHostToDevice,
                     st) != hipSuccess ||
      hipMemsetAsync(lp->prof, 0, sizeof(Prof), st) != hipSuccess)
    return -1;
  k_ctl<<<1, 256, 0, st>>>(a, (uint32_t)lp->blocks);
  if (hipGraphLaunch(lp->graph[key], st) != hipSuccess ||
      hipMemcpyAsync(lp->ahs, lp->st, sizeof(State), hipMemcpyDeviceToHost,
                     st) != hipSuccess ||
      hipEventRecord(lp->aev, st) != hipSuccess)
    return -1;
  lp->alaunches = launches_per_op() * kPairs + 1;
  lp->abusy = 1;
  return 0;
}

int acd_dev_lp_async_wait(struct acd_dev_lp* lp) {
  if (!lp->abusy) return 0;
  return hipEventSynchronize(lp->aev) == hipSuccess ? 0 : -1;
}

int acd_dev_lp_async_poll(struct acd_dev_lp* lp, struct acd_solve_result* r) {
  if (!lp->abusy) return -1;
  const hipError_t q = hipEventQuery(lp->aev);
  if (q == hipErrorNotReady) return 0;
  if (q != hipSuccess) return -1;
  State* hs = lp->ahs;
  if (hs->pc == PC_DONE || hs->error) {
    memset(r, 0, sizeof *r);
    r->error = hs->error;
    r->ret = hs->alm_ret;
    for (int k = 0; k < 3; k++) r->diag[k] = hs->alm_diag[k];
    r->inner_iterations = hs->inner_total;
    r->outer_iterations = hs->outer;
    r->launches = lp->alaunches;
    r->kkt = hs->kkt;
    lp->abusy = 0;
    return 1;
  }
  hipStream_t st = lp->as;
  kick_if_idle(lp, hs, st, make_args(lp, lp->akey));
  for (int c = 0; c < lp->achunk; c++)
    if (hipGraphLaunch(lp->graph[lp->akey], st) != hipSuccess) return -1;
  lp->alaunches += (uint64_t)lp->achunk * launches_per_op() * kPairs;
  if (lp->achunk < 16) lp->achunk *= 2;
  if (hipMemcpyAsync(hs, lp->st, sizeof(State), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipEventRecord(lp->aev, st) != hipSuccess)
    return -1;
  return 0;
}

}  // extern "C"
