// solver_ops.hpp -- the LP hot path of pkhuong/acd-alm as micro-ops.
//
// ALM outer loop (alm.c:247-322), APPROX inner loop (approx/approx.c:
// 417-708) and spectral projected gradient (approx/mpg.inc:155-199,
// approx/approx.c:517-573), on the row/column-permuted working copy of the
// problem (approx/approx.c:84-88) stored as sliced ELL.
//
// The solver is split into
//   * ops: one data-parallel pass each (an SpMV over A or A' fused with its
//     elementwise epilogue, or an elementwise pass), producing up to eight
//     partial sums/maxima per thread; and
//   * a control state machine (Ctl) that consumes the reduced results of
//     the last op, takes every decision of the reference (step acceptance,
//     restarts, stopping tests, ALM updates) and names the next op.
// For a large LP each op is one wide kernel and Ctl runs in a one-block
// kernel after it (the kernel boundary is the global barrier, cheaper on
// MI355X than an in-kernel grid barrier); for a batch of small LPs one
// workgroup loops op -> workgroup reduction -> Ctl.  Ops visit index j
// with a fixed lane in a fixed order and reductions combine in a fixed
// order, so results are deterministic and identical for every launch.
//
// Pass structure of one APPROX iteration in the common (long step)
// case: PASS1 over A' rows -- g2 = A'(y.viol) + c, [g = A'(z.viol) + c],
// zp = clamp(z - g2 * itheta / v), model decrease, c'z, c'zp, fused with
// the previous iteration's x <- (1-t)x + t z -- then PASS2 over A rows --
// zp.viol = w (A zp - b) and its quadratic, plus the violation half of
// the momentum update written speculatively to shadow buffers that are
// adopted by swapping roles when the step is accepted.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "acd_dev.h"

namespace acd {

constexpr int kRes = 8;  // result slots per op
// Relative checkpoint change (over 100 iterations) below which FP32
// iterates count as stalled: 2^-20, about 8 ulps of FP32.
constexpr double kMixedMinDelta = 9.5367431640625e-07;

// n-side and m-side work vector roles.
enum { NX = 0, NZ, NZP, NG, NG2, NPV, kNumVN };
enum { MXV = 0, MYV, MZV, MZPV, MSX, MSY, kNumVM };

enum Op : int32_t {
  OP_NONE = 0,
  OP_INIT_X,
  OP_VIOL,
  OP_PASS1,
  OP_PASS2,
  OP_FIXUP,
  OP_RESTART_COPY,
  OP_CKPT_X,
  OP_GRAD_PG,
  OP_NORM_G,
  OP_FIN_A,
  OP_PERMUTE_OUT,
  OP_MPG_GRAD0,
  OP_MPG_GRAD,
  OP_MPG_ZUPD,
  OP_ALM_LAMPROJ,
  OP_PENALISE,
  OP_PERMUTE_PROBLEM,
  OP_STEP_SIZES,
  OP_ALM_VIOL,
  OP_ARGMAX,
  OP_LAM_W,
  OP_KKT,
  OP_PCD_STEP,
  OP_PCD_SIZES,
  OP_KKT_CK
};

enum OpFlag : uint32_t {
  F_APPROX = 1u << 0,  // INIT_X: approx variant (prev_x copy, c'x)
  F_COPY2 = 1u << 1,   // VIOL: copy the violation into z.viol and y.viol
  F_CX = 1u << 2,      // VIOL / PERMUTE_OUT: also c'x
  F_TWO = 1u << 3,     // PASS1: gradient at z too
  F_LONG = 1u << 4,    // PASS1: long step (model decrease)
  F_PG = 1u << 5,      // PASS1: projected gradient at z
  F_LAZY = 1u << 6,    // PASS1 / CKPT_X: apply pending x update
  F_PCOPY = 1u << 7,   // PASS1 / FIN_A: prev_x <- x pending
  F_SPEC = 1u << 8,    // PASS2: speculative (shadow) momentum update
  F_NEEDG = 1u << 9,   // FIXUP: compute g at z
  F_RESTEP = 1u << 10, // FIXUP: re-step from g2
  F_UPD = 1u << 11,    // LAM_W: scale the weights
  F_BUMP = 1u << 12    // LAM_W: weight[argmax] *= 10
};

// Control continuation points.
enum Pc : int32_t {
  PC_DONE = 0,
  // APPROX
  PC_A_INIT, PC_A_INIT1, PC_A_INIT2,
  PC_I0, PC_I_SHORT, PC_I_LONG1, PC_I_LONG2, PC_I_STEPPED, PC_I_FIX,
  PC_I_DD, PC_I_RESTARTED, PC_I_LIN2, PC_I_COMMIT, PC_I_DRIVER,
  PC_I_CK1, PC_I_CK2, PC_I_CK3, PC_I_CK4, PC_I_PRINT, PC_I_PRINT2, PC_I_NEXT,
  PC_F0, PC_F1, PC_F2, PC_F3,
  // SPG
  PC_M0, PC_M1, PC_M2, PC_M3, PC_MI, PC_MI2, PC_MI3, PC_MI5,
  PC_MF1, PC_MF2,
  // ALM
  PC_L0, PC_L1, PC_LB0, PC_LB1, PC_LB2, PC_LB3,
  PC_LE0, PC_LE1, PC_LE2, PC_LE3, PC_LE4, PC_LE5,
  // approx_update
  PC_U0, PC_U1, PC_U2,
  // parallel coordinate descent
  PC_P0, PC_PZ, PC_P1, PC_P2, PC_PI, PC_PS, PC_PV, PC_PV2, PC_PE, PC_PF,
  PC_PF2
};

// Whether coordinate j is in the sample keyed by `key` (a SplitMix64-style
// finaliser of key + j times the golden-ratio increment): the high 32 bits
// below thresh, i.e. with probability thresh / 2^32.
__device__ __forceinline__ bool pcd_pick(uint64_t key, uint32_t j,
                                         uint32_t thresh) {
  uint64_t h = key + 0x9E3779B97F4A7C15ull * (uint64_t)(j + 1);
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  return (uint32_t)(h >> 32) < thresh;
}

struct Sell {
  const uint64_t* off;
  const uint32_t* col;
  const double* val;       // FP64 values
  const float* val32;      // FP32 copy (mixed precision) or NULL
  const uint16_t* val16;   // BF16 copy when every value is exact, or NULL
  const uint32_t* colsgn;  // value-free entries of a 0/+-1 operator, or NULL
  const uint16_t* col16;   // columns as 16 bits (at most 65,536), or NULL
  uint32_t nrows, nslices;
};

// Matrix-value type of a value-free operator: every stored value is -1, 0
// or +1 and travels in the column word (bit 31: -1, bit 30: 0), so an
// entry costs 4 bytes instead of 6 (BF16) or 8 (FP32).
struct SgnT {};

// Matrix-value type of FP32 values with 16-bit column indices (operators
// of at most 65,536 columns): 6 bytes per entry instead of 8.
struct F32C16 {};

// Column of entry e as stored for precision MT.
template <class MT>
__device__ __forceinline__ uint32_t mcol(const Sell& M, uint64_t e) {
  return M.col[e];
}
template <>
__device__ __forceinline__ uint32_t mcol<SgnT>(const Sell& M, uint64_t e) {
  return M.colsgn[e] & 0x3fffffffu;
}
template <>
__device__ __forceinline__ uint32_t mcol<F32C16>(const Sell& M, uint64_t e) {
  return M.col16[e];
}

// Value of entry e as stored for precision MT (double, float, or bf16
// carried in a uint16_t), widened to double for the FP64 accumulation.
template <class MT>
__device__ __forceinline__ double mval(const Sell& M, uint64_t e);
template <>
__device__ __forceinline__ double mval<double>(const Sell& M, uint64_t e) {
  return M.val[e];
}
template <>
__device__ __forceinline__ double mval<float>(const Sell& M, uint64_t e) {
  return (double)M.val32[e];
}
template <>
__device__ __forceinline__ double mval<uint16_t>(const Sell& M, uint64_t e) {
  return (double)__uint_as_float((uint32_t)M.val16[e] << 16);
}
template <>
__device__ __forceinline__ double mval<F32C16>(const Sell& M, uint64_t e) {
  return (double)M.val32[e];
}
template <>
__device__ __forceinline__ double mval<SgnT>(const Sell& M, uint64_t e) {
  const uint32_t w = M.colsgn[e];
  return (w & 0x40000000u) ? 0.0 : ((w & 0x80000000u) ? -1.0 : 1.0);
}

// Element e of a matrix stream, read as global memory (the streams never
// live in LDS), so the load counts only against the vector-memory counter
// and a wait for it does not also wait for LDS reads.
template <class T>
__device__ __forceinline__ T gload(const T* p, uint64_t e) {
  return ((const __attribute__((address_space(1))) T*)p)[e];
}
// A pointer that is the same in every lane (read from a record all lanes
// share), moved to scalar registers.
template <class T>
__device__ __forceinline__ const T* wave_ptr(const T* p) {
  const uint64_t v = (uint64_t)p;
  const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint64_t hi =
      (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const T*)((hi << 32) | lo);
}

// Base pointers of the entry streams that precision MT reads, kept in
// scalar registers for a whole product (the empty asm statement makes
// them opaque, so the compiler cannot re-read them from the LDS-resident
// argument record between loads).
template <class MT>
struct EntBase {
  const uint32_t* col;
  const MT* val;
  __device__ __forceinline__ explicit EntBase(const Sell& M);
};
template <>
__device__ __forceinline__ EntBase<double>::EntBase(const Sell& M)
    : col(wave_ptr(M.col)), val(wave_ptr(M.val)) {
  asm volatile("" : "+s"(col), "+s"(val));
}
template <>
__device__ __forceinline__ EntBase<float>::EntBase(const Sell& M)
    : col(wave_ptr(M.col)), val(wave_ptr(M.val32)) {
  asm volatile("" : "+s"(col), "+s"(val));
}
template <>
__device__ __forceinline__ EntBase<uint16_t>::EntBase(const Sell& M)
    : col(wave_ptr(M.col)), val(wave_ptr(M.val16)) {
  asm volatile("" : "+s"(col), "+s"(val));
}
template <>
struct EntBase<SgnT> {
  const uint32_t* w;
  __device__ __forceinline__ explicit EntBase(const Sell& M)
      : w(wave_ptr(M.colsgn)) {
    asm volatile("" : "+s"(w));
  }
};
template <>
struct EntBase<F32C16> {
  const uint16_t* col;
  const float* val;
  __device__ __forceinline__ explicit EntBase(const Sell& M)
      : col(wave_ptr(M.col16)), val(wave_ptr(M.val32)) {
    asm volatile("" : "+s"(col), "+s"(val));
  }
};

// One stored entry as loaded, decoded only where it is used.  The
// batched row loop issues the next chunk's entry loads while the current
// chunk's gathers are in flight and waits for both at one point (pin()); a
// conversion at the load (FP32 or BF16 to FP64, the value-free word's bit
// tests) would make the compiler wait for the loads right there, before
// the gathers are issued, serialising the two memory round trips.
template <class MT>
struct Ent {
  uint32_t c;
  MT v;
  __device__ __forceinline__ void load(const EntBase<MT>& B, uint64_t e) {
    c = gload(B.col, e);
    v = gload(B.val, e);
  }
  __device__ __forceinline__ uint32_t col() const { return c; }
  __device__ __forceinline__ double val() const { return (double)v; }
  __device__ __forceinline__ void pin() { asm volatile("" : "+v"(c), "+v"(v)); }
};
template <>
struct Ent<uint16_t> {  // BF16 bits, zero-extended
  uint32_t c, v;
  __device__ __forceinline__ void load(const EntBase<uint16_t>& B,
                                       uint64_t e) {
    c = gload(B.col, e);
    v = gload(B.val, e);
  }
  __device__ __forceinline__ uint32_t col() const { return c; }
  __device__ __forceinline__ double val() const {
    return (double)__uint_as_float(v << 16);
  }
  __device__ __forceinline__ void pin() { asm volatile("" : "+v"(c), "+v"(v)); }
};
template <>
struct Ent<SgnT> {
  uint32_t w;
  __device__ __forceinline__ void load(const EntBase<SgnT>& B, uint64_t e) {
    w = gload(B.w, e);
  }
  __device__ __forceinline__ uint32_t col() const { return w & 0x3fffffffu; }
  __device__ __forceinline__ double val() const {
    return (w & 0x40000000u) ? 0.0 : ((w & 0x80000000u) ? -1.0 : 1.0);
  }
  __device__ __forceinline__ void pin() { asm volatile("" : "+v"(w)); }
};
template <>
struct Ent<F32C16> {
  uint32_t c;  // zero-extended 16-bit column
  float v;
  __device__ __forceinline__ void load(const EntBase<F32C16>& B, uint64_t e) {
    c = gload(B.col, e);
    v = gload(B.val, e);
  }
  __device__ __forceinline__ uint32_t col() const { return c; }
  __device__ __forceinline__ double val() const { return (double)v; }
  __device__ __forceinline__ void pin() { asm volatile("" : "+v"(c), "+v"(v)); }
};

struct State {
  // control
  int32_t pc, mode, log, error;
  int32_t ret, reason, center_x, restart_flag;
  int32_t lazy_x, prevx_pending, violp_zp, events_full;
  // current op
  int32_t op, op_r0, op_r1, pad0;
  uint32_t op_flags, op_maxmask, op_u, pad1;
  double op_a, op_b, op_c;
  double* pn[kNumVN];
  double* pm[kNumVM];
  // APPROX
  uint64_t niter, period, i, k;
  double max_pg, max_value, min_delta, offset;
  double theta, step_length, lazy_theta;
  double quad_x, quad_z, quad_zp;
  double value, pg, delta;
  double diag[5];
  // values carried between continuations of one iteration
  double t_expected, t_cz, t_czp, t_dd, t_pg2, t_len, t_x0, t_x1, t_x2;
  int32_t t_descent, t_got, t_restep, t_have_dd, t_restarted, t_safe;
  // ALM
  uint64_t outer, outer_niter;
  double precision, prev_norm, max_viol, alm_pg, cx, norm, w1_pre;
  double stop_violation, stop_pg, precision_floor, inner_min_delta;
  int32_t mixed;  // FP32 iterates: inner stall test at the FP32 floor
  double stop_kkt;  // > 0: also stop when the relative KKT residual <= it
  double kkt;       // last relative KKT residual (NaN until computed)
  double kkt_cc, kkt_bb;  // sum of c_j^2 and of b_i^2 (set at ALM start)
  double alm_diag[3];
  int32_t alm_ret, inner_reason, upd, bump;
  double scale, argmax;
  // parallel coordinate descent: sampling threshold (of 2^32), steps per
  // epoch, seed, sampling probability
  uint32_t pcd_thresh, pcd_steps;
  uint64_t pcd_seed;
  double pcd_p;
  // bookkeeping
  uint64_t inner_total, events, ops;
  uint64_t t_last;              // wall clock when the last op was issued
  uint64_t t_step;              // wall clock at the step kernel's start
};
static_assert(sizeof(State) % 8 == 0, "State is copied as 64-bit words");

// Per-op counters of one solve, kept beside the state (four arrays of 32,
// updated as 64-bit words by consecutive lanes of the control kernel).
struct Prof {
  uint64_t op_count[32];        // launches per op
  uint64_t op_ticks[32];        // ticks from the op's issue to its control
  uint64_t step_ticks[32];      // ticks from step kernel start to control
  uint64_t op_bytes[32];        // algorithmic bytes moved per op
  uint64_t ctl_phase[4];        // control kernel: loads, combine, decide
};

struct Args {
  uint32_t m, n;
  Sell A, AT, A0, AT0;          // working copy and caller's matrix
  const uint32_t* row_idx;      // working copy permutations
  const uint32_t* col_idx;
  const uint32_t* col_ridx;
  const uint32_t* beta;         // entries per row of the working copy
  // caller-order vectors
  double *rhs, *weight, *linc, *lower, *upper, *alm_c, *lam_lo, *lam_hi;
  double *x_io, *lam, *r;
  // working-copy vectors
  double *p_rhs, *p_w, *p_c, *p_l, *p_u, *p_v, *p_iv;
  State* st;
  Prof* prof;
  double* partials;             // [blocks][kRes]
  acd_event* ev;
  uint32_t ev_cap;              // 0: events off (batched solves)
  uint64_t nnz;                 // nonzeros of A
  uint32_t si, sc, sv;          // bytes per iterate / constant / value
  uint32_t resident;            // batched kernel: vector pairs held in LDS
  uint32_t vwaves;              // batched kernel: the LP's virtual waves
  uint32_t sx;                  // bytes per stored column index
};
static_assert(sizeof(Args) % 8 == 0, "Args is copied as 64-bit words");

// SSE2 maxpd(a, b) / minpd(a, b) of the reference's projections
// (approx/step_project.inc:18-19, :53-54, approx/approx.c:272-273).
__device__ __forceinline__ double maxpd(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double minpd(double a, double b) { return a < b ? a : b; }

// Max-combines propagate NaN (the reference's maxpd chains do), so a
// non-finite iterate reaches the unboundedness test instead of being
// dropped by fmax and leaving the solve to spin on NaN comparisons.
__device__ __forceinline__ double combine(double a, double b, bool is_max) {
  if (!is_max) return a + b;
  if (a != a || b != b) return a + b;
  return fmax(a, b);
}

__device__ __forceinline__ double wave_reduce(double v, bool is_max) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = combine(v, __shfl_xor(v, o, 64), is_max);
  return v;
}

// ---------------------------------------------------------------------------
// Ops.  G supplies lane, gwave (this wave's index among the group's waves)
// and nwaves; slice s of 64 rows goes to wave s mod nwaves.  VI is the
// type of the iterates (x, z, gradients, violation vectors), VC that of the
// working copy's per-column constants (cost, bounds, step sizes) and MT that
// of the stored matrix values (double, float or bf16 bits).  Row-side
// constants (rhs, weights), multipliers and the caller's vectors stay in
// double, and every product is formed and accumulated in double.  The ALM
// residual and penalty passes read the caller's matrix in MT as well, so
// that the outer loop drives the residual of the same (rounded) model the
// inner solver minimises; a mismatch there would be absorbed into the
// multipliers scaled by the row weights.
template <class G, class VI = double, class VC = double, class MT = double>
struct Ops {
  const G& g;
  const Args& a;
  const State& s;
  double p[kRes];

  __device__ __forceinline__ Ops(const G& g_, const Args& a_, const State& s_)
      : g(g_), a(a_), s(s_) {
#pragma unroll
    for (int k = 0; k < kRes; k++) p[k] = 0.0;
  }

  __device__ __forceinline__ VI* VN(int r) const { return (VI*)s.pn[r]; }
  __device__ __forceinline__ VI* VM(int r) const { return (VI*)s.pm[r]; }
  __device__ __forceinline__ VC* P(double* q) const { return (VC*)q; }

  // Row r of M summed lane-sequentially in ascending column order from
  // +0.0 with separate multiply and add (the reference's per-row order,
  // spmv/spmv_block.c:160-254); epi(row, acc0, acc1).  MT2/X are the value
  // and gathered-vector types (the ALM passes use double for both).
  struct NoPre {};
  template <int NRHS, class MT2, class X, class E>
  __device__ __forceinline__ void spmv_rows_t(const Sell& M, const X* x0,
                                              const X* x1, E&& epi) {
    spmv_rows_pt<NRHS, MT2, X>(
        M, x0, x1, [](uint32_t) { return NoPre{}; },
        [&](uint32_t r, double a0, double a1, NoPre) { epi(r, a0, a1); });
  }
  // Same, with pre(row) called at the start of each slice so the row's
  // epilogue operands are loaded while its products are gathered; epi
  // receives what pre returned.
  template <int NRHS, class MT2, class X, class PreF, class E>
  __device__ __forceinline__ void spmv_rows_pt(const Sell& M, const X* x0,
                                               const X* x1, PreF&& pre,
                                               E&& epi) {
    // Round k gives wave w slice k*nwaves + (w + k) mod nwaves.  Slices
    // in a window of 256 rows are ordered by row length (the reference's
    // permutation), so a plain stride of nwaves (a multiple of 4) would
    // give every wave the same position in each window -- some waves all
    // the longest slices; the rotation spreads them.
    uint32_t rk = g.gwave, bk = 0;
    uint32_t sl = rk;
    if (sl >= M.nslices) return;
    uint64_t base = M.off[sl], end = M.off[sl + 1];
    for (;;) {
      // The next slice's bounds and each next chunk of 4 columns/values
      // are loaded while the current chunk's gathers are in flight; the
      // accumulation order is unchanged.
      bk += g.nwaves;
      rk = rk + 1 == g.nwaves ? 0u : rk + 1;
      const uint32_t nsl = bk + rk;
      uint64_t nbase = 0, nend = 0;
      if (nsl < M.nslices) {
        nbase = M.off[nsl];
        nend = M.off[nsl + 1];
      }
      const uint32_t len = (uint32_t)((end - base) >> 6);
      const uint32_t r = sl * 64 + g.lane;
      const bool live = r < M.nrows;
      const auto pv = pre(live ? r : 0u);
      uint64_t e = base + g.lane;
      double a0 = 0.0, a1 = 0.0;
      uint32_t k = 0;
      if (len >= 4 && !G::kPipelinedRows) {
        // Per chunk of 4 entries: the next chunk's entries are loaded
        // (and waited for), then the gathers of its products are issued,
        // then the adds in entry order.
        uint32_t c0 = mcol<MT2>(M, e), c1 = mcol<MT2>(M, e + 64),
                 c2 = mcol<MT2>(M, e + 128), c3 = mcol<MT2>(M, e + 192);
        double v0 = mval<MT2>(M, e), v1 = mval<MT2>(M, e + 64),
               v2 = mval<MT2>(M, e + 128), v3 = mval<MT2>(M, e + 192);
        for (;;) {
          const bool more = k + 8 <= len;
          uint32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0;
          double w0 = 0, w1 = 0, w2 = 0, w3 = 0;
          if (more) {
            const uint64_t en = e + 256;
            d0 = mcol<MT2>(M, en);
            d1 = mcol<MT2>(M, en + 64);
            d2 = mcol<MT2>(M, en + 128);
            d3 = mcol<MT2>(M, en + 192);
            w0 = mval<MT2>(M, en);
            w1 = mval<MT2>(M, en + 64);
            w2 = mval<MT2>(M, en + 128);
            w3 = mval<MT2>(M, en + 192);
          }
          const double p0 = x0[c0], p1 = x0[c1], p2 = x0[c2], p3 = x0[c3];
          a0 = a0 + v0 * p0;
          a0 = a0 + v1 * p1;
          a0 = a0 + v2 * p2;
          a0 = a0 + v3 * p3;
          if (NRHS == 2) {
            const double q0 = x1[c0], q1 = x1[c1], q2 = x1[c2], q3 = x1[c3];
            a1 = a1 + v0 * q0;
            a1 = a1 + v1 * q1;
            a1 = a1 + v2 * q2;
            a1 = a1 + v3 * q3;
          }
          k += 4;
          e += 256;
          if (!more) break;
          c0 = d0; c1 = d1; c2 = d2; c3 = d3;
          v0 = w0; v1 = w1; v2 = w2; v3 = w3;
        }
      } else if (len >= 4 && len < 8) {
        // One full chunk and a tail of len - 4 entries (every A' row of an
        // LP with 5 nonzeros per column): all entries are loaded at once,
        // then all gathers are issued, then the adds in entry order.
        const EntBase<MT2> B(M);
        const uint32_t t = len - 4;
        Ent<MT2> cur[4], tl[3];
#pragma unroll
        for (int q = 0; q < 4; q++) cur[q].load(B, e + 64 * q);
#pragma unroll
        for (int q = 0; q < 3; q++) {
          tl[q] = {};
          if ((uint32_t)q < t) tl[q].load(B, e + 256 + 64 * q);
        }
        X g0[7], g1[7];
#pragma unroll
        for (int q = 0; q < 4; q++) g0[q] = x0[cur[q].col()];
#pragma unroll
        for (int q = 0; q < 3; q++)
          g0[4 + q] = (uint32_t)q < t ? x0[tl[q].col()] : X(0);
        if (NRHS == 2) {
#pragma unroll
          for (int q = 0; q < 4; q++) g1[q] = x1[cur[q].col()];
#pragma unroll
          for (int q = 0; q < 3; q++)
            g1[4 + q] = (uint32_t)q < t ? x1[tl[q].col()] : X(0);
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
          a0 = a0 + cur[q].val() * (double)g0[q];
#pragma unroll
        for (int q = 0; q < 3; q++)
          if ((uint32_t)q < t) a0 = a0 + tl[q].val() * (double)g0[4 + q];
        if (NRHS == 2) {
#pragma unroll
          for (int q = 0; q < 4; q++)
            a1 = a1 + cur[q].val() * (double)g1[q];
#pragma unroll
          for (int q = 0; q < 3; q++)
            if ((uint32_t)q < t) a1 = a1 + tl[q].val() * (double)g1[4 + q];
        }
        k = len;
      } else if (len >= 4) {
        // Per chunk of 4 entries: the gathers of its products are issued,
        // then the next chunk's entry loads (after the last full chunk:
        // the row's last len % 4 entries), then one wait for both (pin:
        // the gathered and the loaded registers pass through an empty asm
        // statement, so neither load can move past it and nothing is
        // computed from them before it), then the adds in entry order.
        const EntBase<MT2> B(M);
        Ent<MT2> cur[4], nxt[4];
#pragma unroll
        for (int q = 0; q < 4; q++) cur[q].load(B, e + 64 * q);
        for (;;) {
          const bool more = k + 8 <= len;
          const uint32_t nn = more ? 4u : len - k - 4;  // entries in nxt
          X g0[4], g1[4];
#pragma unroll
          for (int q = 0; q < 4; q++) g0[q] = x0[cur[q].col()];
          if (NRHS == 2) {
#pragma unroll
            for (int q = 0; q < 4; q++) g1[q] = x1[cur[q].col()];
          }
#pragma unroll
          for (int q = 0; q < 4; q++) {
            nxt[q] = {};
            if ((uint32_t)q < nn) nxt[q].load(B, e + 256 + 64 * q);
          }
#pragma unroll
          for (int q = 0; q < 4; q++) {
            asm volatile("" : "+v"(g0[q]));
            if (NRHS == 2) asm volatile("" : "+v"(g1[q]));
            nxt[q].pin();
          }
#pragma unroll
          for (int q = 0; q < 4; q++)
            a0 = a0 + cur[q].val() * (double)g0[q];
          if (NRHS == 2) {
#pragma unroll
            for (int q = 0; q < 4; q++)
              a1 = a1 + cur[q].val() * (double)g1[q];
          }
          k += 4;
          e += 256;
          if (!more) break;
#pragma unroll
          for (int q = 0; q < 4; q++) cur[q] = nxt[q];
        }
        // The last len % 4 entries, loaded with the last full chunk.
        const uint32_t tail = len - k;
        if (tail) {
          const bool h1 = tail > 1, h2 = tail > 2;
          const double p0 = (double)x0[nxt[0].col()];
          const double p1 = h1 ? (double)x0[nxt[1].col()] : 0.0;
          const double p2 = h2 ? (double)x0[nxt[2].col()] : 0.0;
          a0 = a0 + nxt[0].val() * p0;
          if (h1) a0 = a0 + nxt[1].val() * p1;
          if (h2) a0 = a0 + nxt[2].val() * p2;
          if (NRHS == 2) {
            const double q0 = (double)x1[nxt[0].col()];
            const double q1 = h1 ? (double)x1[nxt[1].col()] : 0.0;
            const double q2 = h2 ? (double)x1[nxt[2].col()] : 0.0;
            a1 = a1 + nxt[0].val() * q0;
            if (h1) a1 = a1 + nxt[1].val() * q1;
            if (h2) a1 = a1 + nxt[2].val() * q2;
          }
          k = len;
        }
      }
      // The last len % 4 entries as one batch: their loads and gathers
      // are issued together, the adds stay in entry order.
      const uint32_t rem = len - k;
      if (rem) {
        const bool h1 = rem > 1, h2 = rem > 2;
        const uint32_t c0 = mcol<MT2>(M, e);
        const uint32_t c1 = h1 ? mcol<MT2>(M, e + 64) : 0u;
        const uint32_t c2 = h2 ? mcol<MT2>(M, e + 128) : 0u;
        const double v0 = mval<MT2>(M, e);
        const double v1 = h1 ? mval<MT2>(M, e + 64) : 0.0;
        const double v2 = h2 ? mval<MT2>(M, e + 128) : 0.0;
        const double p0 = x0[c0];
        const double p1 = h1 ? (double)x0[c1] : 0.0;
        const double p2 = h2 ? (double)x0[c2] : 0.0;
        a0 = a0 + v0 * p0;
        if (h1) a0 = a0 + v1 * p1;
        if (h2) a0 = a0 + v2 * p2;
        if (NRHS == 2) {
          const double q0 = x1[c0];
          const double q1 = h1 ? (double)x1[c1] : 0.0;
          const double q2 = h2 ? (double)x1[c2] : 0.0;
          a1 = a1 + v0 * q0;
          if (h1) a1 = a1 + v1 * q1;
          if (h2) a1 = a1 + v2 * q2;
        }
      }
      if (live) epi(r, a0, a1, pv);
      if (nsl >= M.nslices) break;
      sl = nsl;
      base = nbase;
      end = nend;
    }
  }
  template <int NRHS, class PreF, class E>
  __device__ __forceinline__ void spmv_rows_pre(const Sell& M, const VI* x0,
                                                const VI* x1, PreF&& pre,
                                                E&& epi) {
    spmv_rows_pt<NRHS, MT, VI>(M, x0, x1, pre, epi);
  }
  // Working-copy operators (A, A') in the solver precision.  The ALM
  // passes over the caller's operators (A0, A0') read the same stored
  // precision so that the outer and inner problems agree.
  template <int NRHS, class E>
  __device__ __forceinline__ void spmv_rows(const Sell& M, const VI* x0,
                                            const VI* x1, E&& epi) {
    spmv_rows_t<NRHS, MT, VI>(M, x0, x1, epi);
  }
  template <class F>
  __device__ __forceinline__ void each(uint32_t n, F&& f) {
    const uint32_t ns = (n + 63) >> 6;
    for (uint32_t sl = g.gwave; sl < ns; sl += g.nwaves) {
      const uint32_t r = sl * 64 + g.lane;
      if (r < n) f(r);
    }
  }

  // Clamp to [l, u] as the reference does: maxpd(l, t) then minpd(u, t).
  __device__ __forceinline__ double clamp_lu(uint32_t j, double t) const {
    t = maxpd((double)P(a.p_l)[j], t);
    return minpd((double)P(a.p_u)[j], t);
  }

  template <bool SolveOnly = false>
  __device__ __forceinline__ void run() {
    const uint32_t fl = s.op_flags;
    const VC* pc = P(a.p_c);
    switch (s.op) {
      case OP_INIT_X: {  // approx/approx.c:591-599, :534-537
        VI* X = VN(NX);
        VI* Z = VN(NZ);
        VI* PV = VN(NPV);
        each(a.n, [&](uint32_t j) {
          double xj = a.x_io[a.col_idx[j]];
          const double lo = maxpd((double)P(a.p_l)[j], xj);
          xj = minpd(lo, (double)P(a.p_u)[j]);
          const VI xv = (VI)xj;
          X[j] = xv;
          Z[j] = xv;
          if (fl & F_APPROX) {
            PV[j] = xv;
            p[0] += (double)pc[j] * (double)xv;
          }
        });
        break;
      }
      case OP_VIOL: {  // approx/gradient_value.inc:15-81
        const VI* X = VN(s.op_r0);
        VI* Vv = VM(s.op_r1);
        VI* c1 = VM(MZV);
        VI* c2 = VM(MYV);
        const bool copy2 = fl & F_COPY2;
        const double* b = a.p_rhs;
        const double* w = a.p_w;
        spmv_rows<1>(a.A, X, nullptr, [&](uint32_t i, double ax, double) {
          const double v = ax - (double)b[i];
          const double sv = (double)w[i] * v;
          const VI svv = (VI)sv;
          Vv[i] = svv;
          if (copy2) {
            c1[i] = svv;
            c2[i] = svv;
          }
          p[0] += sv * v;
        });
        if (fl & F_CX)
          each(a.n, [&](uint32_t j) { p[1] += (double)pc[j] * (double)X[j]; });
        break;
      }
      case OP_PASS1: {  // step_project.inc:179-236, :320-375, :452-508
        VI* Zp = VN(NZP);
        const VI* Z = VN(NZ);
        VI* G1 = VN(NG);
        VI* X = VN(NX);
        VI* PV = VN(NPV);
        const VC* iv = P(a.p_iv);
        const VC* vv = P(a.p_v);
        const bool two = fl & F_TWO, lng = fl & F_LONG, want_pg = fl & F_PG;
        const bool lazy = fl & F_LAZY, pcopy = fl & F_PCOPY;
        const double itheta = s.op_a, theta2 = s.op_b;
        const double lt = s.op_c, ls = 1 - s.op_c;
        struct Pre1 {
          double c, z, iv, l, u, v;
        };
        auto pre = [&](uint32_t j) {
          Pre1 q;
          q.c = (double)pc[j];
          q.z = (double)Z[j];
          q.iv = (double)iv[j];
          q.l = (double)P(a.p_l)[j];
          q.u = (double)P(a.p_u)[j];
          q.v = lng ? (double)vv[j] : 0.0;
          return q;
        };
        auto epi = [&](uint32_t j, double gy, double gz, const Pre1& q) {
          const double cj = q.c;
          const double zj = q.z;
          if (pcopy || lazy) {
            const VI xj = X[j];
            if (pcopy) PV[j] = xj;
            if (lazy) X[j] = (VI)(ls * (double)xj + lt * zj);
          }
          // g2 = A'viol(y) + c is not stored: the rare re-step (FIXUP)
          // recomputes it bit for bit from the same viol(y).
          const double g2 = gy + cj;
          const double stp = itheta * q.iv;
          const VI tv = (VI)minpd(q.u, maxpd(q.l, zj - g2 * stp));
          const double t = (double)tv;
          Zp[j] = tv;
          p[4] = maxpd(p[4], fabs(t));
          p[2] += cj * zj;
          p[3] += cj * t;
          if (lng) {
            const double d = t - zj;
            p[0] += g2 * d;
            p[1] += d * d * theta2 * q.v;
          }
          if (two) {
            const double gg = gz + cj;
            G1[j] = (VI)gg;
            p[5] += gg * (t - zj);
            if (want_pg) {
              const double d = zj - minpd(q.u, maxpd(q.l, zj - gg));
              p[6] += d * d;
            }
          }
        };
        if (two)
          spmv_rows_pre<2>(a.AT, VM(MYV), VM(MZV), pre, epi);
        else
          spmv_rows_pre<1>(a.AT, VM(MYV), nullptr, pre, epi);
        break;
      }
      case OP_PASS2: {  // compute_violation(zp) + approx/vector.inc:249-295
        const VI* Zp = VN(NZP);
        VI* ZPV = VM(MZPV);
        VI* XV = VM(MXV);
        VI* YV = VM(MYV);
        VI* SX = VM(MSX);
        VI* SY = VM(MSY);
        const double* b = a.p_rhs;
        const double* w = a.p_w;
        const bool spec = fl & F_SPEC;
        const double th = s.op_a, thn = s.op_b;
        const double sc = 1 - th, nsc = 1 - thn;
        struct Pre2 {
          double b, w, xv;
        };
        auto pre = [&](uint32_t i) {
          Pre2 q;
          q.b = b[i];
          q.w = w[i];
          q.xv = (double)XV[i];
          return q;
        };
        spmv_rows_pre<1>(a.A, Zp, nullptr, pre, [&](uint32_t i, double ax, double, const Pre2& q) {
          const double v = ax - q.b;
          const VI svv = (VI)(q.w * v);
          const double sv = (double)svv;
          ZPV[i] = svv;
          p[0] += sv * v;
          const VI xv = (VI)(sc * q.xv + th * sv);
          const VI yv = (VI)(nsc * (double)xv + thn * sv);
          if (spec) {
            SX[i] = xv;
            SY[i] = yv;
          } else {
            XV[i] = xv;
            YV[i] = yv;
          }
        });
        break;
      }
      case OP_FIXUP: {  // approx/approx.c:383-389, :445-450
        VI* Zp = VN(NZP);
        const VI* Z = VN(NZ);
        VI* G1 = VN(NG);
        const VC* iv = P(a.p_iv);
        const bool needg = fl & F_NEEDG, restep = fl & F_RESTEP;
        const double itheta = s.op_a;
        // gy = A'viol(y): viol(y) is still PASS1's (PASS2 wrote only the
        // shadow copies), so g2 = gy + c is PASS1's value bit for bit.
        auto body = [&](uint32_t j, double gy, double gz) {
          const double zj = (double)Z[j];
          double gg;
          if (needg) {
            gg = gz + (double)pc[j];
            G1[j] = (VI)gg;
          } else {
            gg = (double)G1[j];
          }
          double t;
          if (restep) {
            const double g2 = (double)(VI)(gy + (double)pc[j]);
            const double stp = itheta * (double)iv[j];
            const VI tv = (VI)clamp_lu(j, zj - g2 * stp);
            Zp[j] = tv;
            t = (double)tv;
            p[1] = maxpd(p[1], fabs(t));
          } else {
            t = (double)Zp[j];
          }
          p[0] += gg * (t - zj);
        };
        if (needg && restep)
          spmv_rows<2>(a.AT, VM(MYV), VM(MZV),
                       [&](uint32_t j, double gy, double gz) { body(j, gy, gz); });
        else if (needg)
          spmv_rows<1>(a.AT, VM(MZV), nullptr,
                       [&](uint32_t j, double gz, double) { body(j, 0.0, gz); });
        else if (restep)
          spmv_rows<1>(a.AT, VM(MYV), nullptr,
                       [&](uint32_t j, double gy, double) { body(j, gy, 0.0); });
        else
          each(a.n, [&](uint32_t j) { body(j, 0.0, 0.0); });
        break;
      }
      case OP_RESTART_COPY: {  // approx/approx.c:452-453
        VI* X = VN(NX);
        const VI* Z = VN(NZ);
        VI* XV = VM(MXV);
        VI* YV = VM(MYV);
        const VI* ZV = VM(MZV);
        each(a.n, [&](uint32_t j) { X[j] = Z[j]; });
        each(a.m, [&](uint32_t r) {
          const VI v = ZV[r];
          XV[r] = v;
          YV[r] = v;
        });
        break;
      }
      case OP_CKPT_X: {  // approx/approx.c:641-651
        VI* X = VN(NX);
        const VI* Z = VN(NZ);
        const VI* PV = VN(NPV);
        const bool lazy = fl & F_LAZY;
        const double lt = s.op_c, ls = 1 - s.op_c;
        each(a.n, [&](uint32_t j) {
          double xj = (double)X[j];
          if (lazy) {
            const VI xv = (VI)(ls * xj + lt * (double)Z[j]);
            X[j] = xv;
            xj = (double)xv;
          }
          p[0] += (double)pc[j] * xj;
          const double d = (double)PV[j] - xj;
          p[1] += d * d;
          p[2] += xj * xj;
        });
        break;
      }
      case OP_GRAD_PG: {  // gradient_value.inc:158-181, approx.c:258-278
        const VI* Vv = VM(s.op_r1);
        const VI* X = VN(s.op_r0);
        VI* Gout = VN(NG);
        spmv_rows<1>(a.AT, Vv, nullptr, [&](uint32_t j, double gx, double) {
          const double gg = gx + (double)pc[j];
          Gout[j] = (VI)gg;
          const double xj = (double)X[j];
          const double d = xj - clamp_lu(j, xj - gg);
          p[0] += d * d;
          p[1] += gg * gg;
        });
        break;
      }
      case OP_NORM_G: {
        const VI* Gv = VN(NG);
        each(a.n, [&](uint32_t j) {
          const double gj = (double)Gv[j];
          p[0] += gj * gj;
        });
        break;
      }
      case OP_FIN_A: {  // approx/approx.c:681-683
        const VI* C = VN(s.op_r0);
        const VI* X = VN(NX);
        const VI* PV = VN(NPV);
        const bool pcopy = fl & F_PCOPY;
        each(a.n, [&](uint32_t j) {
          const double cj = (double)C[j];
          const double pv = pcopy ? (double)X[j] : (double)PV[j];
          const double d = pv - cj;
          p[0] += d * d;
          p[1] += cj * cj;
          p[2] += (double)pc[j] * cj;
        });
        break;
      }
      case OP_PERMUTE_OUT: {  // approx/approx.c:693-694
        const VI* C = VN(s.op_r0);
        each(a.n, [&](uint32_t i) { a.x_io[i] = (double)C[a.col_ridx[i]]; });
        if (fl & F_CX)
          each(a.n, [&](uint32_t j) { p[0] += (double)pc[j] * (double)C[j]; });
        break;
      }
      case OP_ALM_LAMPROJ: {  // alm.c:171-180, :314
        each(a.m, [&](uint32_t i) {
          a.lam[i] = fmax(a.lam_lo[i], fmin(a.lam[i], a.lam_hi[i]));
          p[2] += a.rhs[i] * a.rhs[i];
        });
        each(a.n, [&](uint32_t j) {
          p[0] += a.x_io[j] * a.alm_c[j];
          p[1] += a.alm_c[j] * a.alm_c[j];
        });
        break;
      }
      case OP_PENALISE: {  // alm.c:106-119
        spmv_rows_t<1, MT, double>(a.AT0, a.lam, nullptr, [&](uint32_t j, double atl, double) {
          a.linc[j] = a.alm_c[j] - atl;
        });
        each(a.m, [&](uint32_t i) { p[0] += a.lam[i] * a.rhs[i]; });
        break;
      }
      case OP_PERMUTE_PROBLEM: {  // approx/approx.c:192-211
        VI* wb = VM(MSX);
        double* prhs = a.p_rhs;
        double* pw = a.p_w;
        each(a.m, [&](uint32_t i) {
          const uint32_t src = a.row_idx[i];
          prhs[i] = a.rhs[src];
          const double w = a.weight[src];
          pw[i] = w;
          wb[i] = (VI)((double)w * (double)a.beta[i]);
        });
        VC* pcw = P(a.p_c);
        VC* pl = P(a.p_l);
        VC* pu = P(a.p_u);
        each(a.n, [&](uint32_t j) {
          const uint32_t src = a.col_idx[j];
          pcw[j] = (VC)a.linc[src];
          pl[j] = (VC)a.lower[src];
          pu[j] = (VC)a.upper[src];
        });
        break;
      }
      case OP_STEP_SIZES: {  // approx/approx.c:135-186
        const VI* wb = VM(MSX);
        VC* pv = P(a.p_v);
        VC* piv = P(a.p_iv);
        const Sell& M = a.AT;
        for (uint32_t sl = g.gwave; sl < M.nslices; sl += g.nwaves) {
          const uint64_t base = M.off[sl];
          const uint32_t len = (uint32_t)((M.off[sl + 1] - base) >> 6);
          double acc = 0.0;
          for (uint32_t k = 0; k < len; k++) {
            const uint64_t e = base + (uint64_t)k * 64 + g.lane;
            const double val = mval<MT>(M, e);
            acc = acc + ((double)wb[mcol<MT>(M, e)] * val) * val;
          }
          const uint32_t j = sl * 64 + g.lane;
          if (j < M.nrows) {
            pv[j] = (VC)acc;
            piv[j] = (VC)((acc == 0 && (double)pc[j] == 0) ? 1.0 : 1.0 / acc);
          }
        }
        break;
      }
      case OP_ALM_VIOL: {  // alm.c:121-132, :134-169
        p[1] = -HUGE_VAL;
        spmv_rows_t<1, MT, double>(a.A0, a.x_io, nullptr, [&](uint32_t i, double ax, double) {
          const double ri = ax - a.rhs[i];
          a.r[i] = ri;
          p[0] += ri * ri;
          p[1] = fmax(p[1], fabs(ri));
          p[2] += fabs(a.weight[i]);
        });
        each(a.n, [&](uint32_t j) { p[3] += a.x_io[j] * a.alm_c[j]; });
        break;
      }
      case OP_ARGMAX: {  // alm.c:155-166: first index of the maximum
        const double mv = s.op_a;
        p[0] = -HUGE_VAL;
        each(a.m, [&](uint32_t i) {
          if (fabs(a.r[i]) == mv) p[0] = fmax(p[0], -(double)i);
        });
        break;
      }
      case OP_LAM_W: {  // alm.c:182-212
        const bool upd = fl & F_UPD, bump = fl & F_BUMP;
        const double scale = s.op_a;
        const uint32_t am = s.op_u;
        each(a.m, [&](uint32_t i) {
          const double li = a.lam[i] - a.weight[i] * a.r[i];
          a.lam[i] = fmax(a.lam_lo[i], fmin(li, a.lam_hi[i]));
          if (upd) {
            double w = a.weight[i] * scale;
            if (bump && i == am) w *= 10;
            a.weight[i] = w;
          }
          p[0] += fabs(a.weight[i]);
        });
        break;
      }
      case OP_KKT: {  // relative KKT residual of (x, lambda); d = c - A'lambda
        spmv_rows_t<1, MT, double>(a.AT0, a.lam, nullptr, [&](uint32_t j, double atl, double) {
          const double c = a.alm_c[j];
          const double d = c - atl;
          const double x = a.x_io[j], lo = a.lower[j], up = a.upper[j];
          const double pr = fmin(fmax(x - d, lo), up);
          const double dp = fmax(d, 0.0), dn = fmax(-d, 0.0);
          const bool lo_inf = !isfinite(lo), up_inf = !isfinite(up);
          const double inf = (lo_inf ? dp : 0.0) + (up_inf ? dn : 0.0);
          p[0] += (x - pr) * (x - pr);
          p[1] += inf * inf;
          p[2] += (lo_inf ? 0.0 : lo * dp) - (up_inf ? 0.0 : up * dn);
          p[3] += c * c;
        });
        each(a.m, [&](uint32_t i) {
          p[4] += a.rhs[i] * a.lam[i];
          p[5] += a.rhs[i] * a.rhs[i];
        });
        break;
      }
      case OP_KKT_CK: {
        // Relative KKT residual of (x, lambda') at an inner checkpoint of
        // an ALM solve, lambda' = lambda - w (Ax - b) being the multipliers
        // the outer step would set from this x (alm.c:182-192).  The
        // checkpoint has just formed viol = w (Ax - b) and the gradient
        // g = A' viol + c - A' lambda = c - A' lambda', so d = g needs no
        // product; that holds while lambda' needs no projection onto the
        // multiplier bounds, which slot 6 reports (max |projection|).
        // Working-copy order; caller-order data through the index maps.
        const VI* X = VN(NX);
        const VI* Gv = VN(NG);
        each(a.n, [&](uint32_t j) {
          const uint32_t src = a.col_idx[j];
          const double x = (double)X[j], d = (double)Gv[j];
          const double lo = a.lower[src], up = a.upper[src];
          const double pr = fmin(fmax(x - d, lo), up);
          const double dp = fmax(d, 0.0), dn = fmax(-d, 0.0);
          const bool lo_inf = !isfinite(lo), up_inf = !isfinite(up);
          const double inf = (lo_inf ? dp : 0.0) + (up_inf ? dn : 0.0);
          p[0] += (x - pr) * (x - pr);
          p[1] += inf * inf;
          p[2] += (lo_inf ? 0.0 : lo * dp) - (up_inf ? 0.0 : up * dn);
          p[3] += a.alm_c[src] * x;
        });
        const VI* V = VM(MXV);
        p[6] = 0.0;
        each(a.m, [&](uint32_t i) {
          const uint32_t src = a.row_idx[i];
          const double v = (double)V[i];
          const double r = v / a.p_w[i];
          const double l1 = a.lam[src] - v;
          const double l2 = fmax(a.lam_lo[src], fmin(l1, a.lam_hi[src]));
          p[4] += a.p_rhs[i] * l2;
          p[5] += r * r;
          p[6] = fmax(p[6], fabs(l2 - l1));
        });
        break;
      }
      default:
        if constexpr (!SolveOnly) run_rest();
        break;
    }
  }

  // The ops of the SPG and parallel coordinate descent solvers, which the
  // batched kernel (ALM with the APPROX inner solver only) leaves out, so
  // that its code and the control continuation it calls (57 KB together)
  // stay within the 64 KB instruction cache that two compute units share.
  __device__ __forceinline__ void run_rest() {
    const uint32_t fl = s.op_flags;
    const VC* pc = P(a.p_c);
    switch (s.op) {
      case OP_MPG_GRAD0: {  // approx/approx.c:538-539
        VI* G1 = VN(NG);
        VI* G2 = VN(NG2);
        spmv_rows<1>(a.AT, VM(MXV), nullptr, [&](uint32_t j, double gx, double) {
          const VI gg = (VI)(gx + (double)pc[j]);
          G1[j] = gg;
          G2[j] = gg;
        });
        break;
      }
      case OP_MPG_GRAD: {  // approx/mpg.inc:161-174
        const VI* X = VN(NX);
        const VI* Z = VN(NZ);
        VI* G1 = VN(NG);
        const VI* G2 = VN(NG2);
        spmv_rows<1>(a.AT, VM(MXV), nullptr, [&](uint32_t j, double gx, double) {
          const VI ggv = (VI)(gx + (double)pc[j]);
          const double gg = (double)ggv;
          G1[j] = ggv;
          const double xj = (double)X[j];
          const double d = xj - clamp_lu(j, xj - gg);
          p[0] += d * d;
          const double yj = xj - (double)Z[j];
          const double uj = gg - (double)G2[j];
          p[1] += yj * yj;
          p[2] += yj * uj;
        });
        break;
      }
      case OP_MPG_ZUPD: {  // approx/mpg.inc:180-182
        const VI* X = VN(NX);
        VI* Z = VN(NZ);
        const VI* G1 = VN(NG);
        const double st = s.op_a;
        each(a.n, [&](uint32_t j) {
          const double t = (double)X[j] - st * (double)G1[j];
          const double lo = maxpd((double)P(a.p_l)[j], t);
          Z[j] = (VI)minpd(lo, (double)P(a.p_u)[j]);
        });
        break;
      }
      case OP_PCD_STEP: {  // parallel coordinate descent, one sampled step
        // Each coordinate j is in this step's sample with probability
        // thresh / 2^32, drawn from a hash of (seed, epoch, step, j), so the
        // sample does not depend on the launch shape.  Sampled coordinates
        // move by x_j <- clamp(x_j - g_j / v'_j) (v'_j from OP_PCD_SIZES,
        // its inverse kept in z) with g = A'(w (A x - b)) + c at the step's
        // start; unsampled rows of A' issue no loads.
        VI* X = VN(NX);
        const VI* ist = VN(NZ);
        const VI* sv = VM(MXV);
        const Sell& M = a.AT;
        const uint64_t key = s.pcd_seed ^ (s.i << 32) ^ s.k;
        for (uint32_t sl = g.gwave; sl < M.nslices; sl += g.nwaves) {
          const uint32_t j = sl * 64 + g.lane;
          if (j >= M.nrows || !pcd_pick(key, j, s.pcd_thresh)) continue;
          const uint64_t base = M.off[sl];
          const uint32_t len = (uint32_t)((M.off[sl + 1] - base) >> 6);
          double acc = 0.0;
          for (uint32_t k = 0; k < len; k++) {
            const uint64_t e = base + (uint64_t)k * 64 + g.lane;
            acc = acc + mval<MT>(M, e) * (double)sv[mcol<MT>(M, e)];
          }
          const double gj = acc + (double)pc[j];
          X[j] = (VI)clamp_lu(j, (double)X[j] - gj * (double)ist[j]);
        }
        break;
      }
      case OP_PCD_SIZES: {  // ESO curvatures of the sampled steps
        // v'_j = sum_i w_i a_ij^2 (1 + p (omega_i - 1)), omega_i the
        // nonzeros in row i: E over the sample S of the quadratic term at
        // x + h_S is at most p/2 sum_j v'_j h_j^2 (Cauchy-Schwarz over each
        // row's sampled entries).  Its inverse goes to z, which the solve
        // does not otherwise use; zero curvature steps like approx_update.
        VI* ist = VN(NZ);
        const double pr = s.pcd_p;
        const Sell& M = a.AT;
        for (uint32_t sl = g.gwave; sl < M.nslices; sl += g.nwaves) {
          const uint64_t base = M.off[sl];
          const uint32_t len = (uint32_t)((M.off[sl + 1] - base) >> 6);
          double acc = 0.0;
          for (uint32_t k = 0; k < len; k++) {
            const uint64_t e = base + (uint64_t)k * 64 + g.lane;
            const uint32_t i = mcol<MT>(M, e);
            const double val = mval<MT>(M, e);
            const double wi =
                (double)a.p_w[i] * (1.0 + pr * ((double)a.beta[i] - 1.0));
            acc = acc + (wi * val) * val;
          }
          const uint32_t j = sl * 64 + g.lane;
          if (j < M.nrows)
            ist[j] = (VI)((acc == 0 && (double)pc[j] == 0) ? 1.0 : 1.0 / acc);
        }
        break;
      }
      default:
        break;
    }
  }
};

// ---------------------------------------------------------------------------
// Algorithmic bytes of one op: each stored nonzero read once (value +
// 32-bit column), the slice offsets, the gathered vector counted once per
// element, and every streamed vector read or written once.  Padding and
// repeated gathers are not counted: they are the kernel's overhead, which
// the achieved fraction of the bandwidth roofline exposes.
__device__ inline double alg_bytes(int32_t op, uint32_t fl, const Args& a) {
  const double m = a.m, n = a.n, si = a.si, sc = a.sc;
  const double mat = (double)a.nnz * (a.sv + a.sx);
  const double offA = 8.0 * (a.A.nslices + 1), offT = 8.0 * (a.AT.nslices + 1);
  const double spA = mat + offA, spT = mat + offT;  // A rows / A' rows
  switch (op) {
    case OP_INIT_X:
      return n * (8 + 4 + 2 * sc + 2 * si) +
             ((fl & F_APPROX) ? n * (si + sc) : 0);
    case OP_VIOL:
      return spA + n * si + m * (16 + si) + ((fl & F_COPY2) ? 2 * m * si : 0) +
             ((fl & F_CX) ? n * (si + sc) : 0);
    case OP_PASS1: {
      const bool two = fl & F_TWO;
      double b = spT + m * si * (two ? 2 : 1);
      b += n * (4 * sc + 2 * si);  // c, iv, l, u; z, zp
      if (fl & F_LONG) b += n * sc;
      if (two) b += n * si;
      if (fl & (F_PCOPY | F_LAZY)) b += n * si;
      if (fl & F_PCOPY) b += n * si;
      if (fl & F_LAZY) b += n * si;
      return b;
    }
    case OP_PASS2:
      return spA + n * si + m * (16 + 4 * si);
    case OP_FIXUP: {
      const bool g = fl & F_NEEDG, r = fl & F_RESTEP;
      double b = n * si;  // z
      if (g || r) b += spT + m * si * ((g && r) ? 2 : 1) + n * sc;  // + c
      b += g ? n * si : n * si;  // g written, or g read
      b += r ? n * (3 * sc + si) : n * si;  // iv, l, u; zp written / read
      return b;
    }
    case OP_RESTART_COPY: return 2 * n * si + 3 * m * si;
    case OP_CKPT_X:
      return n * (2 * si + sc) + ((fl & F_LAZY) ? 2 * n * si : 0);
    case OP_GRAD_PG: return spT + m * si + n * (3 * sc + 2 * si);
    case OP_NORM_G: return n * si;
    case OP_FIN_A: return n * (2 * si + sc);
    case OP_PERMUTE_OUT:
      return n * (si + 4 + 8) + ((fl & F_CX) ? n * (si + sc) : 0);
    case OP_MPG_GRAD0: return spT + m * si + n * (sc + 2 * si);
    case OP_MPG_GRAD: return spT + m * si + n * (3 * sc + 4 * si);
    case OP_MPG_ZUPD: return n * (3 * si + 2 * sc);
    case OP_ALM_LAMPROJ: return 40 * m + 16 * n;
    case OP_PENALISE: return spT + 8 * m + 16 * n + 16 * m;
    case OP_PERMUTE_PROBLEM: return m * (40 + si) + n * (28 + 3 * sc);
    case OP_STEP_SIZES: return spT + m * si + 3 * n * sc;
    case OP_ALM_VIOL: return spA + 8 * n + 24 * m + 16 * n;
    case OP_ARGMAX: return 8 * m;
    case OP_LAM_W: return 48 * m + ((fl & F_UPD) ? 8 * m : 0);
    case OP_KKT: return spT + 8 * m + 32 * n + 16 * m;
    case OP_PCD_STEP: return spT + m * si + n * (2 * si + 4 * sc);
    case OP_PCD_SIZES: return spT + m * 12 + n * (si + sc);
    case OP_KKT_CK: return n * (2 * si + 4 + 24) + m * (si + 48);
    default: return 0;
  }
}

// ---------------------------------------------------------------------------
// Control: consumes the reduced results of the op just run, advances.
// The control step always works on the state record and the reduced
// results in LDS; saying so in the types lets an out-of-line continuation
// address them with LDS instructions instead of generic ones.
typedef __attribute__((address_space(3))) State LdsState;
typedef __attribute__((address_space(3))) double LdsDouble;

struct Ctl {
  LdsState& s;
  const Args& a;
  const LdsDouble* res;

  __device__ void emit(int32_t type, uint64_t k, double a0 = 0, double a1 = 0,
                       double a2 = 0, double a3 = 0, double a4 = 0,
                       double a5 = 0) {
    if (s.events < a.ev_cap) {
      acd_event* e = &a.ev[s.events];
      e->type = type;
      e->pad = 0;
      e->k = k;
      e->a[0] = a0; e->a[1] = a1; e->a[2] = a2;
      e->a[3] = a3; e->a[4] = a4; e->a[5] = a5;
      s.events++;
    }
  }

  __device__ void issue(int32_t op, uint32_t flags, int32_t next_pc,
                        uint32_t maxmask = 0, double oa = 0, double ob = 0,
                        double oc = 0) {
    s.op = op;
    s.op_flags = flags;
    s.op_maxmask = maxmask;
    s.op_a = oa;
    s.op_b = ob;
    s.op_c = oc;
    s.pc = next_pc;
    s.ops++;
  }

  __device__ void fail(int32_t code) {
    s.error = code;
    s.pc = PC_DONE;
    s.op = OP_NONE;
  }

  template <class T>
  __device__ static void swap_(T& x, T& y) {
    const auto t = +x;  // the value, without the LDS qualifier of x
    x = y;
    y = t;
  }

  __device__ void stop(int32_t reason) {
    s.reason = reason;
    s.pc = PC_F0;
  }

  // Runs continuations until one issues an op (returns) or the solve ends.
  // OutOfLine = true calls the continuation as a function of its own: the
  // caller's register allocation then does not have to hold the
  // continuation's temporaries beside its own (the batched kernel, whose
  // occupancy the register count sets, uses it; the control kernel
  // inlines it).
  // SolveOnly = true keeps only the APPROX, ALM and approx_update
  // continuations (any other program counter fails the solve).
  template <bool OutOfLine = false, bool SolveOnly = false>
  __device__ void advance() {
    s.op = OP_NONE;
    for (int guard = 0; guard < 256 && s.op == OP_NONE && s.pc != PC_DONE;
         guard++) {
      if (OutOfLine)
        step_call<SolveOnly>();
      else
        step<SolveOnly>();
    }
  }

  template <bool SolveOnly>
  __device__ __attribute__((noinline)) void step_call() {
    step<SolveOnly>();
  }

  template <bool SolveOnly = false>
  __device__ void step() {
    switch (s.pc) {
      // ------------------------------------------------------------ APPROX
      case PC_A_INIT:  // approx/approx.c:588-607
        issue(OP_INIT_X, F_APPROX, PC_A_INIT1);
        break;
      case PC_A_INIT1:
        s.cx = res[0];
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_VIOL, F_COPY2, PC_A_INIT2);
        break;
      case PC_A_INIT2: {
        const double q = .5 * res[0];
        s.quad_x = s.quad_z = q;
        s.violp_zp = 0;
        s.value = s.cx + q + s.offset;
        s.i = 0;
        s.k = 0;
        s.theta = 1;
        s.step_length = 1;
        s.lazy_x = 0;
        s.prevx_pending = 0;
        s.center_x = 1;
        s.reason = 0;
        s.restart_flag = 0;
        s.pg = HUGE_VAL;
        s.delta = HUGE_VAL;
        emit(ACD_EV_APPROX_BEGIN, 0);
        s.pc = s.niter > 0 ? PC_I0 : PC_F0;
        break;
      }
      case PC_I0: {  // approx/approx.c:608-616, :417-437
        if (s.i >= s.niter) {
          s.pc = PC_F0;
          break;
        }
        if (a.ev_cap && s.events + 8 > a.ev_cap) {  // let the host drain the log
          s.events_full = 1;
          s.op = OP_NONE;
          s.pc = PC_I0;
          return;
        }
        const bool pg_req = (s.i == 0) || ((s.i + 1) % 10 == 0);
        const double len = s.step_length;
        uint32_t fl = (pg_req ? F_PG : 0) | (s.lazy_x ? F_LAZY : 0) |
                      (s.prevx_pending ? F_PCOPY : 0);
        s.t_len = len;
        if (len <= 1) {
          issue(OP_PASS1, fl | F_TWO, PC_I_SHORT, 1u << 4,
                (1 - 1e-6) / s.theta, 0.0, s.lazy_theta);
        } else {
          const bool got = pg_req || (len <= 1 + 1e-6);
          const bool safe = len <= 1 + 1e-6;
          const double L = safe ? 1.0 : len;
          s.t_got = got;
          s.t_safe = safe;
          issue(OP_PASS1, fl | F_LONG | (got ? F_TWO : 0), PC_I_LONG1, 1u << 4,
                (L - 1e-6) / s.theta, s.theta / L, s.lazy_theta);
        }
        s.lazy_x = 0;
        s.prevx_pending = 0;
        break;
      }
      case PC_I_SHORT: {  // approx/approx.c:394-415
        if (!(res[4] < HUGE_VAL)) return fail(ACD_ERR_UNBOUNDED);
        s.step_length = s.t_len * 1.01;
        if (s.step_length > 1) s.step_length = 1 + 1e-6;
        s.t_descent = 0;
        s.t_got = 1;
        s.t_restep = 0;
        s.t_dd = res[5];
        s.t_have_dd = 1;
        s.violp_zp = 0;
        s.t_cz = res[2];
        s.t_pg2 = res[6];
        s.pc = PC_I_STEPPED;
        break;
      }
      case PC_I_LONG1: {  // approx/approx.c:343-376
        if (!(res[4] < HUGE_VAL)) return fail(ACD_ERR_UNBOUNDED);
        s.t_expected = res[0] + .5 * res[1];
        s.t_cz = res[2];
        s.t_czp = res[3];
        s.t_dd = res[5];
        s.t_pg2 = res[6];
        issue(OP_PASS2, F_SPEC, PC_I_LONG2, 0, s.theta,
              2.0 / (double)(s.k + 1 + 2));
        break;
      }
      case PC_I_LONG2: {  // approx/approx.c:374-391
        const double qzp = .5 * res[0];
        s.quad_zp = qzp;
        s.violp_zp = 1;
        const double L = s.t_safe ? 1.0 : s.t_len;
        const double initial = s.t_cz + s.quad_z;
        const double now = s.t_czp + qzp;
        if (!(s.t_expected <= 0)) return fail(ACD_ERR_ESTIMATE);
        s.t_restep = 0;
        if (now <= initial + s.t_expected) {
          s.step_length = L * 1.01;
          s.t_descent = 1;
        } else {
          s.step_length = .9 * L;
          s.t_descent = 0;
          s.t_restep = !s.t_safe;
        }
        s.t_have_dd = s.t_got && !s.t_restep;
        s.pc = PC_I_STEPPED;
        break;
      }
      case PC_I_STEPPED: {  // approx/approx.c:439-450
        const bool pg_req = (s.i == 0) || ((s.i + 1) % 10 == 0);
        s.pg = pg_req ? sqrt(s.t_pg2) : HUGE_VAL;
        s.delta = HUGE_VAL;
        s.t_restarted = 0;
        if (!s.t_descent && !s.t_have_dd) {
          issue(OP_FIXUP, (s.t_got ? 0 : F_NEEDG) | (s.t_restep ? F_RESTEP : 0),
                PC_I_FIX, 2u, (1 - 1e-6) / s.theta);
        } else {
          s.pc = PC_I_DD;
        }
        break;
      }
      case PC_I_FIX:
        s.t_dd = res[0];
        if (s.t_restep) {
          if (!(res[1] < HUGE_VAL)) return fail(ACD_ERR_UNBOUNDED);
          s.violp_zp = 0;
        }
        s.pc = PC_I_DD;
        break;
      case PC_I_DD:
        if (!s.t_descent && s.t_dd > 0) {  // oscillation: restart
          issue(OP_RESTART_COPY, 0, PC_I_RESTARTED);
        } else if (!s.violp_zp) {
          issue(OP_PASS2, 0, PC_I_LIN2, 0, s.theta,
                2.0 / (double)(s.k + 1 + 2));
        } else {
          swap_(s.pm[MXV], s.pm[MSX]);
          swap_(s.pm[MYV], s.pm[MSY]);
          s.pc = PC_I_COMMIT;
        }
        break;
      case PC_I_RESTARTED:  // approx/approx.c:452-457
        s.quad_x = s.quad_z;
        s.lazy_x = 0;
        s.k = 0;
        s.theta = 1;
        s.center_x = 1;
        s.t_restarted = 1;
        s.pc = PC_I_DRIVER;
        break;
      case PC_I_LIN2:
        s.quad_zp = .5 * res[0];
        s.violp_zp = 1;
        s.pc = PC_I_COMMIT;
        break;
      case PC_I_COMMIT: {  // approx/approx.c:461-477
        const double th = s.theta;
        s.k += 1;
        s.theta = 2.0 / (double)(s.k + 2);
        s.lazy_x = 1;
        s.lazy_theta = th;
        swap_(s.pn[NZ], s.pn[NZP]);
        swap_(s.pm[MZV], s.pm[MZPV]);
        swap_(s.quad_z, s.quad_zp);
        s.center_x = 0;
        s.pc = PC_I_DRIVER;
        break;
      }
      case PC_I_DRIVER: {  // approx/approx.c:617-640
        s.inner_total++;
        if (s.t_restarted) {
          if (s.log) emit(ACD_EV_RESTART, s.i);
          s.restart_flag = 1;
        }
        const double value =
            s.t_cz + (s.center_x ? s.quad_x : s.quad_zp) + s.offset;
        s.value = value;
        if (value < s.max_value) return stop(1);
        if (s.pg < s.max_pg) return stop(2);
        if ((s.i + 1) % 100 == 0) {
          issue(OP_CKPT_X, s.lazy_x ? F_LAZY : 0, PC_I_CK1, 0, 0, 0,
                s.lazy_theta);
        } else {
          s.pc = PC_I_PRINT;
        }
        break;
      }
      case PC_I_CK1:  // approx/approx.c:641-651
        s.t_x0 = res[0];
        s.t_x1 = res[1];
        s.t_x2 = res[2];
        s.lazy_x = 0;
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_VIOL, 0, PC_I_CK2);
        break;
      case PC_I_CK2:
        s.quad_x = .5 * res[0];
        s.value = s.t_x0 + s.quad_x + s.offset;
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_GRAD_PG, 0, PC_I_CK3);
        break;
      case PC_I_CK3:  // approx/approx.c:647-665
        s.pg = sqrt(res[0]);
        s.delta = sqrt(s.t_x1) / (sqrt(s.t_x2) + 1e-10);
        s.center_x = 1;
        if (s.value < s.max_value) return stop(1);
        if (s.pg < s.max_pg) return stop(2);
        if (s.delta < s.min_delta) return stop(3);
        if (s.mode == ACD_MODE_ALM && s.stop_kkt > 0) {
          issue(OP_KKT_CK, 0, PC_I_CK4, 1u << 6);
          break;
        }
        s.prevx_pending = 1;
        s.pc = PC_I_PRINT;
        break;
      case PC_I_CK4: {
        // The ALM's KKT target is met at this x already: end the inner
        // solve here, as if converged; the outer step then sets lambda'
        // and confirms the residual with its own exact evaluation.
        const double rp = sqrt(res[5]) / (1 + sqrt(s.kkt_bb));
        const double rd = (sqrt(res[0]) + sqrt(res[1])) / (1 + sqrt(s.kkt_cc));
        const double dual = res[4] + res[2];
        const double rg = fabs(res[3] - dual) / (1 + fabs(res[3]) + fabs(dual));
        const double k = fmax(rp, fmax(rd, rg));
        if (res[6] == 0.0 && k <= s.stop_kkt) return stop(2);
        s.prevx_pending = 1;
        s.pc = PC_I_PRINT;
        break;
      }
      case PC_I_PRINT:  // approx/approx.c:666-674
        if (s.i == 0 || (s.period && (s.i + 1) % s.period == 0)) {
          if (s.restart_flag) {
            emit(ACD_EV_NEWLINE, s.i);
            s.restart_flag = 0;
          }
          if (s.log) {
            issue(OP_NORM_G, 0, PC_I_PRINT2);
            break;
          }
        }
        s.pc = PC_I_NEXT;
        break;
      case PC_I_PRINT2:
        emit(ACD_EV_PRINT, s.i + 1, s.value, sqrt(res[0]), s.pg, s.step_length,
             s.delta);
        s.pc = PC_I_NEXT;
        break;
      case PC_I_NEXT:
        s.i += 1;
        s.pc = PC_I0;
        break;
      case PC_F0:  // approx/approx.c:676-683
        s.op_r0 = s.center_x ? NX : NZP;
        issue(OP_FIN_A, s.prevx_pending ? F_PCOPY : 0, PC_F1);
        break;
      case PC_F1:
        s.delta = sqrt(res[0]) / (sqrt(res[1]) + 1e-10);
        s.value = res[2] + (s.center_x ? s.quad_x : s.quad_zp) + s.offset;
        s.op_r0 = s.center_x ? NX : NZP;
        s.op_r1 = s.center_x ? MXV : MZPV;
        issue(OP_GRAD_PG, 0, PC_F2);
        break;
      case PC_F2:
        s.diag[1] = sqrt(res[1]);
        s.diag[2] = sqrt(res[0]);
        s.op_r0 = s.center_x ? NX : NZP;
        issue(OP_PERMUTE_OUT, 0, PC_F3);
        break;
      case PC_F3:  // approx/approx.c:684-702
        if (s.restart_flag) {
          emit(ACD_EV_NEWLINE, s.i);
          s.restart_flag = 0;
        }
        s.diag[0] = s.value;
        s.diag[3] = s.delta;
        s.diag[4] = (double)(s.i + 1);
        if (s.log)
          emit(ACD_EV_PRINT, s.i + 1, s.value, s.diag[1], s.diag[2],
               s.step_length, s.delta);
        s.ret = s.reason;
        s.pc = (s.mode == ACD_MODE_ALM) ? PC_LE0 : PC_DONE;
        break;
      // ------------------------------------------------------------ ALM
      case PC_L0:  // alm.c:289-300
        issue(OP_ALM_LAMPROJ, 0, PC_L1);
        break;
      case PC_L1:
        s.cx = res[0];
        s.kkt_cc = res[1];
        s.kkt_bb = res[2];
        s.outer = 0;
        s.precision = 1;
        s.prev_norm = HUGE_VAL;
        s.max_viol = HUGE_VAL;
        s.alm_pg = HUGE_VAL;
        s.alm_ret = 1;
        s.alm_diag[0] = s.cx;
        s.alm_diag[1] = s.max_viol;
        s.alm_diag[2] = s.alm_pg;
        s.pc = s.outer_niter ? PC_LB0 : PC_DONE;
        break;
      case PC_LB0:  // alm.c:247-257
        issue(OP_PENALISE, 0, PC_LB1);
        break;
      case PC_LB1:
        s.offset = res[0];
        issue(OP_PERMUTE_PROBLEM, 0, PC_LB2);
        break;
      case PC_LB2:
        issue(OP_STEP_SIZES, 0, PC_LB3);
        break;
      case PC_LB3:
        s.niter = 4294967295ull;  // -1u (alm.c:255)
        s.max_pg = fmin(1.0, s.precision);
        s.max_value = -HUGE_VAL;
        // With FP32 iterates the checkpoint delta cannot fall below a few
        // FP32 ulps, so a stall is declared there instead of at 1e-11.
        s.min_delta = s.mixed ? fmax(s.inner_min_delta, kMixedMinDelta)
                              : s.inner_min_delta;
        s.period = 10000;
        s.pc = PC_A_INIT;
        break;
      case PC_LE0:  // alm.c:258-264
        s.inner_reason = s.ret;
        issue(OP_ALM_VIOL, 0, PC_LE1, 2u);
        break;
      case PC_LE1: {  // alm.c:258-285, :194-212
        s.norm = sqrt(res[0]);
        s.max_viol = res[1];
        s.w1_pre = res[2];
        s.cx = res[3];
        s.alm_pg = s.diag[2];
        if (s.log)
          emit(ACD_EV_ALM_ITER, s.outer + 1, s.max_viol, s.norm, s.diag[2],
               s.diag[0], s.cx);
        const bool stalled = (s.inner_reason != 2);
        s.upd = 0;
        s.bump = 0;
        s.scale = 1.0;
        s.argmax = -1.0;
        if (!stalled) {
          if (!(s.norm < .5 * s.prev_norm)) {
            s.upd = 1;
            s.scale = 2 * fmin(1.0, s.norm / s.prev_norm);
            s.bump = s.max_viol > .5 * s.norm;
          }
          s.prev_norm = s.norm;
        }
        if (s.bump)
          issue(OP_ARGMAX, 0, PC_LE2, 1u, s.max_viol);
        else
          s.pc = PC_LE3;
        break;
      }
      case PC_LE2:
        s.argmax = -res[0];
        s.pc = PC_LE3;
        break;
      case PC_LE3:
        s.op_u = (uint32_t)(s.argmax < 0 ? 0xffffffffu : (uint32_t)s.argmax);
        issue(OP_LAM_W, (s.upd ? F_UPD : 0) | (s.bump ? F_BUMP : 0), PC_LE4, 0,
              s.scale);
        break;
      case PC_LE4: {  // alm.c:235-245, :273-285, :303-319
        const bool stalled = (s.inner_reason != 2);
        const double w1 = stalled ? s.w1_pre : res[0];
        const double wv = (double)a.m / w1;
        s.precision = fmin(s.precision, fmin(pow(s.max_viol, 1.0), pow(wv, 1.5)));
        s.precision = fmax(s.precision, s.precision_floor);
        if (s.log)
          emit(ACD_EV_ALM_TAIL, s.outer + 1, stalled ? 1.0 : 0.0,
               res[0] / (double)a.m);
        s.outer++;
        s.alm_diag[0] = s.cx;
        s.alm_diag[1] = s.max_viol;
        s.alm_diag[2] = s.alm_pg;
        if (s.stop_kkt > 0) {  // the KKT target replaces the absolute rule
          issue(OP_KKT, 0, PC_LE5);
        } else if (s.max_viol < s.stop_violation &&
                   (stalled || s.alm_pg < s.stop_pg)) {
          s.alm_ret = 0;
          s.pc = PC_DONE;
        } else if (s.outer >= s.outer_niter) {
          s.alm_ret = 1;
          s.pc = PC_DONE;
        } else {
          s.pc = PC_LB0;
        }
        break;
      }
      case PC_LE5: {  // stop on the relative KKT residual (stop_kkt > 0)
        const double rp = s.norm / (1 + sqrt(res[5]));
        const double rd = (sqrt(res[0]) + sqrt(res[1])) / (1 + sqrt(res[3]));
        const double dual = res[4] + res[2];
        const double rg = fabs(s.cx - dual) / (1 + fabs(s.cx) + fabs(dual));
        s.kkt = fmax(rp, fmax(rd, rg));
        if (s.kkt <= s.stop_kkt) {
          s.alm_ret = 0;
          s.pc = PC_DONE;
        } else if (s.outer >= s.outer_niter) {
          s.alm_ret = 1;
          s.pc = PC_DONE;
        } else {
          s.pc = PC_LB0;
        }
        break;
      }
      // ------------------------------------------------------ approx_update
      case PC_U0:
        issue(OP_PERMUTE_PROBLEM, 0, PC_U1);
        break;
      case PC_U1:
        issue(OP_STEP_SIZES, 0, PC_U2);
        break;
      case PC_U2:
        s.pc = PC_DONE;
        break;
      default:
        if constexpr (SolveOnly)
          fail(ACD_ERR_LAUNCH);
        else
          step_rest();
        break;
    }
  }

  // The SPG and parallel coordinate descent continuations (left out of
  // the batched kernel with their ops).
  __device__ void step_rest() {
    switch (s.pc) {
      // ------------------------------------------------------------ SPG
      case PC_M0:  // approx/approx.c:534-541
        issue(OP_INIT_X, 0, PC_M1);
        break;
      case PC_M1:
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_VIOL, 0, PC_M2);
        break;
      case PC_M2:
        s.quad_x = .5 * res[0];
        issue(OP_MPG_GRAD0, 0, PC_M3);
        break;
      case PC_M3:
        s.i = 0;
        s.reason = 0;
        s.pc = PC_MI;
        break;
      case PC_MI:  // approx/approx.c:543-545; mpg.inc:155-199
        if (s.i >= s.niter || s.reason) {
          s.op_r0 = NX;
          issue(OP_PERMUTE_OUT, F_CX, PC_MF1);
          break;
        }
        if (a.ev_cap && s.events + 8 > a.ev_cap) {
          s.events_full = 1;
          s.op = OP_NONE;
          return;
        }
        issue(OP_MPG_GRAD, 0, PC_MI2);
        break;
      case PC_MI2: {
        s.pg = sqrt(res[0]);
        double st = res[1] / res[2];
        if (isnan(st)) st = 1;
        if (!(st > 1e8)) {
          issue(OP_MPG_ZUPD, 0, PC_MI3, 0, st);
        } else {
          s.op_r0 = NX;
          s.op_r1 = MXV;
          issue(OP_VIOL, F_CX, PC_MI5);
        }
        break;
      }
      case PC_MI3:
        swap_(s.pn[NZ], s.pn[NX]);
        swap_(s.pn[NG2], s.pn[NG]);
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_VIOL, F_CX, PC_MI5);
        break;
      case PC_MI5: {  // approx/approx.c:546-553
        s.quad_x = .5 * res[0];
        const double value = s.offset + (res[1] + s.quad_x);
        if (s.pg < s.max_pg) s.reason = 2;
        if (value < s.max_value) s.reason = 1;
        if (s.log &&
            ((s.period && (s.i % s.period == 0)) || (s.i == 0) || s.reason))
          emit(ACD_EV_MPG, s.i + 1, s.pg, value);
        s.value = value;
        s.inner_total++;
        s.i += 1;
        s.pc = PC_MI;
        break;
      }
      case PC_MF1:  // approx/approx.c:559-569
        s.cx = res[0];
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_GRAD_PG, 0, PC_MF2);
        break;
      case PC_MF2:
        s.diag[0] = s.offset + (s.cx + s.quad_x);
        s.diag[1] = sqrt(res[1]);
        s.diag[2] = sqrt(res[0]);
        s.diag[3] = HUGE_VAL;
        s.diag[4] = (double)(s.i + 1);
        s.ret = 2;
        s.pc = PC_DONE;
        break;
      // ------------------------------------------- parallel coordinate descent
      // Epochs of pcd_steps sampled steps each; after every step the
      // violation and value at x are recomputed; after every epoch the
      // projected gradient is, and the stopping tests of approx/cd.c:
      // 353-366 apply (value < max_value: 1, pg < max_pg: 2).
      case PC_P0:
        issue(OP_INIT_X, 0, PC_PZ);
        break;
      case PC_PZ:
        issue(OP_PCD_SIZES, 0, PC_P1);
        break;
      case PC_P1:
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_VIOL, F_CX, PC_P2);
        break;
      case PC_P2:
        s.quad_x = .5 * res[0];
        s.value = res[1] + s.quad_x + s.offset;
        s.i = 0;
        s.reason = 0;
        s.pg = HUGE_VAL;
        s.pc = PC_PI;
        break;
      case PC_PI:
        if (s.i >= s.niter) {
          s.pc = PC_PF;
          break;
        }
        s.k = 0;
        s.pc = PC_PS;
        break;
      case PC_PS:
        if (s.k < s.pcd_steps) {
          issue(OP_PCD_STEP, 0, PC_PV);
        } else {
          s.op_r0 = NX;
          s.op_r1 = MXV;
          issue(OP_GRAD_PG, 0, PC_PE);
        }
        break;
      case PC_PV:
        s.k += 1;
        s.op_r0 = NX;
        s.op_r1 = MXV;
        issue(OP_VIOL, F_CX, PC_PV2);
        break;
      case PC_PV2:
        s.quad_x = .5 * res[0];
        s.value = res[1] + s.quad_x + s.offset;
        s.pc = PC_PS;
        break;
      case PC_PE:
        s.pg = sqrt(res[0]);
        s.diag[1] = sqrt(res[1]);
        s.diag[2] = s.pg;
        s.inner_total++;
        s.i += 1;
        if (s.value < s.max_value) {
          s.reason = 1;
          s.pc = PC_PF;
        } else if (s.pg < s.max_pg) {
          s.reason = 2;
          s.pc = PC_PF;
        } else {
          s.pc = PC_PI;
        }
        break;
      case PC_PF:
        s.op_r0 = NX;
        issue(OP_PERMUTE_OUT, 0, PC_PF2);
        break;
      case PC_PF2:
        s.diag[0] = s.value;
        s.diag[3] = 0;
        s.diag[4] = (double)(s.reason ? s.i : s.i + 1);  // cd.c:369-378
        s.ret = s.reason;
        s.pc = PC_DONE;
        break;
      default:
        fail(ACD_ERR_LAUNCH);
        break;
    }
  }
};

}  // namespace acd
