// solver_core.hpp -- the LP hot path as one device-resident state machine:
// ALM outer loop (alm.c:247-322), APPROX inner loop (approx/approx.c:
// 417-708) and spectral projected gradient (approx/mpg.inc:155-199,
// approx/approx.c:517-573) of pkhuong/acd-alm.
//
// The solver works on the row/column-permuted working copy of the problem
// (approx/approx.c:84-88) stored as sliced ELL (64 rows per slice).  Each
// phase below is one or more passes over rows, separated by group-wide
// reductions.  Control decisions (step acceptance, restarts, stopping)
// are taken from reduced scalars that every thread holds bit-identically,
// so every thread follows the same branch and no decision goes back to
// the host.  The state persists in device memory between launches: a
// launch runs until the iteration budget is used, the event buffer is
// nearly full, or the solve ends.
//
// Pass structure of one APPROX iteration (common long-step case):
//   pass 1 (over A' rows j): g2 = A'(y.viol) + c, [g = A'(z.viol) + c],
//          zp = clamp(z - g2 * itheta / v), model-decrease partials, c'z,
//          c'zp, and the pending x <- (1-t)x + t z update of the previous
//          iteration (fused: that update reads the vector pass 1 reads);
//   pass 2 (over A rows i):  zp.viol = w (A zp - b), 1/2 sum zp.viol*(A zp-b)
//          and, speculatively, the violation part of the momentum update
//          into shadow buffers that are adopted by swapping roles when the
//          step is accepted.
// Rejected long steps, restarts and the every-100-iteration checkpoint
// take extra passes, exactly where the reference takes extra work.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "acd_dev.h"
#include "solver_grid.hpp"

namespace acd {

enum Phase : int32_t {
  PH_DONE = 0,
  PH_APPROX_INIT,
  PH_APPROX_ITER,
  PH_APPROX_FINISH,
  PH_MPG_INIT,
  PH_MPG_ITER,
  PH_MPG_FINISH,
  PH_ALM_START,
  PH_ALM_BEGIN,
  PH_ALM_END,
  PH_UPDATE_ONLY
};

// n-side and m-side work vector roles (indices into Args::vn / Args::vm).
enum { NX = 0, NZ, NZP, NG, NG2, NPV, NY, NU, kNumVN };
enum { MXV = 0, MYV, MZV, MZPV, MSX, MSY, kNumVM };

struct Sell {
  const uint64_t* off;
  const uint32_t* col;
  const double* val;
  uint32_t nrows, nslices;
};

struct State {
  int32_t phase, mode, log, error;
  int32_t ret, reason, center_x, restart_flag;
  int32_t lazy_x, prevx_pending, violp_zp, pad0;
  double* pn[kNumVN];   // current vector of each n-side role
  double* pm[kNumVM];   // current vector of each m-side role
  uint64_t niter, period, i, k;
  double max_pg, max_value, min_delta, offset;
  double theta, step_length, lazy_theta;
  double quad_x, quad_z, quad_zp;
  double value, pg, delta;
  double diag[5];
  // ALM
  uint64_t outer, outer_niter;
  double precision, prev_norm, max_viol, alm_pg, cx;
  double stop_violation, stop_pg, precision_floor, inner_min_delta;
  double alm_diag[3];
  int32_t alm_ret, inner_reason;
  // bookkeeping
  uint64_t inner_total, events;
};

struct Args {
  uint32_t m, n;
  Sell A, AT, A0, AT0;          // working copy and caller's matrix
  const uint32_t* row_idx;      // working copy row permutation
  const uint32_t* col_idx;
  const uint32_t* col_ridx;
  const uint32_t* beta;         // entries per row of the working copy
  // caller-order vectors
  double *rhs, *weight, *linc, *lower, *upper, *alm_c, *lam_lo, *lam_hi;
  double *x_io, *lam, *r;
  // working-copy vectors
  double *p_rhs, *p_w, *p_c, *p_l, *p_u, *p_v, *p_iv;
  State* st;
  SyncArea* sa;
  acd_event* ev;
  uint32_t ev_cap;
  uint64_t budget;
};

// SSE2 maxpd(a, b) / minpd(a, b) as used by the reference's box
// projections (approx/step_project.inc:18-19, :53-54).
__device__ __forceinline__ double maxpd(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double minpd(double a, double b) { return a < b ? a : b; }

template <class G>
struct Solver {
  G& g;
  const Args& a;
  State s;

  __device__ Solver(G& g_, const Args& a_) : g(g_), a(a_) {}

  __device__ __forceinline__ double* VN(int role) const { return s.pn[role]; }
  __device__ __forceinline__ double* VM(int role) const { return s.pm[role]; }

  // ---- traversal helpers ----
  // Row r of M, lane-sequential in ascending column order, starting from
  // +0.0; F(acc, value, column) folds one entry, E(row, acc0, acc1) is
  // the epilogue.
  template <int NRHS, class E>
  __device__ __forceinline__ void spmv_rows(const Sell& M, const double* x0,
                                            const double* x1, E&& epi) {
    for (uint32_t sl = g.gwave; sl < M.nslices; sl += g.nwaves) {
      const uint64_t base = M.off[sl];
      const uint32_t len = (uint32_t)((M.off[sl + 1] - base) >> 6);
      const double* v = M.val + base + g.lane;
      const uint32_t* c = M.col + base + g.lane;
      double a0 = 0.0, a1 = 0.0;
      uint32_t k = 0;
      for (; k + 4 <= len; k += 4, v += 256, c += 256) {
        const double v0 = __builtin_nontemporal_load(v);
        const double v1 = __builtin_nontemporal_load(v + 64);
        const double v2 = __builtin_nontemporal_load(v + 128);
        const double v3 = __builtin_nontemporal_load(v + 192);
        const uint32_t c0 = __builtin_nontemporal_load(c);
        const uint32_t c1 = __builtin_nontemporal_load(c + 64);
        const uint32_t c2 = __builtin_nontemporal_load(c + 128);
        const uint32_t c3 = __builtin_nontemporal_load(c + 192);
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
      }
      for (; k < len; k++, v += 64, c += 64) {
        const double vv = *v;
        const uint32_t cc = *c;
        a0 = a0 + vv * x0[cc];
        if (NRHS == 2) a1 = a1 + vv * x1[cc];
      }
      const uint32_t r = sl * 64 + g.lane;
      if (r < M.nrows) epi(r, a0, a1);
    }
  }

  template <class F>
  __device__ __forceinline__ void each(uint32_t n, F&& f) {
    const uint32_t ns = (n + 63) >> 6;
    for (uint32_t sl = g.gwave; sl < ns; sl += g.nwaves) {
      const uint32_t r = sl * 64 + g.lane;
      if (r < n) f(r);
    }
  }

  __device__ __forceinline__ void emit(int32_t type, uint64_t k, double a0 = 0, double a1 = 0,
                       double a2 = 0, double a3 = 0, double a4 = 0,
                       double a5 = 0) {
    if (s.events < a.ev_cap) {
      if (g.block == 0 && threadIdx.x == 0) {
        acd_event* e = &a.ev[s.events];
        e->type = type;
        e->pad = 0;
        e->k = k;
        e->a[0] = a0; e->a[1] = a1; e->a[2] = a2;
        e->a[3] = a3; e->a[4] = a4; e->a[5] = a5;
      }
      s.events++;
    }
  }

  __device__ __forceinline__ void fail(int32_t code) {
    s.error = code;
    s.phase = PH_DONE;
  }

  // ---- building blocks ----

  // x.viol = w (A x - b) on the working copy; returns 1/2 sum viol*(Ax-b)
  // (approx/gradient_value.inc:15-81).  Optional copies of the violation.
  __device__ __forceinline__ double violation(const double* X, double* V, double* copy1,
                              double* copy2) {
    double q[1] = {0.0};
    const double* b = a.p_rhs;
    const double* w = a.p_w;
    spmv_rows<1>(a.A, X, nullptr, [&](uint32_t i, double ax, double) {
      const double v = ax - b[i];
      const double sv = w[i] * v;
      V[i] = sv;
      if (copy1) copy1[i] = sv;
      if (copy2) copy2[i] = sv;
      q[0] += sv * v;
    });
    g.reduce(q, 0);
    return .5 * q[0];
  }

  // g = A'(viol) + c, with pg^2 = sum (x - clamp(x - g))^2 and |g|^2
  // (approx/gradient_value.inc:158-181, approx/approx.c:258-278).
  __device__ __forceinline__ void gradient_pg(const double* V, const double* X, double* Gout,
                              double& pg2, double& gn2) {
    double r[2] = {0.0, 0.0};
    spmv_rows<1>(a.AT, V, nullptr, [&](uint32_t j, double gx, double) {
      const double gg = gx + a.p_c[j];
      Gout[j] = gg;
      const double xj = X[j];
      double xp = xj - gg;
      xp = maxpd(a.p_l[j], xp);
      xp = minpd(a.p_u[j], xp);
      const double d = xj - xp;
      r[0] += d * d;
      r[1] += gg * gg;
    });
    g.reduce(r, 0);
    pg2 = r[0];
    gn2 = r[1];
  }

  __device__ __forceinline__ double norm2_sq(const double* X) {
    double r[1] = {0.0};
    each(a.n, [&](uint32_t j) { r[0] += X[j] * X[j]; });
    g.reduce(r, 0);
    return r[0];
  }

  // Step sizes of the working copy (approx/approx.c:135-186):
  // v_j = sum over the column, ascending row, of ((w*beta)*a)*a.
  __device__ __forceinline__ void step_sizes() {
    double* wb = VM(MSX);
    each(a.m, [&](uint32_t i) { wb[i] = a.p_w[i] * (double)a.beta[i]; });
    g.sync();
    const Sell& M = a.AT;
    for (uint32_t sl = g.gwave; sl < M.nslices; sl += g.nwaves) {
      const uint64_t base = M.off[sl];
      const uint32_t len = (uint32_t)((M.off[sl + 1] - base) >> 6);
      double acc = 0.0;
      for (uint32_t k = 0; k < len; k++) {
        const uint64_t e = base + (uint64_t)k * 64 + g.lane;
        const double val = M.val[e];
        acc = acc + (wb[M.col[e]] * val) * val;
      }
      const uint32_t j = sl * 64 + g.lane;
      if (j < M.nrows) {
        a.p_v[j] = acc;
        a.p_iv[j] = (acc == 0 && a.p_c[j] == 0) ? 1.0 : 1.0 / acc;
      }
    }
    g.sync();
  }

  // Permute b, w, c', l, u into the working copy
  // (approx/approx.c:192-211; spmv/permutation.c:136-148).
  __device__ __forceinline__ void permute_problem() {
    each(a.m, [&](uint32_t i) {
      const uint32_t src = a.row_idx[i];
      a.p_rhs[i] = a.rhs[src];
      a.p_w[i] = a.weight[src];
    });
    each(a.n, [&](uint32_t j) {
      const uint32_t src = a.col_idx[j];
      a.p_c[j] = a.linc[src];
      a.p_l[j] = a.lower[src];
      a.p_u[j] = a.upper[src];
    });
    g.sync();
  }

  // ---- APPROX ----

  // approx/approx.c:588-607
  __device__ __forceinline__ void approx_init() {
    double* X = VN(NX);
    double* Z = VN(NZ);
    double* PV = VN(NPV);
    double r[1] = {0.0};
    each(a.n, [&](uint32_t j) {
      double xj = a.x_io[a.col_idx[j]];
      const double lo = maxpd(a.p_l[j], xj);
      xj = minpd(lo, a.p_u[j]);
      X[j] = xj;
      Z[j] = xj;
      PV[j] = xj;
      r[0] += a.p_c[j] * xj;
    });
    g.reduce(r, 0);
    const double cx = r[0];
    const double q = violation(X, VM(MXV), VM(MZV), VM(MYV));
    s.quad_x = s.quad_z = q;
    s.violp_zp = 0;
    s.value = cx + q + s.offset;
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
    s.phase = s.niter > 0 ? PH_APPROX_ITER : PH_APPROX_FINISH;
  }

  struct P1 {
    double lin, quad, cz, czp, maxz, dd, pg2;
  };

  // Pass 1 (approx/step_project.inc:179-236, :320-375, :452-508).
  __device__ __forceinline__ void pass1(bool two, bool lng, double itheta, double theta2,
                        bool want_pg, P1& o) {
    double* Zp = VN(NZP);
    const double* Z = VN(NZ);
    double* G1 = VN(NG);
    double* G2 = VN(NG2);
    double* X = VN(NX);
    double* PV = VN(NPV);
    const double* YV = VM(MYV);
    const double* ZV = VM(MZV);
    const bool lazy = s.lazy_x, pcopy = s.prevx_pending;
    const double lt = s.lazy_theta, ls = 1 - s.lazy_theta;
    double r[7] = {0, 0, 0, 0, 0, 0, 0};
    auto epi = [&](uint32_t j, double gy, double gz) {
      const double cj = a.p_c[j];
      const double zj = Z[j];
      if (pcopy || lazy) {
        double xj = X[j];
        if (pcopy) PV[j] = xj;
        if (lazy) X[j] = ls * xj + lt * zj;
      }
      const double g2 = gy + cj;
      G2[j] = g2;
      const double stp = itheta * a.p_iv[j];
      double t = zj - g2 * stp;
      t = maxpd(a.p_l[j], t);
      t = minpd(a.p_u[j], t);
      Zp[j] = t;
      r[4] = maxpd(r[4], fabs(t));
      r[2] += cj * zj;
      r[3] += cj * t;
      if (lng) {
        const double d = t - zj;
        r[0] += g2 * d;
        r[1] += d * d * theta2 * a.p_v[j];
      }
      if (two) {
        const double gg = gz + cj;
        G1[j] = gg;
        r[5] += gg * (t - zj);
        if (want_pg) {
          double xp = zj - gg;
          xp = maxpd(a.p_l[j], xp);
          xp = minpd(a.p_u[j], xp);
          const double d = zj - xp;
          r[6] += d * d;
        }
      }
    };
    if (two)
      spmv_rows<2>(a.AT, YV, ZV, epi);
    else
      spmv_rows<1>(a.AT, YV, nullptr, epi);
    g.reduce(r, 1u << 4);
    s.lazy_x = 0;
    s.prevx_pending = 0;
    o.lin = r[0];
    o.quad = r[1];
    o.cz = r[2];
    o.czp = r[3];
    o.maxz = r[4];
    o.dd = r[5];
    o.pg2 = r[6];
  }

  // Pass 2: zp.viol and quad(zp); momentum update of x.viol / y.viol
  // (approx/vector.inc:249-295), into shadows when speculative.
  __device__ __forceinline__ double pass2(bool spec, double th, double thn) {
    const double* Zp = VN(NZP);
    double* ZPV = VM(MZPV);
    double* XV = VM(MXV);
    double* YV = VM(MYV);
    double* SX = VM(MSX);
    double* SY = VM(MSY);
    const double sc = 1 - th, nsc = 1 - thn;
    double q[1] = {0.0};
    spmv_rows<1>(a.A, Zp, nullptr, [&](uint32_t i, double ax, double) {
      const double v = ax - a.p_rhs[i];
      const double sv = a.p_w[i] * v;
      ZPV[i] = sv;
      q[0] += sv * v;
      const double xv = sc * XV[i] + th * sv;
      const double yv = nsc * xv + thn * sv;
      if (spec) {
        SX[i] = xv;
        SY[i] = yv;
      } else {
        XV[i] = xv;
        YV[i] = yv;
      }
    });
    g.reduce(q, 0);
    return .5 * q[0];
  }

  // After a rejected step: [g at z], [re-step from g2], dot_diff
  // (approx/approx.c:383-389, :445-450, step_project.inc:34-94).
  __device__ __forceinline__ double fixup(bool need_g, bool restep, double itheta,
                          double& maxz) {
    double* Zp = VN(NZP);
    const double* Z = VN(NZ);
    double* G1 = VN(NG);
    const double* G2 = VN(NG2);
    double r[2] = {0.0, 0.0};
    auto body = [&](uint32_t j, double gz) {
      const double zj = Z[j];
      double gg;
      if (need_g) {
        gg = gz + a.p_c[j];
        G1[j] = gg;
      } else {
        gg = G1[j];
      }
      double t;
      if (restep) {
        const double stp = itheta * a.p_iv[j];
        t = zj - G2[j] * stp;
        t = maxpd(a.p_l[j], t);
        t = minpd(a.p_u[j], t);
        Zp[j] = t;
        r[1] = maxpd(r[1], fabs(t));
      } else {
        t = Zp[j];
      }
      r[0] += gg * (t - zj);
    };
    if (need_g)
      spmv_rows<1>(a.AT, VM(MZV), nullptr,
                   [&](uint32_t j, double gz, double) { body(j, gz); });
    else
      each(a.n, [&](uint32_t j) { body(j, 0.0); });
    g.reduce(r, 2u);
    maxz = r[1];
    return r[0];
  }

  template <class T>
  __device__ static __forceinline__ void swap_(T& x, T& y) {
    T t = x;
    x = y;
    y = t;
  }

  // One iteration of approx/approx.c:417-478 followed by the per-iteration
  // checks of approx/approx.c:608-675.
  __device__ __forceinline__ void approx_iter() {
    if (s.i >= s.niter) {
      s.phase = PH_APPROX_FINISH;
      return;
    }
    const uint64_t i = s.i;
    const bool pg_req = (i == 0) || ((i + 1) % 10 == 0);
    const double len = s.step_length;
    bool descent, got_gradient, restep = false;
    double pg = HUGE_VAL, dd = 0.0, cz = 0.0;
    bool have_dd = false;
    P1 o;
    if (len <= 1) {
      pass1(true, false, (1 - 1e-6) / s.theta, 0.0, pg_req, o);
      if (g.aborted) return;
      if (!(o.maxz < HUGE_VAL)) return fail(ACD_ERR_UNBOUNDED);
      s.step_length = len * 1.01;
      if (s.step_length > 1) s.step_length = 1 + 1e-6;
      descent = false;
      got_gradient = true;
      dd = o.dd;
      have_dd = true;
      s.violp_zp = 0;
    } else {
      got_gradient = pg_req || (len <= 1 + 1e-6);
      const bool safe = (len <= 1 + 1e-6);
      const double L = safe ? 1.0 : len;
      pass1(got_gradient, true, (L - 1e-6) / s.theta, s.theta / L, pg_req, o);
      if (g.aborted) return;
      const double expected = (o.lin) + .5 * (o.quad);
      const double thn = 2.0 / (double)(s.k + 1 + 2);
      const double qzp = pass2(true, s.theta, thn);
      if (g.aborted) return;
      s.quad_zp = qzp;
      s.violp_zp = 1;
      const double initial = o.cz + s.quad_z;
      const double now = o.czp + qzp;
      if (!(expected <= 0)) return fail(ACD_ERR_ESTIMATE);
      if (now <= initial + expected) {
        s.step_length = L * 1.01;
        descent = true;
      } else {
        s.step_length = .9 * L;
        descent = false;
        restep = !safe;
      }
      if (got_gradient && !restep) {
        dd = o.dd;
        have_dd = true;
      }
    }
    cz = o.cz;
    if (pg_req) pg = sqrt(o.pg2);
    s.pg = pg;
    s.delta = HUGE_VAL;
    bool restarted = false;
    if (!descent) {
      if (!have_dd) {
        double maxz = 0.0;
        dd = fixup(!got_gradient, restep, (1 - 1e-6) / s.theta, maxz);
        if (g.aborted) return;
        if (restep) {
          if (!(maxz < HUGE_VAL)) return fail(ACD_ERR_UNBOUNDED);
          s.violp_zp = 0;
        }
      }
      if (dd > 0) {
        // Oscillation: x = y = z, theta = 1 (approx/approx.c:449-457).
        double* X = VN(NX);
        const double* Z = VN(NZ);
        double* XV = VM(MXV);
        double* YV = VM(MYV);
        const double* ZV = VM(MZV);
        each(a.n, [&](uint32_t j) { X[j] = Z[j]; });
        each(a.m, [&](uint32_t r) {
          const double v = ZV[r];
          XV[r] = v;
          YV[r] = v;
        });
        g.sync();
        if (g.aborted) return;
        s.quad_x = s.quad_z;
        s.lazy_x = 0;
        s.k = 0;
        s.theta = 1;
        s.center_x = 1;
        restarted = true;
      }
    }
    if (!restarted) {
      const double th = s.theta;
      const double thn = 2.0 / (double)(s.k + 1 + 2);
      if (!s.violp_zp) {
        s.quad_zp = pass2(false, th, thn);
        if (g.aborted) return;
        s.violp_zp = 1;
      } else {
        swap_(s.pm[MXV], s.pm[MSX]);
        swap_(s.pm[MYV], s.pm[MSY]);
      }
      s.k += 1;
      s.theta = thn;
      s.lazy_x = 1;
      s.lazy_theta = th;
      // swap z and zp: the centre is the old z
      swap_(s.pn[NZ], s.pn[NZP]);
      swap_(s.pm[MZV], s.pm[MZPV]);
      swap_(s.quad_z, s.quad_zp);
      s.center_x = 0;
    }
    s.inner_total++;
    // driver: value of the centre (approx/approx.c:617-639)
    if (restarted) {
      if (s.log) emit(ACD_EV_RESTART, i);
      s.restart_flag = 1;
    }
    double value = cz + (s.center_x ? s.quad_x : s.quad_zp) + s.offset;
    s.value = value;
    if (value < s.max_value) return stop(1);
    if (pg < s.max_pg) return stop(2);
    if ((i + 1) % 100 == 0) {
      checkpoint();
      if (g.aborted || s.phase == PH_DONE) return;
      if (s.value < s.max_value) return stop(1);
      if (s.pg < s.max_pg) return stop(2);
      if (s.delta < s.min_delta) return stop(3);
      s.prevx_pending = 1;
    }
    if (i == 0 || (s.period && (i + 1) % s.period == 0)) {
      if (s.restart_flag) {
        emit(ACD_EV_NEWLINE, i);
        s.restart_flag = 0;
      }
      if (s.log) {
        const double ng = sqrt(norm2_sq(VN(NG)));
        emit(ACD_EV_PRINT, i + 1, s.value, ng, s.pg, s.step_length, s.delta);
      }
    }
    s.i = i + 1;
  }

  __device__ __forceinline__ void stop(int32_t reason) {
    s.reason = reason;
    s.phase = PH_APPROX_FINISH;
  }

  // Every 100 iterations: evaluate at x (approx/approx.c:641-665).
  __device__ __forceinline__ void checkpoint() {
    double* X = VN(NX);
    const double* Z = VN(NZ);
    const double* PV = VN(NPV);
    const bool lazy = s.lazy_x;
    const double lt = s.lazy_theta, ls = 1 - s.lazy_theta;
    double r[3] = {0.0, 0.0, 0.0};
    each(a.n, [&](uint32_t j) {
      double xj = X[j];
      if (lazy) {
        xj = ls * xj + lt * Z[j];
        X[j] = xj;
      }
      r[0] += a.p_c[j] * xj;
      const double d = PV[j] - xj;
      r[1] += d * d;
      r[2] += xj * xj;
    });
    g.reduce(r, 0);
    if (g.aborted) return;
    s.lazy_x = 0;
    const double q = violation(X, VM(MXV), nullptr, nullptr);
    if (g.aborted) return;
    s.quad_x = q;
    s.value = r[0] + q + s.offset;
    double pg2, gn2;
    gradient_pg(VM(MXV), X, VN(NG), pg2, gn2);
    s.pg = sqrt(pg2);
    s.delta = sqrt(r[1]) / (sqrt(r[2]) + 1e-10);
    s.center_x = 1;
  }

  // approx/approx.c:676-707
  __device__ __forceinline__ void approx_finish() {
    const double* C = s.center_x ? VN(NX) : VN(NZP);
    const double* CV = s.center_x ? VM(MXV) : VM(MZPV);
    const double cq = s.center_x ? s.quad_x : s.quad_zp;
    const double* X = VN(NX);
    const double* PV = VN(NPV);
    const bool pcopy = s.prevx_pending;
    double r[3] = {0.0, 0.0, 0.0};
    each(a.n, [&](uint32_t j) {
      const double cj = C[j];
      const double pv = pcopy ? X[j] : PV[j];
      const double d = pv - cj;
      r[0] += d * d;
      r[1] += cj * cj;
      r[2] += a.p_c[j] * cj;
    });
    g.reduce(r, 0);
    if (g.aborted) return;
    const double delta = sqrt(r[0]) / (sqrt(r[1]) + 1e-10);
    const double value = r[2] + cq + s.offset;
    double pg2, gn2;
    gradient_pg(CV, C, VN(NG), pg2, gn2);
    if (g.aborted) return;
    each(a.n, [&](uint32_t i) { a.x_io[i] = C[a.col_ridx[i]]; });
    g.sync();
    if (g.aborted) return;
    if (s.restart_flag) {
      emit(ACD_EV_NEWLINE, s.i);
      s.restart_flag = 0;
    }
    s.diag[0] = value;
    s.diag[1] = sqrt(gn2);
    s.diag[2] = sqrt(pg2);
    s.diag[3] = delta;
    s.diag[4] = (double)(s.i + 1);
    if (s.log)
      emit(ACD_EV_PRINT, s.i + 1, value, s.diag[1], s.diag[2], s.step_length,
           delta);
    s.ret = s.reason;
    s.phase = (s.mode == ACD_MODE_ALM) ? PH_ALM_END : PH_DONE;
  }

  // ---- SPG (approx/approx.c:517-573, approx/mpg.inc:155-199) ----
  __device__ __forceinline__ void mpg_init() {
    double* X = VN(NX);
    double* Z = VN(NZ);
    each(a.n, [&](uint32_t j) {
      double xj = a.x_io[a.col_idx[j]];
      const double lo = maxpd(a.p_l[j], xj);
      xj = minpd(lo, a.p_u[j]);
      X[j] = xj;
      Z[j] = xj;
    });
    g.sync();
    s.quad_x = violation(X, VM(MXV), nullptr, nullptr);
    double* G1 = VN(NG);
    double* G2 = VN(NG2);
    spmv_rows<1>(a.AT, VM(MXV), nullptr, [&](uint32_t j, double gx, double) {
      const double gg = gx + a.p_c[j];
      G1[j] = gg;
      G2[j] = gg;
    });
    g.sync();
    s.i = 0;
    s.reason = 0;
    s.phase = s.niter > 0 ? PH_MPG_ITER : PH_MPG_FINISH;
  }

  __device__ __forceinline__ void mpg_iter() {
    if (s.i >= s.niter || s.reason) {
      s.phase = PH_MPG_FINISH;
      return;
    }
    const uint64_t i = s.i;
    double* X = VN(NX);
    double* Z = VN(NZ);
    double* G1 = VN(NG);
    const double* G2 = VN(NG2);
    // x is already inside the box (projection is idempotent), and its
    // violation is current, so g = A'(x.viol) + c directly.
    double r[3] = {0.0, 0.0, 0.0};
    spmv_rows<1>(a.AT, VM(MXV), nullptr, [&](uint32_t j, double gx, double) {
      const double gg = gx + a.p_c[j];
      G1[j] = gg;
      const double xj = X[j];
      double xp = xj - gg;
      xp = maxpd(a.p_l[j], xp);
      xp = minpd(a.p_u[j], xp);
      const double d = xj - xp;
      r[0] += d * d;
      const double yj = xj - Z[j];
      const double uj = gg - G2[j];
      r[1] += yj * yj;
      r[2] += yj * uj;
    });
    g.reduce(r, 0);
    if (g.aborted) return;
    const double pg = sqrt(r[0]);
    double st = r[1] / r[2];
    if (isnan(st)) st = 1;
    if (!(st > 1e8)) {
      each(a.n, [&](uint32_t j) {
        double t = X[j] - st * G1[j];
        const double lo = maxpd(a.p_l[j], t);
        Z[j] = minpd(lo, a.p_u[j]);
      });
      swap_(s.pn[NZ], s.pn[NX]);
      swap_(s.pn[NG2], s.pn[NG]);
      g.sync();
      if (g.aborted) return;
    }
    // value = offset + c'x + quad(x) (approx/approx.c:546)
    const double* Xn = VN(NX);
    double cx[1] = {0.0};
    each(a.n, [&](uint32_t j) { cx[0] += a.p_c[j] * Xn[j]; });
    g.reduce(cx, 0);
    s.quad_x = violation(Xn, VM(MXV), nullptr, nullptr);
    if (g.aborted) return;
    const double value = s.offset + (cx[0] + s.quad_x);
    if (pg < s.max_pg) s.reason = 2;
    if (value < s.max_value) s.reason = 1;
    if (s.log && ((s.period && (i % s.period == 0)) || (i == 0) || s.reason))
      emit(ACD_EV_MPG, i + 1, pg, value);
    s.value = value;
    s.inner_total++;
    s.i = i + 1;
  }

  __device__ __forceinline__ void mpg_finish() {
    const double* X = VN(NX);
    each(a.n, [&](uint32_t i) { a.x_io[i] = X[a.col_ridx[i]]; });
    double cx[1] = {0.0};
    each(a.n, [&](uint32_t j) { cx[0] += a.p_c[j] * X[j]; });
    g.reduce(cx, 0);
    double pg2, gn2;
    gradient_pg(VM(MXV), X, VN(NG), pg2, gn2);
    if (g.aborted) return;
    s.diag[0] = s.offset + (cx[0] + s.quad_x);
    s.diag[1] = sqrt(gn2);
    s.diag[2] = sqrt(pg2);
    s.diag[3] = HUGE_VAL;
    s.diag[4] = (double)(s.i + 1);
    s.ret = 2;
    s.phase = PH_DONE;
  }

  // ---- ALM (alm.c:106-322) ----
  __device__ __forceinline__ void alm_start() {
    each(a.m, [&](uint32_t i) {
      a.lam[i] = fmax(a.lam_lo[i], fmin(a.lam[i], a.lam_hi[i]));
    });
    double cx[1] = {0.0};
    each(a.n, [&](uint32_t j) { cx[0] += a.x_io[j] * a.alm_c[j]; });
    g.reduce(cx, 0);
    s.cx = cx[0];
    s.outer = 0;
    s.precision = 1;
    s.prev_norm = HUGE_VAL;
    s.max_viol = HUGE_VAL;
    s.alm_pg = HUGE_VAL;
    s.alm_ret = 1;
    s.phase = s.outer_niter ? PH_ALM_BEGIN : PH_DONE;
    s.alm_diag[0] = s.cx;
    s.alm_diag[1] = s.max_viol;
    s.alm_diag[2] = s.alm_pg;
  }

  __device__ __forceinline__ void alm_begin() {
    // c' = c - A' lambda (caller's matrix), offset = lambda . b
    double r[1] = {0.0};
    spmv_rows<1>(a.AT0, a.lam, nullptr, [&](uint32_t j, double atl, double) {
      a.linc[j] = a.alm_c[j] - atl;
    });
    each(a.m, [&](uint32_t i) { r[0] += a.lam[i] * a.rhs[i]; });
    g.reduce(r, 0);
    if (g.aborted) return;
    s.offset = r[0];
    permute_problem();
    step_sizes();
    if (g.aborted) return;
    s.niter = 4294967295ull;  // -1u (alm.c:255)
    s.max_pg = fmin(1.0, s.precision);
    s.max_value = -HUGE_VAL;
    s.min_delta = s.inner_min_delta;
    s.period = 10000;
    s.phase = PH_APPROX_INIT;
  }

  __device__ __forceinline__ void alm_end() {
    s.inner_reason = s.ret;
    double r[4] = {0.0, -HUGE_VAL, 0.0, 0.0};  // |r|^2, max|r|, |w|_1, c.x
    spmv_rows<1>(a.A0, a.x_io, nullptr, [&](uint32_t i, double ax, double) {
      const double ri = ax - a.rhs[i];
      a.r[i] = ri;
      r[0] += ri * ri;
      r[1] = fmax(r[1], fabs(ri));
      r[2] += fabs(a.weight[i]);
    });
    each(a.n, [&](uint32_t j) { r[3] += a.x_io[j] * a.alm_c[j]; });
    g.reduce(r, 2u);
    if (g.aborted) return;
    const double norm = sqrt(r[0]);
    const double max_viol = r[1];
    s.cx = r[3];
    s.max_viol = max_viol;
    s.alm_pg = s.diag[2];
    if (s.log)
      emit(ACD_EV_ALM_ITER, s.outer + 1, max_viol, norm, s.diag[2], s.diag[0],
           s.cx);
    const bool stalled = (s.inner_reason != 2);
    bool upd = false, bump = false;
    double scale = 1.0;
    double argmax = -1.0;
    if (!stalled) {
      if (!(norm < .5 * s.prev_norm)) {
        upd = true;
        scale = 2 * fmin(1.0, norm / s.prev_norm);
        bump = max_viol > .5 * norm;
      }
      s.prev_norm = norm;
      if (bump) {
        // first index attaining the max (alm.c:155-166)
        double mi[1] = {-HUGE_VAL};
        each(a.m, [&](uint32_t i) {
          if (fabs(a.r[i]) == max_viol) mi[0] = fmax(mi[0], -(double)i);
        });
        g.reduce(mi, 1u);
        if (g.aborted) return;
        argmax = -mi[0];
      }
    }
    double wsum[1] = {0.0};
    const uint32_t am = (uint32_t)argmax;
    each(a.m, [&](uint32_t i) {
      double li = a.lam[i] - a.weight[i] * a.r[i];
      a.lam[i] = fmax(a.lam_lo[i], fmin(li, a.lam_hi[i]));
      if (upd) {
        double w = a.weight[i] * scale;
        if (bump && i == am) w *= 10;
        a.weight[i] = w;
      }
      wsum[0] += fabs(a.weight[i]);
    });
    g.reduce(wsum, 0);
    if (g.aborted) return;
    const double w1 = stalled ? r[2] : wsum[0];
    const double wv = (double)a.m / w1;
    s.precision = fmin(s.precision, fmin(pow(max_viol, 1.0), pow(wv, 1.5)));
    s.precision = fmax(s.precision, s.precision_floor);
    if (s.log) emit(ACD_EV_ALM_TAIL, s.outer + 1, stalled ? 1.0 : 0.0,
                    wsum[0] / (double)a.m);
    s.outer++;
    s.alm_diag[0] = s.cx;
    s.alm_diag[1] = s.max_viol;
    s.alm_diag[2] = s.alm_pg;
    if (max_viol < s.stop_violation && (stalled || s.alm_pg < s.stop_pg)) {
      s.alm_ret = 0;
      s.phase = PH_DONE;
    } else if (s.outer >= s.outer_niter) {
      s.alm_ret = 1;
      s.phase = PH_DONE;
    } else {
      s.phase = PH_ALM_BEGIN;
    }
  }

  __device__ __forceinline__ void update_only() {
    permute_problem();
    step_sizes();
    s.phase = PH_DONE;
  }

  // ---- driver ----
  __device__ __forceinline__ void run() {
    uint64_t spent = 0;
    while (!g.aborted && s.phase != PH_DONE) {
      if (s.events + 8 > a.ev_cap) break;
      const int32_t ph = s.phase;
      if (ph == PH_APPROX_ITER || ph == PH_MPG_ITER) {
        if (spent >= a.budget) break;
        spent++;
      }
      switch (ph) {
        case PH_APPROX_INIT: approx_init(); break;
        case PH_APPROX_ITER: approx_iter(); break;
        case PH_APPROX_FINISH: approx_finish(); break;
        case PH_MPG_INIT: mpg_init(); break;
        case PH_MPG_ITER: mpg_iter(); break;
        case PH_MPG_FINISH: mpg_finish(); break;
        case PH_ALM_START: alm_start(); break;
        case PH_ALM_BEGIN: alm_begin(); break;
        case PH_ALM_END: alm_end(); break;
        case PH_UPDATE_ONLY: update_only(); break;
        default: fail(ACD_ERR_LAUNCH); break;
      }
    }
    if (g.aborted && s.error == 0) s.error = ACD_ERR_BARRIER;
  }
};

}  // namespace acd
