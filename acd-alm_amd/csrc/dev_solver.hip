// dev_solver.hip -- the device-resident LP solver behind approx_solve,
// mpg_solve and alm_solve: one cooperative (co-resident) grid per LP
// running the state machine of solver_core.hpp, relaunched in chunks of
// iterations until the solve ends.  Between chunks the host only reads
// the small state record and replays log events.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acd_dev.h"
#include "solver_core.hpp"

using namespace acd;

namespace {

inline hipStream_t S() { return (hipStream_t)acd_dev_stream(); }

__global__ void __launch_bounds__(kBlock) k_solver(Args a) {
  GridGroup g;
  g.init(a.sa);
  Solver<GridGroup> sv(g, a);
  sv.s = *a.st;
  sv.run();
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.st = sv.s;
}

Sell view(const acd_sell& op) {
  Sell v;
  v.off = op.off;
  v.col = op.col;
  v.val = op.val;
  v.nrows = op.nrows;
  v.nslices = op.nslices;
  return v;
}

constexpr uint32_t kEventCap = 4096;

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
  SyncArea* sa;
  acd_event* ev;
  int max_blocks;  // co-resident workgroups of k_solver
};

extern "C" {

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
  lp->st = (State*)acd_dev_malloc(sizeof(State));
  lp->sa = (SyncArea*)acd_dev_malloc(sizeof(SyncArea));
  lp->ev = (acd_event*)acd_dev_malloc(sizeof(acd_event) * kEventCap);
  if (!ok || !lp->st || !lp->sa || !lp->ev) {
    acd_dev_lp_free(lp);
    return nullptr;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_solver, kBlock,
                                                   0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  lp->max_blocks = per_cu * acd_dev_cu_count();
  if (lp->max_blocks > kMaxBlocks) lp->max_blocks = kMaxBlocks;
  return lp;
}

void acd_dev_lp_free(struct acd_dev_lp* lp) {
  if (!lp) return;
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
  acd_dev_free(lp->st);
  acd_dev_free(lp->sa);
  acd_dev_free(lp->ev);
  free(lp);
}

double* acd_dev_lp_vec(struct acd_dev_lp* lp, int which) {
  return (which >= 0 && which < ACD_V_COUNT) ? lp->vec[which] : nullptr;
}

}  // extern "C"

namespace {

Args make_args(acd_dev_lp* lp, uint64_t budget) {
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
  a.sa = lp->sa;
  a.ev = lp->ev;
  a.ev_cap = kEventCap;
  a.budget = budget;
  return a;
}

// Workgroups for this LP: enough waves for the largest operator, within
// what can be co-resident.
int pick_blocks(const acd_dev_lp* lp) {
  uint32_t slices = lp->perm->at.nslices;
  if (lp->perm->a.nslices > slices) slices = lp->perm->a.nslices;
  if (lp->orig->at.nslices > slices) slices = lp->orig->at.nslices;
  if (lp->orig->a.nslices > slices) slices = lp->orig->a.nslices;
  int want = (int)((slices + kWPB - 1) / kWPB);
  if (want < 1) want = 1;
  if (want > lp->max_blocks) want = lp->max_blocks;
  return want;
}

int launch(acd_dev_lp* lp, uint64_t budget, int blocks) {
  Args a = make_args(lp, budget);
  void* params[] = {&a};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_solver, dim3(blocks),
                                            dim3(kBlock), params, 0, S());
  if (e != hipSuccess) {
    acd_dev_record_error((int)e, "cooperative launch of the solver");
    return (int)e;
  }
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
  switch (q->mode) {
    case ACD_MODE_APPROX: s->phase = PH_APPROX_INIT; break;
    case ACD_MODE_MPG: s->phase = PH_MPG_INIT; break;
    case ACD_MODE_ALM:
      s->phase = PH_ALM_START;
      s->outer_niter = q->niter;
      break;
    default: s->phase = PH_UPDATE_ONLY; break;
  }
}

int run_to_completion(acd_dev_lp* lp, State* hs, acd_solve_result* res,
                      acd_event_sink sink, void* ctx, uint64_t budget) {
  const int blocks = pick_blocks(lp);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  acd_event* evbuf = (acd_event*)malloc(sizeof(acd_event) * kEventCap);
  int rc = 0;
  double ms_total = 0;
  uint64_t launches = 0;
  if (hipMemcpyAsync(lp->st, hs, sizeof(State), hipMemcpyHostToDevice, S()) !=
          hipSuccess ||
      hipMemsetAsync(lp->sa, 0, sizeof(uint32_t) * 4, S()) != hipSuccess) {
    rc = -1;
  }
  while (!rc) {
    hipEventRecord(e0, S());
    if ((rc = launch(lp, budget, blocks))) break;
    hipEventRecord(e1, S());
    if (hipMemcpyAsync(hs, lp->st, sizeof(State), hipMemcpyDeviceToHost,
                       S()) != hipSuccess ||
        hipStreamSynchronize(S()) != hipSuccess) {
      acd_dev_record_error(-1, "solver chunk");
      rc = -1;
      break;
    }
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    ms_total += ms;
    launches++;
    if (hs->events) {
      uint64_t ne = hs->events < kEventCap ? hs->events : kEventCap;
      hipMemcpy(evbuf, lp->ev, sizeof(acd_event) * ne, hipMemcpyDeviceToHost);
      if (sink)
        for (uint64_t i = 0; i < ne; i++) sink(ctx, &evbuf[i]);
      hs->events = 0;
      hipMemcpyAsync(&lp->st->events, &hs->events, sizeof(uint64_t),
                     hipMemcpyHostToDevice, S());
    }
    if (hs->error || hs->phase == PH_DONE) break;
  }
  hipStreamSynchronize(S());
  free(evbuf);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (res) {
    res->error = hs->error ? hs->error : (rc ? ACD_ERR_LAUNCH : 0);
    res->launches = launches;
    res->device_ms = ms_total;
    res->inner_iterations = hs->inner_total;
    res->outer_iterations = hs->outer;
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

}  // namespace

extern "C" {

int acd_dev_approx_update(struct acd_dev_lp* lp) {
  acd_solve_request q;
  memset(&q, 0, sizeof q);
  q.mode = 0;
  State hs;
  init_state(&hs, lp, &q);
  acd_solve_result res;
  return run_to_completion(lp, &hs, &res, nullptr, nullptr, 1);
}

int acd_dev_solve(struct acd_dev_lp* lp, const struct acd_solve_request* q,
                  struct acd_solve_result* res, acd_event_sink sink,
                  void* ctx) {
  State hs;
  init_state(&hs, lp, q);
  const uint64_t budget = q->log ? 2000 : 20000;
  return run_to_completion(lp, &hs, res, sink, ctx, budget);
}

}  // extern "C"
