// dev_cd.hip -- device side of cd_solve, the randomised coordinate descent
// that pkhuong/acd-alm ships as its ALM inner solver (approx/cd.c:27-385).
//
// Every value is computed in the reference's order so the trajectory is
// bit-identical to the serial C code:
//   - column set-up (cd.c:164-204): norm_i = sum over the rows r of column
//     i, ascending, of (w_r a_ri) a_ri; q_i = A' (w o a_i) summed per output
//     column j over ascending r (the terms with a zero factor that the
//     reference also adds are +-0 and change nothing), then entries with
//     |q| < 1e-10 dropped.  The products are generated as (i, j) keyed
//     triples in ascending r, stably radix-sorted by (i, j) and summed
//     sequentially per key;
//   - gradient and value (cd.c:80-109) use the bit-exact SpMV and a single
//     wavefront that adds the per-row terms in index order;
//   - the coordinate steps (cd.c:252-282) run in one workgroup, one step at
//     a time, in the sampled order; the 256 threads split the update of g
//     along q_c, whose indices are distinct.
// The host draws the sampled column order with the process's random(),
// exactly as the reference does, and drives the major steps.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "acd_dev.h"

namespace {

#define CK(call, what)                        \
  do {                                        \
    hipError_t e_ = (call);                   \
    if (e_ != hipSuccess) {                   \
      acd_dev_record_error((int)e_, what);    \
      return -1;                              \
    }                                         \
  } while (0)

inline hipStream_t S() { return (hipStream_t)acd_dev_stream(); }

inline unsigned blocks_for(uint64_t n, unsigned bs) {
  uint64_t b = (n + bs - 1) / bs;
  return (unsigned)(b ? b : 1);
}

constexpr uint32_t kLdsG = 8000;  // g kept in LDS (64,000 B) up to this n

// Entry k of row r of a sliced-ELL operator.
__device__ __forceinline__ uint64_t sell_entry(const uint64_t* off, uint32_t r,
                                               uint32_t k) {
  return off[r >> 6] + (uint64_t)k * 64 + (r & 63);
}

// Per column i: the number of (i, j, r) product triples, the column norm
// sum_r (w_r a_ri) a_ri in ascending r, its probability norm^2 + 1e-2 and
// 1 / norm (cd.c:177-200).
__global__ void k_cd_columns(acd_sell at, const uint32_t* rowlen_a,
                             const double* w, uint32_t n, uint64_t* cnt,
                             double* prob, double* inv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t len = at.rowlen[i];
  uint64_t c = 0;
  double norm = 0;
  for (uint32_t k = 0; k < len; k++) {
    const uint64_t e = sell_entry(at.off, i, k);
    const uint32_t r = at.col[e];
    const double d = at.val[e];
    c += rowlen_a[r];
    norm = norm + (w[r] * d) * d;
  }
  cnt[i] = c;
  prob[i] = (norm * norm) + 1e-2;
  inv[i] = 1.0 / norm;
}

// The triples of column i in ascending r, then ascending j within a row.
__global__ void k_cd_fill(acd_sell at, acd_sell a, const double* w, uint32_t n,
                          const uint64_t* start, uint64_t* key, double* term) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t t = start[i];
  const uint32_t len = at.rowlen[i];
  for (uint32_t k = 0; k < len; k++) {
    const uint64_t e = sell_entry(at.off, i, k);
    const uint32_t r = at.col[e];
    const double dw = w[r] * at.val[e];
    const uint32_t lr = a.rowlen[r];
    for (uint32_t kk = 0; kk < lr; kk++) {
      const uint64_t e2 = sell_entry(a.off, r, kk);
      key[t] = ((uint64_t)i << 32) | a.col[e2];
      term[t] = a.val[e2] * dw;
      t++;
    }
  }
}

__global__ void k_cd_heads(const uint64_t* key, uint64_t nt, uint32_t* head) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt) head[t] = (t == 0 || key[t] != key[t - 1]) ? 1u : 0u;
}

// One thread per run of equal keys: the sequential sum from +0.0, kept
// unless |q| < 1e-10 (cd.c:128-146).
__global__ void k_cd_segsum(const uint64_t* key, const double* term,
                            uint64_t nt, const uint32_t* head, double* qv,
                            uint32_t* keep) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  if (!head[t]) {
    keep[t] = 0;
    return;
  }
  double acc = 0.0;
  uint64_t u = t;
  do {
    acc = acc + term[u];
    u++;
  } while (u < nt && key[u] == key[t]);
  qv[t] = acc;
  keep[t] = !(fabs(acc) < 1e-10) ? 1u : 0u;
}

__global__ void k_cd_compact(const uint64_t* key, const double* qv,
                             const uint32_t* keep, const uint32_t* pos,
                             uint64_t nt, uint32_t* qcol, uint32_t* qidx,
                             double* qval) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt || !keep[t]) return;
  const uint32_t p = pos[t];
  qcol[p] = (uint32_t)(key[t] >> 32);
  qidx[p] = (uint32_t)key[t];
  qval[p] = qv[t];
}

// qoff[c] = first kept entry of column >= c (binary search).
__global__ void k_cd_qoff(const uint32_t* qcol, uint32_t nq, uint32_t n,
                          uint32_t* qoff) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > n) return;
  uint32_t lo = 0, hi = nq;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (qcol[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  qoff[c] = lo;
}

// Smallest 1/norm over the columns, folded with fmin in index order
// (cd.c:230-236); one wavefront.
__global__ void k_cd_min_inv(const double* inv, uint32_t n, double* out) {
  double acc = HUGE_VAL;
  for (uint32_t b = 0; b < n; b += 64) {
    const uint32_t i = b + threadIdx.x;
    const double v = i < n ? inv[i] : HUGE_VAL;
    for (int l = 0; l < 64 && b + l < n; l++) acc = fmin(acc, __shfl(v, l, 64));
  }
  if (threadIdx.x == 0) *out = acc;
}

// x = fmin(fmax(l, x), u) (cd.c:42-52).
__global__ void k_cd_clamp(double* x, const double* lo, const double* hi,
                           uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = fmin(fmax(lo[i], x[i]), hi[i]);
}

// viol <- A x - b, scaled <- w viol (cd.c:87-94).
__global__ void k_cd_scale(double* viol, const double* rhs, const double* w,
                           double* scaled, uint32_t m) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double v = viol[i] - rhs[i];
  viol[i] = v;
  scaled[i] = w[i] * v;
}

// value = .5 sum_i (w v) v + sum_j l_j x_j with g_j += l_j (cd.c:89-106);
// one wavefront adds the terms in index order.
__global__ void k_cd_value(const double* viol, const double* w, uint32_t m,
                           double* g, const double* lin, const double* x,
                           uint32_t n, double* out) {
  const uint32_t lane = threadIdx.x;
  double acc = 0;
  for (uint32_t b = 0; b < m; b += 64) {
    const uint32_t i = b + lane;
    double t = 0;
    if (i < m) {
      const double v = viol[i];
      t = w[i] * v * v;
    }
    for (int l = 0; l < 64 && b + l < m; l++) acc = acc + __shfl(t, l, 64);
  }
  acc = acc * .5;
  for (uint32_t b = 0; b < n; b += 64) {
    const uint32_t j = b + lane;
    double t = 0;
    if (j < n) {
      const double li = lin[j];
      g[j] = g[j] + li;
      t = li * x[j];
    }
    for (int l = 0; l < 64 && b + l < n; l++) acc = acc + __shfl(t, l, 64);
  }
  if (lane == 0) *out = acc;
}

// out[0] = ||clamp(x - g) - x||_2 (cd.c:64-78), out[1] = ||g||_2,
// out[2] = ||viol||_2 (cd.c:54-62), each summed in index order.
__global__ void k_cd_norms(const double* g, const double* x, const double* lo,
                           const double* hi, uint32_t n, const double* viol,
                           uint32_t m, double* out) {
  const uint32_t lane = threadIdx.x;
  double pg = 0, ng = 0, nv = 0;
  for (uint32_t b = 0; b < n; b += 64) {
    const uint32_t j = b + lane;
    double d2 = 0, g2 = 0;
    if (j < n) {
      const double xi = x[j], gi = g[j];
      double xp = xi - gi;
      if (xp < lo[j]) xp = lo[j];
      if (xp > hi[j]) xp = hi[j];
      const double d = xp - xi;
      d2 = d * d;
      g2 = gi * gi;
    }
    for (int l = 0; l < 64 && b + l < n; l++) {
      pg = pg + __shfl(d2, l, 64);
      ng = ng + __shfl(g2, l, 64);
    }
  }
  for (uint32_t b = 0; b < m; b += 64) {
    const uint32_t i = b + lane;
    const double v = i < m ? viol[i] : 0.0;
    const double v2 = v * v;
    for (int l = 0; l < 64 && b + l < m; l++) nv = nv + __shfl(v2, l, 64);
  }
  if (lane == 0) {
    out[0] = sqrt(pg);
    out[1] = sqrt(ng);
    out[2] = sqrt(nv);
  }
}

// The sampled coordinate steps (cd.c:252-282, :318-320), in order.  The
// constants of the next step are loaded while the current one runs.
template <bool kLds>
__global__ void __launch_bounds__(256)
k_cd_steps(const uint32_t* cols, uint32_t ncols, double scale,
           const double* inv_p, const double* lo, const double* hi,
           const uint32_t* qoff, const uint32_t* qidx, const double* qval,
           double* g_glob, double* x, uint32_t n) {
  extern __shared__ double g_lds[];
  double* G = kLds ? g_lds : g_glob;
  const uint32_t tid = threadIdx.x;
  if (kLds) {
    for (uint32_t j = tid; j < n; j += 256) g_lds[j] = g_glob[j];
    __syncthreads();
  }
  const double inv = *inv_p;
  uint32_t c = ncols ? cols[0] : 0;
  double l = 0, h = 0;
  uint32_t q0 = 0, q1 = 0;
  if (ncols) {
    l = lo[c];
    h = hi[c];
    q0 = qoff[c];
    q1 = qoff[c + 1];
  }
  for (uint32_t s = 0; s < ncols; s++) {
    // this step's share of q_c, and the next step's constants
    const uint32_t k = q0 + tid;
    uint32_t qi = 0;
    double qv = 0;
    if (k < q1) {
      qi = qidx[k];
      qv = qval[k];
    }
    uint32_t cn = c, q0n = 0, q1n = 0;
    double ln = 0, hn = 0;
    if (s + 1 < ncols) {
      cn = cols[s + 1];
      ln = lo[cn];
      hn = hi[cn];
      q0n = qoff[cn];
      q1n = qoff[cn + 1];
    }
    const double xi = x[c], gi = G[c];
    bool moved = false;
    double step = 0, xp = xi;
    if (gi != 0) {
      xp = xi - scale * (gi * inv);
      if (xp < l) xp = l;
      if (xp > h) xp = h;
      step = xp - xi;
      moved = step != 0;
    }
    if (moved) {
      __syncthreads();  // every thread has read x[c] and g[c]
      if (tid == 0) x[c] = xp;
      if (k < q1) G[qi] = G[qi] + step * qv;
      for (uint32_t kk = k + 256; kk < q1; kk += 256)
        G[qidx[kk]] = G[qidx[kk]] + step * qval[kk];
      __syncthreads();
    }
    c = cn;
    l = ln;
    h = hn;
    q0 = q0n;
    q1 = q1n;
  }
  if (kLds) {
    __syncthreads();
    for (uint32_t j = tid; j < n; j += 256) g_glob[j] = g_lds[j];
  }
}

}  // namespace

struct acd_dev_cd {
  const acd_dev_matrix* a;
  struct acd_dev_lp* lp;
  uint32_t m, n;
  double *g, *viol, *scaled, *prob, *inv, *inv_min, *out;
  uint32_t *qoff, *qidx, *cols;
  double* qval;
  uint32_t nq, cols_cap;
};

extern "C" {

struct acd_dev_cd* acd_dev_cd_create(const struct acd_dev_matrix* a,
                                     struct acd_dev_lp* lp) {
  if (acd_dev_init()) return nullptr;
  acd_dev_cd* cd = (acd_dev_cd*)calloc(1, sizeof(acd_dev_cd));
  if (!cd) return nullptr;
  cd->a = a;
  cd->lp = lp;
  cd->m = a->nrows;
  cd->n = a->ncols;
  const size_t m = cd->m + 1, n = cd->n + 1;
  cd->g = (double*)acd_dev_malloc(n * 8);
  cd->viol = (double*)acd_dev_malloc(m * 8);
  cd->scaled = (double*)acd_dev_malloc(m * 8);
  cd->prob = (double*)acd_dev_malloc(n * 8);
  cd->inv = (double*)acd_dev_malloc(n * 8);
  cd->inv_min = (double*)acd_dev_malloc(8);
  cd->out = (double*)acd_dev_malloc(4 * 8);
  cd->qoff = (uint32_t*)acd_dev_malloc((n + 1) * 4);
  if (!cd->g || !cd->viol || !cd->scaled || !cd->prob || !cd->inv ||
      !cd->inv_min || !cd->out || !cd->qoff) {
    acd_dev_cd_free(cd);
    return nullptr;
  }
  return cd;
}

void acd_dev_cd_free(struct acd_dev_cd* cd) {
  if (!cd) return;
  acd_dev_sync();
  acd_dev_free(cd->g);
  acd_dev_free(cd->viol);
  acd_dev_free(cd->scaled);
  acd_dev_free(cd->prob);
  acd_dev_free(cd->inv);
  acd_dev_free(cd->inv_min);
  acd_dev_free(cd->out);
  acd_dev_free(cd->qoff);
  acd_dev_free(cd->qidx);
  acd_dev_free(cd->qval);
  acd_dev_free(cd->cols);
  free(cd);
}

int acd_dev_cd_setup(struct acd_dev_cd* cd, double* h_prob) {
  const uint32_t n = cd->n;
  const acd_sell& at = cd->a->at;
  const acd_sell& a = cd->a->a;
  const double* w = acd_dev_lp_vec(cd->lp, ACD_V_WEIGHT);
  uint64_t* cnt = nullptr;
  uint64_t* start = nullptr;
  CK(hipMalloc((void**)&cnt, (size_t)(n + 1) * 8), "cd counts");
  CK(hipMalloc((void**)&start, (size_t)(n + 1) * 8), "cd starts");
  int rc = 0;
  void* tmp = nullptr;
  size_t need = 0;
  uint64_t nt = 0;
  uint64_t *key = nullptr, *key2 = nullptr;
  double *term = nullptr, *term2 = nullptr, *qv = nullptr;
  uint32_t *head = nullptr, *keep = nullptr, *pos = nullptr, *qcol = nullptr;
  do {
    hipMemsetAsync(cnt, 0, (size_t)(n + 1) * 8, S());
    if (n)
      k_cd_columns<<<blocks_for(n, 256), 256, 0, S()>>>(
          at, a.rowlen, w, n, cnt, cd->prob, cd->inv);
    k_cd_min_inv<<<1, 64, 0, S()>>>(cd->inv, n, cd->inv_min);
    if (rocprim::exclusive_scan(nullptr, need, cnt, start, (uint64_t)0,
                                (size_t)n + 1, rocprim::plus<uint64_t>(),
                                S()) != hipSuccess ||
        hipMalloc(&tmp, need ? need : 1) != hipSuccess ||
        rocprim::exclusive_scan(tmp, need, cnt, start, (uint64_t)0,
                                (size_t)n + 1, rocprim::plus<uint64_t>(),
                                S()) != hipSuccess) {
      rc = -1;
      break;
    }
    hipFree(tmp);
    tmp = nullptr;
    if (hipMemcpyAsync(&nt, start + n, 8, hipMemcpyDeviceToHost, S()) !=
            hipSuccess ||
        hipStreamSynchronize(S()) != hipSuccess) {
      rc = -1;
      break;
    }
    if (nt >= 0xffffffffull) {
      acd_dev_record_error(-1, "cd_solve: too many column products");
      rc = -1;
      break;
    }
    const size_t nb = (size_t)(nt ? nt : 1);
    if (hipMalloc((void**)&key, nb * 8) || hipMalloc((void**)&key2, nb * 8) ||
        hipMalloc((void**)&term, nb * 8) ||
        hipMalloc((void**)&term2, nb * 8) || hipMalloc((void**)&qv, nb * 8) ||
        hipMalloc((void**)&head, nb * 4) ||
        hipMalloc((void**)&keep, (nb + 1) * 4) ||
        hipMalloc((void**)&pos, (nb + 1) * 4)) {
      acd_dev_record_error(-1, "cd_solve: hipMalloc");
      rc = -1;
      break;
    }
    uint32_t nq = 0;
    if (nt) {
      k_cd_fill<<<blocks_for(n, 256), 256, 0, S()>>>(at, a, w, n, start, key,
                                                    term);
      unsigned bits = 32, nn = n ? n - 1 : 0;
      while (nn) {
        bits++;
        nn >>= 1;
      }
      need = 0;
      if (rocprim::radix_sort_pairs(nullptr, need, key, key2, term, term2,
                                    (size_t)nt, 0u, bits, S()) != hipSuccess ||
          hipMalloc(&tmp, need ? need : 1) != hipSuccess ||
          rocprim::radix_sort_pairs(tmp, need, key, key2, term, term2,
                                    (size_t)nt, 0u, bits, S()) != hipSuccess) {
        rc = -1;
        break;
      }
      hipFree(tmp);
      tmp = nullptr;
      k_cd_heads<<<blocks_for(nt, 256), 256, 0, S()>>>(key2, nt, head);
      k_cd_segsum<<<blocks_for(nt, 256), 256, 0, S()>>>(key2, term2, nt, head,
                                                       qv, keep);
      need = 0;
      if (rocprim::exclusive_scan(nullptr, need, keep, pos, 0u, (size_t)nt + 1,
                                  rocprim::plus<uint32_t>(), S()) !=
              hipSuccess ||
          hipMalloc(&tmp, need ? need : 1) != hipSuccess) {
        rc = -1;
        break;
      }
      // keep[nt] is read by the scan of nt + 1 elements: the last count.
      hipMemsetAsync(keep + nt, 0, 4, S());
      if (rocprim::exclusive_scan(tmp, need, keep, pos, 0u, (size_t)nt + 1,
                                  rocprim::plus<uint32_t>(), S()) !=
          hipSuccess) {
        rc = -1;
        break;
      }
      hipFree(tmp);
      tmp = nullptr;
      if (hipMemcpyAsync(&nq, pos + nt, 4, hipMemcpyDeviceToHost, S()) !=
              hipSuccess ||
          hipStreamSynchronize(S()) != hipSuccess) {
        rc = -1;
        break;
      }
    }
    acd_dev_free(cd->qidx);
    acd_dev_free(cd->qval);
    cd->qidx = (uint32_t*)acd_dev_malloc((size_t)(nq ? nq : 1) * 4);
    cd->qval = (double*)acd_dev_malloc((size_t)(nq ? nq : 1) * 8);
    if (hipMalloc((void**)&qcol, (size_t)(nq ? nq : 1) * 4) != hipSuccess ||
        !cd->qidx || !cd->qval) {
      rc = -1;
      break;
    }
    cd->nq = nq;
    if (nt)
      k_cd_compact<<<blocks_for(nt, 256), 256, 0, S()>>>(
          key2, qv, keep, pos, nt, qcol, cd->qidx, cd->qval);
    k_cd_qoff<<<blocks_for((uint64_t)n + 1, 256), 256, 0, S()>>>(qcol, nq, n,
                                                                 cd->qoff);
    if (hipMemcpyAsync(h_prob, cd->prob, (size_t)n * 8, hipMemcpyDeviceToHost,
                       S()) != hipSuccess ||
        hipStreamSynchronize(S()) != hipSuccess)
      rc = -1;
  } while (0);
  if (tmp) hipFree(tmp);
  hipFree(cnt);
  hipFree(start);
  hipFree(key);
  hipFree(key2);
  hipFree(term);
  hipFree(term2);
  hipFree(qv);
  hipFree(head);
  hipFree(keep);
  hipFree(pos);
  hipFree(qcol);
  if (rc) acd_dev_record_error(rc, "cd_solve set-up");
  return rc;
}

int acd_dev_cd_clamp(struct acd_dev_cd* cd) {
  const uint32_t n = cd->n;
  if (n)
    k_cd_clamp<<<blocks_for(n, 256), 256, 0, S()>>>(
        acd_dev_lp_vec(cd->lp, ACD_V_X), acd_dev_lp_vec(cd->lp, ACD_V_LOWER),
        acd_dev_lp_vec(cd->lp, ACD_V_UPPER), n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int acd_dev_cd_gradient(struct acd_dev_cd* cd, double* value) {
  double* x = acd_dev_lp_vec(cd->lp, ACD_V_X);
  const double* rhs = acd_dev_lp_vec(cd->lp, ACD_V_RHS);
  const double* w = acd_dev_lp_vec(cd->lp, ACD_V_WEIGHT);
  const double* lin = acd_dev_lp_vec(cd->lp, ACD_V_LINEAR);
  if (acd_dev_spmv(cd->a, 0, cd->viol, x)) return -1;
  if (cd->m)
    k_cd_scale<<<blocks_for(cd->m, 256), 256, 0, S()>>>(cd->viol, rhs, w,
                                                        cd->scaled, cd->m);
  if (acd_dev_spmv(cd->a, 1, cd->g, cd->scaled)) return -1;
  k_cd_value<<<1, 64, 0, S()>>>(cd->viol, w, cd->m, cd->g, lin, x, cd->n,
                                cd->out + 3);
  CK(hipMemcpyAsync(value, cd->out + 3, 8, hipMemcpyDeviceToHost, S()),
     "cd value");
  CK(hipStreamSynchronize(S()), "cd gradient");
  return 0;
}

int acd_dev_cd_norms(struct acd_dev_cd* cd, double* pg, double* ng,
                     double* nviol) {
  k_cd_norms<<<1, 64, 0, S()>>>(cd->g, acd_dev_lp_vec(cd->lp, ACD_V_X),
                                acd_dev_lp_vec(cd->lp, ACD_V_LOWER),
                                acd_dev_lp_vec(cd->lp, ACD_V_UPPER), cd->n,
                                cd->viol, cd->m, cd->out);
  double h[3];
  CK(hipMemcpyAsync(h, cd->out, 24, hipMemcpyDeviceToHost, S()), "cd norms");
  CK(hipStreamSynchronize(S()), "cd norms");
  if (pg) *pg = h[0];
  if (ng) *ng = h[1];
  if (nviol) *nviol = h[2];
  return 0;
}

int acd_dev_cd_steps(struct acd_dev_cd* cd, const uint32_t* h_cols,
                     size_t ncols, double scale) {
  if (ncols > cd->cols_cap) {
    acd_dev_free(cd->cols);
    cd->cols = (uint32_t*)acd_dev_malloc(ncols * 4);
    if (!cd->cols) return -1;
    cd->cols_cap = (uint32_t)ncols;
  }
  if (ncols == 0) return 0;
  if (acd_dev_h2d(cd->cols, h_cols, ncols * 4)) return -1;
  double* x = acd_dev_lp_vec(cd->lp, ACD_V_X);
  const double* lo = acd_dev_lp_vec(cd->lp, ACD_V_LOWER);
  const double* hi = acd_dev_lp_vec(cd->lp, ACD_V_UPPER);
  if (cd->n <= kLdsG)
    k_cd_steps<true><<<1, 256, (size_t)cd->n * 8, S()>>>(
        cd->cols, (uint32_t)ncols, scale, cd->inv_min, lo, hi, cd->qoff,
        cd->qidx, cd->qval, cd->g, x, cd->n);
  else
    k_cd_steps<false><<<1, 256, 0, S()>>>(
        cd->cols, (uint32_t)ncols, scale, cd->inv_min, lo, hi, cd->qoff,
        cd->qidx, cd->qval, cd->g, x, cd->n);
  CK(hipGetLastError(), "cd steps launch");
  return 0;
}

}  // extern "C"
