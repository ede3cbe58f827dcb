// dev_spmv.hip -- y = A x and y = A' x on the sliced-ELL operators, plus
// the vector gather used by the permutation entry points.
//
// Replaces mult_subblock / mult2_subblock (spmv/spmv_block.c:160-254) and
// the thread_pool_for dispatch of sparse_matrix_multiply(_2)
// (spmv/spmv.c:159-233) of pkhuong/acd-alm.  One wavefront owns one
// 64-row slice and one lane owns one row: the lane adds value*x[column]
// over its row in ascending column order, starting from +0.0, with a
// separate multiply and add (the build disables FP contraction), which is
// the reference's per-row sum bit for bit.  Value and index streams are
// read as coalesced 64-lane columns; four columns are loaded ahead so each
// lane keeps four gathers in flight.
#include <hip/hip_runtime.h>
#include "acd_dev.h"

namespace {

inline hipStream_t S() { return (hipStream_t)acd_dev_stream(); }

constexpr int kWaves = 4;  // wavefronts per workgroup

template <int NRHS>
__global__ void __launch_bounds__(64 * kWaves)
k_spmv(const uint64_t* __restrict__ off, const uint32_t* __restrict__ col,
       const double* __restrict__ val, uint32_t nrows, uint32_t nslices,
       const double* __restrict__ x0, const double* __restrict__ x1,
       double* __restrict__ y0, double* __restrict__ y1) {
  const uint32_t s = blockIdx.x * kWaves + threadIdx.x / 64;
  const uint32_t lane = threadIdx.x & 63;
  if (s >= nslices) return;
  const uint64_t base = off[s];
  const uint32_t len = (uint32_t)((off[s + 1] - base) >> 6);
  const double* v = val + base + lane;
  const uint32_t* c = col + base + lane;
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
  const uint32_t r = s * 64 + lane;
  if (r < nrows) {
    y0[r] = a0;
    if (NRHS == 2) y1[r] = a1;
  }
}

__global__ void k_gather(double* __restrict__ dst, const double* __restrict__ src,
                         const uint32_t* __restrict__ p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[p[i]];
}

const struct acd_sell* pick(const struct acd_dev_matrix* m, int transpose) {
  return transpose ? &m->at : &m->a;
}

template <int NRHS>
int launch(const struct acd_sell* op, double* y0, double* y1, const double* x0,
           const double* x1) {
  if (op->nslices == 0) return 0;
  unsigned grid = (op->nslices + kWaves - 1) / kWaves;
  k_spmv<NRHS><<<grid, 64 * kWaves, 0, S()>>>(op->off, op->col, op->val,
                                              op->nrows, op->nslices, x0, x1,
                                              y0, y1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    acd_dev_record_error((int)e, "spmv launch");
    return (int)e;
  }
  return 0;
}

}  // namespace

extern "C" {

int acd_dev_spmv(const struct acd_dev_matrix* m, int transpose, double* y,
                 const double* x) {
  return launch<1>(pick(m, transpose), y, nullptr, x, nullptr);
}

int acd_dev_spmv2(const struct acd_dev_matrix* m, int transpose, double* y0,
                  double* y1, const double* x0, const double* x1) {
  return launch<2>(pick(m, transpose), y0, y1, x0, x1);
}

int acd_dev_gather(double* dest, const double* src, const uint32_t* p,
                   uint32_t n) {
  if (n == 0) return 0;
  k_gather<<<(n + 255) / 256, 256, 0, S()>>>(dest, src, p, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    acd_dev_record_error((int)e, "gather launch");
    return (int)e;
  }
  return 0;
}

int acd_dev_spmv_time(const struct acd_dev_matrix* m, int transpose, double* y,
                      const double* x, int reps, double* OUT_ms) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    return -1;
  const struct acd_sell* op = pick(m, transpose);
  int rc = launch<1>(op, y, nullptr, x, nullptr);  // warm
  hipEventRecord(a, S());
  for (int r = 0; r < reps && !rc; r++) rc = launch<1>(op, y, nullptr, x, nullptr);
  hipEventRecord(b, S());
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  if (OUT_ms) *OUT_ms = reps > 0 ? ms / reps : 0.0;
  hipEventDestroy(a);
  hipEventDestroy(b);
  return rc;
}

}  // extern "C"
