// dev_spmv.hip -- y = A x and y = A' x on the sliced-ELL operators, plus
// the vector gather used by the permutation entry points.
//
// Replaces mult_subblock / mult2_subblock (spmv/spmv_block.c:160-254) and
// the thread_pool_for dispatch of sparse_matrix_multiply(_2)
// (spmv/spmv.c:159-233) of pkhuong/acd-alm.  One lane owns one row and
// adds value*x[column] over its row in ascending column order, starting
// from +0.0, with a separate multiply and add (the build disables FP
// contraction): the reference's per-row sum bit for bit.  Value and index
// streams are read as coalesced 64-lane columns of a 64-row slice.
//
// Because the per-row sum is sequential, throughput comes from memory-
// level parallelism: a lane loads U columns of its slice (values and
// indices) at once, then issues the U gathers, then adds in order; and a
// wavefront can interleave R slices so R independent rows are in flight
// per lane.  acd_dev_spmv_variant() exposes the (U, R) choices for tuning;
// the default is the measured best on MI355X.
#include <hip/hip_runtime.h>
#include "acd_dev.h"

namespace {

inline hipStream_t S() { return (hipStream_t)acd_dev_stream(); }

constexpr int kWaves = 4;  // wavefronts per workgroup

template <int U, bool NT>
__device__ __forceinline__ double row_dot(const double* __restrict__ v,
                                          const uint32_t* __restrict__ c,
                                          uint32_t len,
                                          const double* __restrict__ x) {
  double acc = 0.0;
  uint32_t k = 0;
  for (; k + U <= len; k += U, v += 64 * U, c += 64 * U) {
    double vv[U];
    uint32_t cc[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      vv[u] = NT ? __builtin_nontemporal_load(v + 64 * u) : v[64 * u];
      cc[u] = NT ? __builtin_nontemporal_load(c + 64 * u) : c[64 * u];
    }
    double xx[U];
#pragma unroll
    for (int u = 0; u < U; u++) xx[u] = x[cc[u]];
#pragma unroll
    for (int u = 0; u < U; u++) acc = acc + vv[u] * xx[u];
  }
  for (; k < len; k++, v += 64, c += 64) acc = acc + (*v) * x[*c];
  return acc;
}

// Workgroups are dealt round-robin to the 8 XCDs (block b runs on XCD
// b mod 8).  With XCD = true, block b takes the tile of block
// (b mod 8) * per + b / 8 (per = blocks per XCD; the grid is a multiple of
// 8), so each XCD works on one contiguous eighth of the rows: where row
// blocks touch their own column blocks (the multicommodity flow LP's
// commodities), each XCD's gathers then stay within a slice of x that
// its own 4 MB L2 holds.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t tile_of_block(bool xcd) {
  if (!xcd) return blockIdx.x;
  const uint32_t per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// R slices per wavefront, interleaved chunk by chunk so R independent
// row sums are in flight per lane.
template <int U, int R, bool NT, bool XCD = false>
__global__ void __launch_bounds__(64 * kWaves)
k_spmv1(const uint64_t* __restrict__ off, const uint32_t* __restrict__ col,
        const double* __restrict__ val, uint32_t nrows, uint32_t nslices,
        const double* __restrict__ x, double* __restrict__ y) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t s0 = (tile_of_block(XCD) * kWaves + threadIdx.x / 64) * R;
  if (R == 1) {
    if (s0 >= nslices) return;
    const uint64_t base = off[s0];
    const uint32_t len = (uint32_t)((off[s0 + 1] - base) >> 6);
    const double a = row_dot<U, NT>(val + base + lane, col + base + lane, len, x);
    const uint32_t r = s0 * 64 + lane;
    if (r < nrows) y[r] = a;
    return;
  }
  const double* v[R];
  const uint32_t* c[R];
  uint32_t len[R];
  double acc[R];
  uint32_t maxlen = 0;
#pragma unroll
  for (int q = 0; q < R; q++) {
    const uint32_t s = s0 + q;
    uint64_t base = 0;
    len[q] = 0;
    if (s < nslices) {
      base = off[s];
      len[q] = (uint32_t)((off[s + 1] - base) >> 6);
    }
    v[q] = val + base + lane;
    c[q] = col + base + lane;
    acc[q] = 0.0;
    maxlen = len[q] > maxlen ? len[q] : maxlen;
  }
  for (uint32_t k = 0; k < maxlen; k += U) {
    double vv[R][U], xx[R][U];
    uint32_t cc[R][U];
#pragma unroll
    for (int q = 0; q < R; q++)
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool in = k + u < len[q];
        vv[q][u] = in ? (NT ? __builtin_nontemporal_load(v[q] + 64 * u) : v[q][64 * u]) : 0.0;
        cc[q][u] = in ? (NT ? __builtin_nontemporal_load(c[q] + 64 * u) : c[q][64 * u]) : 0u;
      }
#pragma unroll
    for (int q = 0; q < R; q++)
#pragma unroll
      for (int u = 0; u < U; u++) xx[q][u] = k + u < len[q] ? x[cc[q][u]] : 0.0;
#pragma unroll
    for (int q = 0; q < R; q++) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (k + u < len[q]) acc[q] = acc[q] + vv[q][u] * xx[q][u];
      v[q] += 64 * U;
      c[q] += 64 * U;
    }
  }
#pragma unroll
  for (int q = 0; q < R; q++) {
    const uint32_t r = (s0 + q) * 64 + lane;
    if (s0 + q < nslices && r < nrows) y[r] = acc[q];
  }
}

template <int U>
__global__ void __launch_bounds__(64 * kWaves)
k_spmv2(const uint64_t* __restrict__ off, const uint32_t* __restrict__ col,
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
  for (; k + U <= len; k += U, v += 64 * U, c += 64 * U) {
    double vv[U], p[U], q[U];
    uint32_t cc[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      vv[u] = __builtin_nontemporal_load(v + 64 * u);
      cc[u] = __builtin_nontemporal_load(c + 64 * u);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      p[u] = x0[cc[u]];
      q[u] = x1[cc[u]];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      a0 = a0 + vv[u] * p[u];
      a1 = a1 + vv[u] * q[u];
    }
  }
  for (; k < len; k++, v += 64, c += 64) {
    const double vv = *v;
    const uint32_t cc = *c;
    a0 = a0 + vv * x0[cc];
    a1 = a1 + vv * x1[cc];
  }
  const uint32_t r = s * 64 + lane;
  if (r < nrows) {
    y0[r] = a0;
    y1[r] = a1;
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

int check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    acd_dev_record_error((int)e, what);
    return (int)e;
  }
  return 0;
}

template <int U, int R, bool NT, bool XCD = false>
int launch1_t(const struct acd_sell* op, double* y, const double* x) {
  const uint32_t waves = (op->nslices + R - 1) / R;
  unsigned grid = (waves + kWaves - 1) / kWaves;
  if (XCD) grid = (grid + kXcds - 1) / kXcds * kXcds;  // tiles past the
                                                        // end return
  k_spmv1<U, R, NT, XCD><<<grid, 64 * kWaves, 0, S()>>>(
      op->off, op->col, op->val, op->nrows, op->nslices, x, y);
  return check("spmv launch");
}

// Variant table: (unroll, slices per wave, nontemporal loads, XCD-contiguous
// tiles).
int launch1(const struct acd_sell* op, double* y, const double* x, int variant) {
  if (op->nslices == 0) return 0;
  switch (variant) {
    case 0: return launch1_t<4, 1, true>(op, y, x);
    case 1: return launch1_t<4, 1, false>(op, y, x);
    case 2: return launch1_t<8, 1, true>(op, y, x);
    case 3: return launch1_t<8, 1, false>(op, y, x);
    case 4: return launch1_t<4, 2, true>(op, y, x);
    case 5: return launch1_t<8, 2, true>(op, y, x);
    case 6: return launch1_t<4, 4, true>(op, y, x);
    case 7: return launch1_t<16, 1, true>(op, y, x);
    case 8: return launch1_t<2, 4, true>(op, y, x);
    case 9: return launch1_t<4, 1, false, true>(op, y, x);
    case 10: return launch1_t<8, 1, false, true>(op, y, x);
    case 11: return launch1_t<8, 1, true, true>(op, y, x);
    case 12: return launch1_t<4, 2, true, true>(op, y, x);
    default: return launch1_t<4, 1, true>(op, y, x);
  }
}

// Plain loads, 4-way unrolled: the best or within 5% of the best variant on
// the set-cover and multicommodity operators measured on MI355X.
int g_default_variant = 1;

// Staging memory of the allocation-free product (acd_dev_spmv_noalloc):
// x in the first ACD_NOALLOC_MAX doubles, y in the second.  Each device
// holds its own copy from the library's load until the process ends.  In
// host code `scratch` names a host-side stand-in, never device memory: its
// device address is looked up per device (scratch_ptr) and only that
// address is passed to copies and kernels.
static __device__ __align__(256) double scratch[2 * (size_t)ACD_NOALLOC_MAX];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices];

// Device address of this device's `scratch`, looked up once per device
// with that device current (the library's device).
double* scratch_ptr() {
  int dev = acd_dev_device();
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  if (!g_scratch[dev]) {
    void* p = nullptr;
    int cur = -1;
    hipGetDevice(&cur);
    if (cur != dev) hipSetDevice(dev);
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev && cur >= 0) hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      acd_dev_record_error((int)e, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    g_scratch[dev] = (double*)p;
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" {

int acd_dev_spmv(const struct acd_dev_matrix* m, int transpose, double* y,
                 const double* x) {
  const struct acd_sell* op = pick(m, transpose);
  return launch1(op, y, x, op->tuned ? op->tuned - 1 : g_default_variant);
}

int acd_dev_spmv2(const struct acd_dev_matrix* m, int transpose, double* y0,
                  double* y1, const double* x0, const double* x1) {
  const struct acd_sell* op = pick(m, transpose);
  if (op->nslices == 0) return 0;
  const unsigned grid = (op->nslices + kWaves - 1) / kWaves;
  k_spmv2<4><<<grid, 64 * kWaves, 0, S()>>>(op->off, op->col, op->val,
                                            op->nrows, op->nslices, x0, x1, y0,
                                            y1);
  return check("spmv2 launch");
}

int acd_dev_gather(double* dest, const double* src, const uint32_t* p,
                   uint32_t n) {
  if (n == 0) return 0;
  k_gather<<<(n + 255) / 256, 256, 0, S()>>>(dest, src, p, n);
  return check("gather launch");
}

// Mean milliseconds per launch of SpMV variant `variant` (internal: used
// to choose the default; variant < 0 selects the default).
int acd_dev_spmv_time_variant(const struct acd_dev_matrix* m, int transpose,
                              double* y, const double* x, int reps,
                              int variant, double* OUT_ms) {
  if (variant < 0) variant = g_default_variant;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    return -1;
  const struct acd_sell* op = pick(m, transpose);
  int rc = launch1(op, y, x, variant);  // warm
  hipEventRecord(a, S());
  for (int r = 0; r < reps && !rc; r++) rc = launch1(op, y, x, variant);
  hipEventRecord(b, S());
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  if (OUT_ms) *OUT_ms = reps > 0 ? ms / reps : 0.0;
  hipEventDestroy(a);
  hipEventDestroy(b);
  return rc;
}

int acd_dev_spmv_time(const struct acd_dev_matrix* m, int transpose, double* y,
                      const double* x, int reps, double* OUT_ms) {
  return acd_dev_spmv_time_variant(m, transpose, y, x, reps, -1, OUT_ms);
}

int acd_dev_spmv_time_rotate(const struct acd_dev_matrix* const* m, size_t k,
                             int transpose, double* const* y,
                             const double* const* x, int reps,
                             double* OUT_ms) {
  hipEvent_t a, b;
  if (k == 0 || hipEventCreate(&a) != hipSuccess ||
      hipEventCreate(&b) != hipSuccess)
    return -1;
  int rc = 0;
  for (size_t i = 0; i < k && !rc; i++) {  // warm: one product each
    const struct acd_sell* op = pick(m[i], transpose);
    rc = launch1(op, y[i], x[i], op->tuned ? op->tuned - 1 : g_default_variant);
  }
  hipEventRecord(a, S());
  for (int r = 0; r < reps && !rc; r++)
    for (size_t i = 0; i < k && !rc; i++) {
      const struct acd_sell* op = pick(m[i], transpose);
      rc = launch1(op, y[i], x[i],
                   op->tuned ? op->tuned - 1 : g_default_variant);
    }
  hipEventRecord(b, S());
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  if (OUT_ms) *OUT_ms = reps > 0 ? ms / (reps * (double)k) : 0.0;
  hipEventDestroy(a);
  hipEventDestroy(b);
  return rc;
}

int acd_dev_spmv_autotune(struct acd_dev_matrix* m, int reps, int* out) {
  if (reps < 1) reps = 1;
  const size_t nx = (size_t)(m->ncols > m->nrows ? m->ncols : m->nrows) + 1;
  double *x = nullptr, *y = nullptr;
  if (hipMalloc((void**)&x, nx * 8) != hipSuccess ||
      hipMalloc((void**)&y, nx * 8) != hipSuccess) {
    hipFree(x);
    acd_dev_record_error(-1, "autotune buffers");
    return -1;
  }
  hipMemsetAsync(x, 0, nx * 8, S());
  int rc = 0;
  for (int tr = 0; tr < 2 && !rc; tr++) {
    struct acd_sell* op = tr ? &m->at : &m->a;
    int best = g_default_variant;
    double best_ms = 0;
    for (int v = 0; v < ACD_DEV_SPMV_VARIANTS && !rc; v++) {
      double ms = 0;
      if (acd_dev_spmv_time_variant(m, tr, y, x, reps, v, &ms)) {
        rc = -1;
        break;
      }
      if (v == 0 || ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    op->tuned = best + 1;
    if (out) out[tr] = best;
  }
  hipStreamSynchronize(S());
  hipFree(x);
  hipFree(y);
  return rc;
}

int acd_dev_scratch_setup(void) { return scratch_ptr() ? 0 : -1; }

int acd_dev_spmv_noalloc(const struct acd_dev_matrix* m, int transpose,
                         double* y, size_t ny, const double* x, size_t nx) {
  const struct acd_sell* op = pick(m, transpose);
  if (nx > ACD_NOALLOC_MAX || ny > ACD_NOALLOC_MAX) {
    acd_dev_record_error(-1, "allocation-free product: vector longer than "
                             "the 4,194,304-element staging maximum");
    return -1;
  }
  if (ny != op->nrows || nx != op->ncols) {
    acd_dev_record_error(-1, "allocation-free product: length mismatch");
    return -1;
  }
  if (ny == 0) return 0;
  double* sx = scratch_ptr();
  if (!sx) return -1;
  double* sy = sx + ACD_NOALLOC_MAX;
  // Every y[0..ny) is written by the product (each row of every slice,
  // empty rows included) and x is fully uploaded, so nothing left in the
  // staging memory by an earlier call is read.
  if (acd_dev_h2d(sx, x, nx * sizeof(double))) return -1;
  if (launch1(op, sy, sx, op->tuned ? op->tuned - 1 : g_default_variant))
    return -1;
  return acd_dev_d2h(y, sy, ny * sizeof(double));
}

int acd_dev_spmv_set_default_variant(int variant) {
  g_default_variant = variant;
  return 0;
}

}  // extern "C"
