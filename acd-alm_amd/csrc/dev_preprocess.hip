// dev_preprocess.hip -- matrix build on the device, bit-identical to the
// host preprocessing of pkhuong/acd-alm:
//
//   * permutation     spmv/permutation.c:63-125  (count per index, then
//                     windows of 256 indices ordered by (count, index))
//   * renumbering     spmv/spmv.c:49-59
//   * CSR order       spmv/spmv_csr.c:27-80  (sort by row<<32 | column)
//   * Z-order         spmv/spmv.swizzle.inc:1-52 (interleaved bits)
//
// Every sort key is unique (no repeated (row, column) pair, which the
// interface requires -- spmv/spmv.h:15-18), so the radix sorts used here
// produce exactly the reference's qsort order.  Instead of the reference's
// 4-row slabs (spmv/spmv_block.c:58-149) the operators are stored as
// sliced ELL with 64 rows per slice: slice s holds len_s = max row length
// columns of 64 entries, entry k of row 64s+l at off[s] + 64k + l, so a
// wavefront reads each column of the slice as one coalesced 512-B (value)
// and 256-B (index) load.  Padding entries carry value 0 and the row's
// last column index (a cache-resident gather).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <stdlib.h>
#include <string.h>
#include "acd_dev.h"

namespace {

#define CK(call, what)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) {                                   \
      acd_dev_record_error((int)e_, what);                    \
      return (int)e_;                                         \
    }                                                         \
  } while (0)

inline hipStream_t S() { return (hipStream_t)acd_dev_stream(); }

inline unsigned blocks_for(uint64_t n, unsigned bs) {
  uint64_t b = (n + bs - 1) / bs;
  return (unsigned)(b ? b : 1);
}

unsigned bits_for(uint64_t maxval) {  // bits to represent values < maxval+1
  unsigned b = 0;
  while (b < 64 && (maxval >> b)) b++;
  return b ? b : 1;
}

__global__ void k_iota(uint32_t* a, uint32_t* b, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i] = i;
}

__global__ void k_histogram(const uint32_t* keys, uint64_t nnz,
                            uint32_t* count) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) atomicAdd(&count[keys[i]], 1u);
}

// One workgroup per window of 256 indices: each index's position is its
// rank under (count, index) within the window.
__global__ void __launch_bounds__(256)
k_window_sort(const uint32_t* count, uint32_t n, uint32_t* idx,
              uint32_t* ridx) {
  __shared__ uint32_t c[256];
  uint32_t base = blockIdx.x * 256u;
  uint32_t t = threadIdx.x;
  uint32_t len = min(256u, n - base);
  if (t < len) c[t] = count[base + t];
  __syncthreads();
  if (t < len) {
    uint32_t mine = c[t], rank = 0;
    for (uint32_t j = 0; j < len; j++) {
      uint32_t o = c[j];
      rank += (o < mine) || (o == mine && j < t);
    }
    idx[base + rank] = base + t;
    ridx[base + t] = base + rank;
  }
}

__global__ void k_remap(uint32_t* keys, uint64_t nnz, const uint32_t* ridx) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) keys[i] = ridx[keys[i]];
}

__global__ void k_major_minor_key(const uint32_t* major, const uint32_t* minor,
                                  uint64_t nnz, uint64_t* key, uint32_t* pos) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) {
    key[i] = ((uint64_t)major[i] << 32) | minor[i];
    pos[i] = (uint32_t)i;
  }
}

__device__ __forceinline__ uint64_t spread_bits(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// Column bits at even positions, row bits at odd positions.
__global__ void k_morton_key(const uint32_t* rows, const uint32_t* cols,
                             uint64_t nnz, uint64_t* key, uint32_t* pos) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) {
    key[i] = spread_bits(cols[i]) | (spread_bits(rows[i]) << 1);
    pos[i] = (uint32_t)i;
  }
}

__global__ void k_gather_u32(uint32_t* dst, const uint32_t* src,
                             const uint32_t* pos, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[pos[i]];
}

__global__ void k_gather_f64(double* dst, const double* src,
                             const uint32_t* pos, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[pos[i]];
}

// Slice length = longest row of the slice; width in entries = 64 * len.
__global__ void k_slice_width(const uint32_t* rowlen, uint32_t nrows,
                              uint32_t nslices, uint64_t* width,
                              uint32_t* maxlen_out) {
  uint32_t s = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  uint32_t lane = threadIdx.x & 63;
  if (s >= nslices) return;
  uint32_t r = s * 64 + lane;
  uint32_t len = r < nrows ? rowlen[r] : 0;
  for (int o = 32; o > 0; o >>= 1) len = max(len, (uint32_t)__shfl_xor(len, o));
  if (lane == 0) {
    width[s] = (uint64_t)len * 64u;
    atomicMax(maxlen_out, len);
  }
}

// Scatter the CSR rows into their slices; pad to the slice length.
__global__ void k_fill_sell(const uint32_t* rowptr, const uint32_t* rowlen,
                            const uint32_t* ccol, const double* cval,
                            uint32_t nrows, const uint64_t* off,
                            uint32_t* col, double* val) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t s = r / 64, lane = r & 63;
  if (r >= ((nrows + 63) / 64) * 64) return;
  uint64_t base = off[s];
  uint32_t len = (uint32_t)((off[s + 1] - base) / 64u);
  uint32_t have = r < nrows ? rowlen[r] : 0;
  uint32_t start = r < nrows ? rowptr[r] : 0;
  uint32_t pad_col = have ? ccol[start + have - 1] : 0;
  for (uint32_t k = 0; k < len; k++) {
    uint64_t e = base + (uint64_t)k * 64u + lane;
    if (k < have) {
      col[e] = ccol[start + k];
      val[e] = cval[start + k];
    } else {
      col[e] = pad_col;
      val[e] = 0.0;
    }
  }
}

__global__ void k_to_f32(const double* src, float* dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// bf16 = the high half of the fp32 bit pattern; flags any value whose
// double -> fp32 -> bf16 truncation is not exact.
__global__ void k_to_bf16(const double* src, uint16_t* dst, uint32_t* inexact,
                          uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = src[i];
  const float f = (float)d;
  const uint32_t bits = __float_as_uint(f);
  dst[i] = (uint16_t)(bits >> 16);
  if ((double)f != d || (bits & 0xffffu)) atomicOr(inexact, 1u);
}

// Value-free entry words (acd_dev_sell_make_sign): flags any value other
// than -1, 0 or +1 and any column of 30 bits or more.
__global__ void k_to_sign(const double* val, const uint32_t* col,
                          uint32_t* dst, uint32_t* inexact, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = val[i];
  const uint32_t c = col[i];
  uint32_t w = c & 0x3fffffffu;
  if (d == -1.0)
    w |= 0x80000000u;
  else if (d == 0.0)
    w |= 0x40000000u;
  else if (d != 1.0)
    atomicOr(inexact, 1u);
  if (c > 0x3fffffffu) atomicOr(inexact, 1u);
  dst[i] = w;
}

__global__ void k_to_u16(const uint32_t* src, uint16_t* dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (uint16_t)src[i];
}

struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return 0;
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
      acd_dev_record_error((int)e, "hipMalloc sort scratch");
      return (int)e;
    }
    bytes = need;
    return 0;
  }
  ~Scratch() {
    if (p) hipFree(p);
  }
};

int sort_pairs(Scratch& tmp, uint64_t* kin, uint64_t* kout, uint32_t* vin,
               uint32_t* vout, uint64_t n, unsigned end_bit) {
  size_t need = 0;
  CK(rocprim::radix_sort_pairs(nullptr, need, kin, kout, vin, vout, (size_t)n,
                               0u, end_bit, S()),
     "radix_sort_pairs size");
  if (tmp.ensure(need)) return -1;
  CK(rocprim::radix_sort_pairs(tmp.p, need, kin, kout, vin, vout, (size_t)n,
                               0u, end_bit, S()),
     "radix_sort_pairs");
  return 0;
}

// Build one sliced-ELL operator from triplets (major = output row).
int build_sell(struct acd_sell* op, uint32_t nmajor, uint32_t nminor,
               const uint32_t* d_major, const uint32_t* d_minor,
               const double* d_vals, uint64_t nnz, Scratch& tmp) {
  memset(op, 0, sizeof *op);
  op->nrows = nmajor;
  op->ncols = nminor;
  op->nslices = (nmajor + 63) / 64;
  const unsigned B = 256;
  uint64_t *key = nullptr, *key2 = nullptr, *width = nullptr;
  uint32_t *pos = nullptr, *pos2 = nullptr, *ccol = nullptr, *rowptr = nullptr,
           *maxlen = nullptr;
  double* cval = nullptr;
  int rc = 0;
  size_t nn = nnz ? nnz : 1;
#define AL(ptr, count)                                                   \
  if (hipMalloc((void**)&ptr, sizeof(*ptr) * (size_t)(count)) != hipSuccess) { \
    acd_dev_record_error(-3, "hipMalloc in matrix build");              \
    rc = -3;                                                             \
    goto done;                                                           \
  }
  AL(key, nn);
  AL(key2, nn);
  AL(pos, nn);
  AL(pos2, nn);
  AL(ccol, nn);
  AL(cval, nn);
  AL(rowptr, (size_t)nmajor + 1);
  AL(op->rowlen, nmajor ? nmajor : 1);
  AL(width, (size_t)op->nslices + 1);
  AL(op->off, (size_t)op->nslices + 1);
  AL(maxlen, 1);
  if (hipMemsetAsync(op->rowlen, 0, sizeof(uint32_t) * (nmajor ? nmajor : 1), S()) ||
      hipMemsetAsync(width, 0, sizeof(uint64_t) * (op->nslices + 1), S()) ||
      hipMemsetAsync(op->off, 0, sizeof(uint64_t) * (op->nslices + 1), S()) ||
      hipMemsetAsync(maxlen, 0, sizeof(uint32_t), S())) {
    rc = -4;
    goto done;
  }
  if (nnz) {
    k_major_minor_key<<<blocks_for(nnz, B), B, 0, S()>>>(d_major, d_minor, nnz,
                                                         key, pos);
    if ((rc = sort_pairs(tmp, key, key2, pos, pos2, nnz,
                         32 + bits_for(nmajor ? nmajor - 1 : 0))))
      goto done;
    k_gather_u32<<<blocks_for(nnz, B), B, 0, S()>>>(ccol, d_minor, pos2, nnz);
    k_gather_f64<<<blocks_for(nnz, B), B, 0, S()>>>(cval, d_vals, pos2, nnz);
    k_histogram<<<blocks_for(nnz, B), B, 0, S()>>>(d_major, nnz, op->rowlen);
  }
  {
    size_t need = 0;
    if (rocprim::exclusive_scan(nullptr, need, op->rowlen, rowptr, 0u,
                                (size_t)nmajor + 1, rocprim::plus<uint32_t>(),
                                S()) != hipSuccess) {
      rc = -5;
      goto done;
    }
    if ((rc = tmp.ensure(need))) goto done;
    // rowlen[nmajor] does not exist: scan nmajor entries, then the total.
    if (nmajor &&
        rocprim::exclusive_scan(tmp.p, need, op->rowlen, rowptr, 0u,
                                (size_t)nmajor, rocprim::plus<uint32_t>(),
                                S()) != hipSuccess) {
      rc = -5;
      goto done;
    }
  }
  if (op->nslices) {
    k_slice_width<<<blocks_for((uint64_t)op->nslices * 64, B), B, 0, S()>>>(
        op->rowlen, nmajor, op->nslices, width, maxlen);
    size_t need = 0;
    if (rocprim::exclusive_scan(nullptr, need, width, op->off, (uint64_t)0,
                                (size_t)op->nslices + 1,
                                rocprim::plus<uint64_t>(), S()) != hipSuccess) {
      rc = -6;
      goto done;
    }
    if ((rc = tmp.ensure(need))) goto done;
    if (rocprim::exclusive_scan(tmp.p, need, width, op->off, (uint64_t)0,
                                (size_t)op->nslices + 1,
                                rocprim::plus<uint64_t>(), S()) != hipSuccess) {
      rc = -6;
      goto done;
    }
  }
  if (hipMemcpyAsync(&op->nent, op->off + op->nslices, sizeof(uint64_t),
                     hipMemcpyDeviceToHost, S()) ||
      hipMemcpyAsync(&op->max_len, maxlen, sizeof(uint32_t),
                     hipMemcpyDeviceToHost, S()) ||
      hipStreamSynchronize(S())) {
    rc = -7;
    goto done;
  }
  AL(op->col, op->nent ? op->nent : 1);
  AL(op->val, op->nent ? op->nent : 1);
  if (op->nslices) {
    // rowptr of the last row + its length is nnz; k_fill_sell only reads
    // rowptr[r] for r < nmajor.
    k_fill_sell<<<blocks_for((uint64_t)op->nslices * 64, B), B, 0, S()>>>(
        rowptr, op->rowlen, ccol, cval, nmajor, op->off, op->col, op->val);
  }
  if (hipStreamSynchronize(S()) != hipSuccess) rc = -8;
done:
#undef AL
  hipFree(key);
  hipFree(key2);
  hipFree(pos);
  hipFree(pos2);
  hipFree(ccol);
  hipFree(cval);
  hipFree(rowptr);
  hipFree(width);
  hipFree(maxlen);
  return rc;
}

void free_sell(struct acd_sell* op) {
  hipFree(op->off);
  hipFree(op->col);
  hipFree(op->val);
  hipFree(op->val32);
  hipFree(op->val16);
  hipFree(op->colsgn);
  hipFree(op->col16);
  hipFree(op->rowlen);
  memset(op, 0, sizeof *op);
}

}  // namespace

extern "C" {

int acd_dev_matrix_build(struct acd_dev_matrix* out, uint32_t nrows,
                         uint32_t ncols, uint64_t nnz, uint32_t* rows,
                         uint32_t* cols, double* vals, int permute,
                         uint32_t* h_row_idx, uint32_t* h_row_ridx,
                         uint32_t* h_col_idx, uint32_t* h_col_ridx) {
  if (acd_dev_init()) return -1;
  memset(out, 0, sizeof *out);
  out->nrows = nrows;
  out->ncols = ncols;
  out->nnz = nnz;
  const unsigned B = 256;
  size_t nn = nnz ? nnz : 1;
  uint32_t *d_rows = nullptr, *d_cols = nullptr, *cnt = nullptr,
           *pos = nullptr, *pos2 = nullptr, *t32 = nullptr;
  uint64_t *key = nullptr, *key2 = nullptr;
  double *d_vals = nullptr, *t64 = nullptr;
  Scratch tmp;
  int rc = 0;
#define AL(ptr, count)                                                   \
  if (hipMalloc((void**)&ptr, sizeof(*ptr) * (size_t)(count)) != hipSuccess) { \
    acd_dev_record_error(-3, "hipMalloc in matrix build");              \
    rc = -3;                                                             \
    goto fail;                                                           \
  }
  AL(d_rows, nn);
  AL(d_cols, nn);
  AL(d_vals, nn);
  AL(out->row_idx, nrows ? nrows : 1);
  AL(out->row_ridx, nrows ? nrows : 1);
  AL(out->col_idx, ncols ? ncols : 1);
  AL(out->col_ridx, ncols ? ncols : 1);
  if (nnz &&
      (hipMemcpyAsync(d_rows, rows, 4 * nnz, hipMemcpyHostToDevice, S()) ||
       hipMemcpyAsync(d_cols, cols, 4 * nnz, hipMemcpyHostToDevice, S()) ||
       hipMemcpyAsync(d_vals, vals, 8 * nnz, hipMemcpyHostToDevice, S()))) {
    rc = -4;
    goto fail;
  }
  if (permute) {
    uint32_t nmax = nrows > ncols ? nrows : ncols;
    AL(cnt, nmax ? nmax : 1);
    for (int which = 0; which < 2; which++) {  // 0: rows, 1: columns
      uint32_t n = which ? ncols : nrows;
      uint32_t* keys = which ? d_cols : d_rows;
      uint32_t* idx = which ? out->col_idx : out->row_idx;
      uint32_t* ridx = which ? out->col_ridx : out->row_ridx;
      if (n == 0) continue;
      hipMemsetAsync(cnt, 0, 4 * (size_t)n, S());
      if (nnz) k_histogram<<<blocks_for(nnz, B), B, 0, S()>>>(keys, nnz, cnt);
      k_window_sort<<<blocks_for(n, 256), 256, 0, S()>>>(cnt, n, idx, ridx);
    }
    if (nnz) {
      k_remap<<<blocks_for(nnz, B), B, 0, S()>>>(d_rows, nnz, out->row_ridx);
      k_remap<<<blocks_for(nnz, B), B, 0, S()>>>(d_cols, nnz, out->col_ridx);
    }
  } else {
    if (nrows)
      k_iota<<<blocks_for(nrows, B), B, 0, S()>>>(out->row_idx, out->row_ridx,
                                                  nrows);
    if (ncols)
      k_iota<<<blocks_for(ncols, B), B, 0, S()>>>(out->col_idx, out->col_ridx,
                                                  ncols);
  }
  if ((rc = build_sell(&out->a, nrows, ncols, d_rows, d_cols, d_vals, nnz, tmp)))
    goto fail;
  if ((rc = build_sell(&out->at, ncols, nrows, d_cols, d_rows, d_vals, nnz, tmp)))
    goto fail;
  // Z-order of the renumbered triplets, returned to the host arrays.
  if (nnz) {
    AL(key, nn);
    AL(key2, nn);
    AL(pos, nn);
    AL(pos2, nn);
    AL(t32, nn);
    AL(t64, nn);
    k_morton_key<<<blocks_for(nnz, B), B, 0, S()>>>(d_rows, d_cols, nnz, key,
                                                    pos);
    uint32_t mx = nrows > ncols ? nrows : ncols;
    if ((rc = sort_pairs(tmp, key, key2, pos, pos2, nnz,
                         2 * bits_for(mx ? mx - 1 : 0))))
      goto fail;
    k_gather_u32<<<blocks_for(nnz, B), B, 0, S()>>>(t32, d_rows, pos2, nnz);
    if (hipMemcpyAsync(rows, t32, 4 * nnz, hipMemcpyDeviceToHost, S()) ||
        hipStreamSynchronize(S())) {
      rc = -9;
      goto fail;
    }
    k_gather_u32<<<blocks_for(nnz, B), B, 0, S()>>>(t32, d_cols, pos2, nnz);
    if (hipMemcpyAsync(cols, t32, 4 * nnz, hipMemcpyDeviceToHost, S()) ||
        hipStreamSynchronize(S())) {
      rc = -9;
      goto fail;
    }
    k_gather_f64<<<blocks_for(nnz, B), B, 0, S()>>>(t64, d_vals, pos2, nnz);
    if (hipMemcpyAsync(vals, t64, 8 * nnz, hipMemcpyDeviceToHost, S())) {
      rc = -9;
      goto fail;
    }
  }
  if (hipMemcpyAsync(h_row_idx, out->row_idx, 4 * (size_t)nrows,
                     hipMemcpyDeviceToHost, S()) ||
      hipMemcpyAsync(h_row_ridx, out->row_ridx, 4 * (size_t)nrows,
                     hipMemcpyDeviceToHost, S()) ||
      hipMemcpyAsync(h_col_idx, out->col_idx, 4 * (size_t)ncols,
                     hipMemcpyDeviceToHost, S()) ||
      hipMemcpyAsync(h_col_ridx, out->col_ridx, 4 * (size_t)ncols,
                     hipMemcpyDeviceToHost, S()) ||
      hipStreamSynchronize(S())) {
    rc = -10;
    goto fail;
  }
  goto done;
fail:
  acd_dev_matrix_free(out);
done:
#undef AL
  hipFree(d_rows);
  hipFree(d_cols);
  hipFree(d_vals);
  hipFree(cnt);
  hipFree(key);
  hipFree(key2);
  hipFree(pos);
  hipFree(pos2);
  hipFree(t32);
  hipFree(t64);
  return rc;
}

void acd_dev_matrix_free(struct acd_dev_matrix* m) {
  hipStreamSynchronize(S());
  free_sell(&m->a);
  free_sell(&m->at);
  hipFree(m->row_idx);
  hipFree(m->row_ridx);
  hipFree(m->col_idx);
  hipFree(m->col_ridx);
  m->row_idx = m->row_ridx = m->col_idx = m->col_ridx = nullptr;
}

int acd_dev_sell_make_fp32(struct acd_sell* op) {
  if (op->val32) return 0;
  CK(hipMalloc((void**)&op->val32, sizeof(float) * (op->nent ? op->nent : 1)),
     "hipMalloc fp32 values");
  if (op->nent)
    k_to_f32<<<blocks_for(op->nent, 256), 256, 0, S()>>>(op->val, op->val32,
                                                         op->nent);
  CK(hipStreamSynchronize(S()), "fp32 conversion");
  return 0;
}

int acd_dev_sell_make_bf16(struct acd_sell* op) {
  if (op->val16) return 1;
  uint16_t* v16 = nullptr;
  uint32_t* inexact = nullptr;
  CK(hipMalloc((void**)&v16, sizeof(uint16_t) * (op->nent ? op->nent : 1)),
     "hipMalloc bf16 values");
  if (hipMalloc((void**)&inexact, sizeof(uint32_t)) != hipSuccess) {
    hipFree(v16);
    acd_dev_record_error(-1, "hipMalloc bf16 flag");
    return -1;
  }
  uint32_t h = 0;
  hipMemsetAsync(inexact, 0, sizeof(uint32_t), S());
  if (op->nent)
    k_to_bf16<<<blocks_for(op->nent, 256), 256, 0, S()>>>(op->val, v16,
                                                          inexact, op->nent);
  hipMemcpyAsync(&h, inexact, sizeof h, hipMemcpyDeviceToHost, S());
  if (hipStreamSynchronize(S()) != hipSuccess) {
    hipFree(v16);
    hipFree(inexact);
    acd_dev_record_error(-1, "bf16 conversion");
    return -1;
  }
  hipFree(inexact);
  if (h) {
    hipFree(v16);
    return 0;
  }
  op->val16 = v16;
  return 1;
}

int acd_dev_sell_make_sign(struct acd_sell* op) {
  if (op->colsgn) return 1;
  uint32_t* w = nullptr;
  uint32_t* inexact = nullptr;
  CK(hipMalloc((void**)&w, sizeof(uint32_t) * (op->nent ? op->nent : 1)),
     "hipMalloc value-free entries");
  if (hipMalloc((void**)&inexact, sizeof(uint32_t)) != hipSuccess) {
    hipFree(w);
    acd_dev_record_error(-1, "hipMalloc value-free flag");
    return -1;
  }
  uint32_t h = 0;
  hipMemsetAsync(inexact, 0, sizeof(uint32_t), S());
  if (op->nent)
    k_to_sign<<<blocks_for(op->nent, 256), 256, 0, S()>>>(op->val, op->col, w,
                                                          inexact, op->nent);
  hipMemcpyAsync(&h, inexact, sizeof h, hipMemcpyDeviceToHost, S());
  if (hipStreamSynchronize(S()) != hipSuccess) {
    hipFree(w);
    hipFree(inexact);
    acd_dev_record_error(-1, "value-free conversion");
    return -1;
  }
  hipFree(inexact);
  if (h) {
    hipFree(w);
    return 0;
  }
  op->colsgn = w;
  return 1;
}

int acd_dev_sell_make_col16(struct acd_sell* op) {
  if (op->col16) return 1;
  if (op->ncols > 65536u) return 0;
  uint16_t* c = nullptr;
  if (hipMalloc((void**)&c, sizeof(uint16_t) * (op->nent ? op->nent : 1)) !=
      hipSuccess) {
    acd_dev_record_error(-1, "hipMalloc 16-bit columns");
    return -1;
  }
  if (op->nent)
    k_to_u16<<<blocks_for(op->nent, 256), 256, 0, S()>>>(op->col, c,
                                                         op->nent);
  if (hipGetLastError() != hipSuccess) {
    hipFree(c);
    acd_dev_record_error(-1, "16-bit columns");
    return -1;
  }
  op->col16 = c;
  return 1;
}

}  // extern "C"
