// dev_runtime.hip -- HIP runtime plumbing behind acd_dev.h: device
// selection, the library stream, zero-filled allocations, copies and
// error capture.  Replaces huge_alloc/huge_alloc.default.inc:6-32 of
// pkhuong/acd-alm (64-B aligned zeroed allocation) with hipMalloc, whose
// allocations are 256-B aligned, plus an explicit zero fill.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "acd_dev.h"

namespace {
int g_device = 0;
int g_inited = 0;
hipStream_t g_stream = nullptr;
int g_cus = 0;
int g_err = 0;
char g_errtext[256];

int note(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    if (g_err == 0) {
      g_err = (int)e;
      snprintf(g_errtext, sizeof g_errtext, "%s: %s", what,
               hipGetErrorString(e));
    }
    return (int)e;
  }
  return 0;
}
}  // namespace

extern "C" {

int acd_dev_set_device(int device) {
  if (g_inited && device != g_device) return -1;
  g_device = device;
  return 0;
}

int acd_dev_device(void) { return g_device; }

int acd_dev_init(void) {
  if (g_inited) return 0;
  int n = 0;
  if (note(hipGetDeviceCount(&n), "hipGetDeviceCount")) return g_err;
  if (n <= 0 || g_device >= n) {
    acd_dev_record_error(-2, "no HIP device visible");
    return -2;
  }
  if (note(hipSetDevice(g_device), "hipSetDevice")) return g_err;
  if (note(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking),
           "hipStreamCreate"))
    return g_err;
  hipDeviceProp_t prop;
  if (note(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties"))
    return g_err;
  g_cus = prop.multiProcessorCount;
  g_inited = 1;
  return 0;
}

int acd_dev_cu_count(void) { return g_cus; }

void* acd_dev_stream(void) { return (void*)g_stream; }

void* acd_dev_stream_create(void) {
  if (acd_dev_init()) return nullptr;
  hipStream_t s = nullptr;
  if (note(hipStreamCreateWithFlags(&s, hipStreamNonBlocking),
           "hipStreamCreate"))
    return nullptr;
  return (void*)s;
}

void acd_dev_stream_destroy(void* s) {
  if (s) hipStreamDestroy((hipStream_t)s);
}

int acd_dev_set_stream(void* s) {
  g_stream = (hipStream_t)s;
  return 0;
}

void* acd_dev_malloc(size_t bytes) {
  if (acd_dev_init()) return nullptr;
  if (bytes == 0) bytes = 256;
  void* p = nullptr;
  if (note(hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
  if (note(hipMemsetAsync(p, 0, bytes, g_stream), "hipMemsetAsync")) {
    hipFree(p);
    return nullptr;
  }
  return p;
}

void acd_dev_free(void* p) {
  if (p == nullptr) return;
  // Frees are ordered after pending work on the library stream.
  hipStreamSynchronize(g_stream);
  note(hipFree(p), "hipFree");
}

int acd_dev_h2d(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  return note(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream),
              "hipMemcpy H2D") ||
         note(hipStreamSynchronize(g_stream), "sync after H2D");
}

int acd_dev_d2h(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  return note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream),
              "hipMemcpy D2H") ||
         note(hipStreamSynchronize(g_stream), "sync after D2H");
}

}  // extern "C"

namespace {
// Plain device copy, 16 bytes per lane per step, grid-stride (the
// reference kernel of known traffic for calibrating byte counters).
__global__ void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src,
                         size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" {

int acd_dev_copy_kernel(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) {
    acd_dev_record_error(-1, "copy kernel: size or alignment not 16 bytes");
    return -1;
  }
  const size_t n = bytes / 16;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_copy16<<<(unsigned)blocks, 256, 0, g_stream>>>((uint4*)dst,
                                                   (const uint4*)src, n);
  return note(hipGetLastError(), "copy kernel");
}

int acd_dev_d2d(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  return note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                             g_stream),
              "hipMemcpy D2D");
}

int acd_dev_memset(void* dst, int byte, size_t bytes) {
  if (bytes == 0) return 0;
  return note(hipMemsetAsync(dst, byte, bytes, g_stream), "hipMemsetAsync");
}

int acd_dev_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (acd_dev_init()) return -1;
  size_t f = 0, t = 0;
  if (note(hipMemGetInfo(&f, &t), "hipMemGetInfo")) return -1;
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return 0;
}

int acd_dev_sync(void) {
  return note(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
}

int acd_dev_last_error(char* buf, size_t len) {
  int e = g_err;
  if (buf && len) {
    strncpy(buf, e ? g_errtext : "", len - 1);
    buf[len - 1] = 0;
  }
  g_err = 0;
  g_errtext[0] = 0;
  return e;
}

void acd_dev_record_error(int code, const char* what) {
  if (g_err == 0) {
    g_err = code;
    snprintf(g_errtext, sizeof g_errtext, "%s", what);
  }
}

void acd_dev_check(int code, const char* what) {
  if (code == 0) return;
  char buf[256];
  acd_dev_last_error(buf, sizeof buf);
  fprintf(stderr, "acd-alm-amd: %s failed (%d): %s\n", what, code, buf);
  abort();
}

}  // extern "C"
