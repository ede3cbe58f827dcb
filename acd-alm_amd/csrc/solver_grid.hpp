// solver_grid.hpp -- execution groups for the device-resident solver.
//
// A "group" is the set of threads that cooperate on one LP: the whole
// grid (GridGroup, large LPs) or one workgroup (BlockGroup, batched small
// LPs).  Both give the solver the same three things:
//   * a wavefront enumeration (gwave / nwaves) used to deal out 64-row
//     slices: slice s is handled by wave s mod nwaves, lane l owns row
//     64s+l, so every pass over the same index space visits every index
//     with the same lane in the same order;
//   * sync(): a barrier over the group;
//   * reduce<K>(): K simultaneous reductions (sum or max per slot) with a
//     fixed combination order -- lanes by xor-butterfly (commutative adds,
//     so every lane ends with the same bits), waves in index order, blocks
//     in index order -- so results never depend on timing.
//
// The grid barrier is a generation counter in device memory.  Each wave
// fences its stores at agent scope before its workgroup arrives; the last
// arriving workgroup combines the per-block partials and publishes the
// results before releasing the generation.  Waiters spin with s_sleep and
// give up after a wall-clock timeout, raising an abort flag that every
// thread checks, so a fault in the protocol ends the kernel instead of
// hanging it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace acd {

constexpr int kWPB = 4;               // wavefronts per workgroup
constexpr int kBlock = 64 * kWPB;
constexpr int kMaxRed = 8;            // simultaneous reductions
constexpr int kMaxBlocks = 4096;
constexpr uint64_t kBarrierTimeoutTicks = 20ull * 100000000ull;  // 20 s @100 MHz

struct SyncArea {
  uint32_t count;
  uint32_t gen;
  uint32_t abort;
  uint32_t pad;
  double result[kMaxRed];
  double partial[kMaxBlocks * kMaxRed];
};

__device__ __forceinline__ uint32_t ld_acquire(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_f64(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p),
                     (uint64_t)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_f64(double* p) {
  return __longlong_as_double((long long)__hip_atomic_load(
      reinterpret_cast<uint64_t*>(p), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT));
}

// Reduction slot kinds.
enum RedOp : uint32_t { kSum = 0, kMax = 1 };

__device__ __forceinline__ double combine(double a, double b, bool is_max) {
  return is_max ? fmax(a, b) : a + b;
}

__device__ __forceinline__ double wave_reduce(double v, bool is_max) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = combine(v, w, is_max);
  }
  return v;
}

// Make a value provably uniform (it already is by construction).
__device__ __forceinline__ double uni(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
  int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)(unsigned)lo) |
                              ((long long)hi << 32));
}

struct GridGroup {
  uint32_t lane, wib, block, nblocks, gwave, nwaves;
  SyncArea* sa;
  bool aborted;

  __device__ __forceinline__ void init(SyncArea* s) {
    sa = s;
    lane = threadIdx.x & 63;
    wib = threadIdx.x >> 6;
    block = blockIdx.x;
    nblocks = gridDim.x;
    gwave = block * kWPB + wib;
    nwaves = nblocks * kWPB;
    aborted = false;
  }

  // K reductions; bit j of maxmask selects max for slot j.
  template <int K>
  __device__ __forceinline__ void reduce(double (&v)[K], uint32_t maxmask) {
    __shared__ double lds[kWPB * kMaxRed];
    __shared__ uint32_t sh_last, sh_gen;
#pragma unroll
    for (int j = 0; j < K; j++) v[j] = wave_reduce(v[j], (maxmask >> j) & 1);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < K; j++) lds[wib * kMaxRed + j] = v[j];
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int j = 0; j < K; j++) {
        double b = lds[j];
        for (int w = 1; w < kWPB; w++)
          b = combine(b, lds[w * kMaxRed + j], (maxmask >> j) & 1);
        st_f64(&sa->partial[block * kMaxRed + j], b);
      }
      uint32_t g0 = ld_acquire(&sa->gen);
      __threadfence();
      uint32_t old = __hip_atomic_fetch_add(&sa->count, 1u, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT);
      sh_last = (old == nblocks - 1);
      sh_gen = g0;
    }
    __syncthreads();
    if (sh_last) {
      if (wib == 0) {
#pragma unroll
        for (int j = 0; j < K; j++) {
          bool mx = (maxmask >> j) & 1;
          double acc = mx ? -HUGE_VAL : 0.0;
          bool first = true;
          for (uint32_t b = lane; b < nblocks; b += 64) {
            double p = ld_f64(&sa->partial[b * kMaxRed + j]);
            acc = first ? p : combine(acc, p, mx);
            first = false;
          }
          // lanes without a block contribute the identity
          acc = wave_reduce(first ? (mx ? -HUGE_VAL : 0.0) : acc, mx);
          if (lane == 0) st_f64(&sa->result[j], acc);
        }
        if (lane == 0) {
          __threadfence();
          st_relaxed(&sa->count, 0u);
          st_release(&sa->gen, sh_gen + 1);
        }
      }
    } else if (threadIdx.x == 0) {
      uint64_t t0 = wall_clock64();
      while (ld_acquire(&sa->gen) == sh_gen) {
        __builtin_amdgcn_s_sleep(1);
        if (ld_acquire(&sa->abort)) break;
        if (wall_clock64() - t0 > kBarrierTimeoutTicks) {
          st_release(&sa->abort, 1u);
          break;
        }
      }
      __threadfence();
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; j++) v[j] = uni(ld_f64(&sa->result[j]));
    if (ld_acquire(&sa->abort)) aborted = true;
  }

  __device__ __forceinline__ void sync() {
    double none[1] = {0.0};
    reduce<1>(none, 0);
  }
};

}  // namespace acd
