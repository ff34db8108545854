// Device-side helpers shared by the hand-written kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace taco_hip {

constexpr int kWave = 64;

// Streaming loads for data touched exactly once (CSR crd/vals): keep them
// from displacing the gathered operand (x or B rows) in L2/MALL.
template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Merge-path split of diagonal d over the lists (row ends pos[1..m],
// entry indices 0..nnz-1): returns r, the number of row ends consumed
// before diagonal d (the entry count consumed is d - r).  Ties go to the
// row end, i.e. entry z belongs to row r iff pos[r] <= z < pos[r+1].
// All 64 lanes of the wave cooperate: each round tests 64 evenly spaced
// pivots of the remaining range, so a 2M-row matrix resolves in 4 rounds of
// one gather each instead of 21 dependent loads.  Every lane returns r.
__device__ __forceinline__ int64_t wave_merge_search(const int32_t* pos,
                                                     int64_t m, int64_t nnz,
                                                     int64_t d) {
  int64_t lo = d - nnz > 0 ? d - nnz : 0;
  int64_t hi = d < m ? d : m;
  const int lane = (int)__lane_id();
  while (lo < hi) {
    const int64_t step = (hi - lo + kWave - 1) / kWave;
    const int64_t p = lo + (int64_t)lane * step;
    bool pred = false;
    if (p < hi) pred = (int64_t)pos[p + 1] <= d - p - 1;
    const uint64_t mask = __ballot(pred);
    const int c = __popcll(mask);
    const int64_t nlo = c == 0 ? lo : lo + (int64_t)(c - 1) * step + 1;
    int64_t nhi = lo + (int64_t)c * step;
    if (nhi > hi) nhi = hi;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// Inclusive segmented sum over the 64 lanes of a wave.  Keys must be
// non-decreasing with the lane index; lanes with equal keys form a segment.
template <typename T>
__device__ __forceinline__ T wave_segmented_inclusive_sum(T v, int key) {
  const int lane = (int)__lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T up = __shfl_up(v, off, kWave);
    int ukey = __shfl_up(key, off, kWave);
    if (lane >= off && ukey == key) v += up;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Float atomic add by an integer compare-and-swap loop.  Valid on every kind
// of memory a kernel can reach, including managed and fine-grained host
// memory, where the hardware float atomics may not take effect.
__device__ __forceinline__ void atomic_add_cas(double* p, double v) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(p);
  unsigned long long old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long assumed;
  do {
    assumed = old;
    old = atomicCAS(a, assumed,
                    (unsigned long long)__double_as_longlong(__longlong_as_double((long long)assumed) + v));
  } while (assumed != old);
}
__device__ __forceinline__ void atomic_add_cas(float* p, float v) {
  unsigned int* a = reinterpret_cast<unsigned int*>(p);
  unsigned int old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned int assumed;
  do {
    assumed = old;
    old = atomicCAS(a, assumed, __float_as_uint(__uint_as_float(assumed) + v));
  } while (assumed != old);
}

// Float atomic add on an output array.  coarse: the array is device memory
// from hipMalloc (coarse-grained), where the hardware float atomics
// (global_atomic_add_f64 / _f32) are valid; otherwise (managed memory) the
// compare-and-swap form is used.  `coarse` is uniform over the launch.
__device__ __forceinline__ void atomic_add(double* p, double v, bool coarse) {
  if (coarse)
    unsafeAtomicAdd(p, v);
  else
    atomic_add_cas(p, v);
}
__device__ __forceinline__ void atomic_add(float* p, float v, bool coarse) {
  if (coarse)
    unsafeAtomicAdd(p, v);
  else
    atomic_add_cas(p, v);
}

}  // namespace taco_hip
