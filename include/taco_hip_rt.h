/* Runtime helpers for modules emitted by CodeGen_HIP (and shared with the
 * hand-written kernels).  The counterpart of the helper prelude taco prints
 * into every CUDA module (reference src/codegen/codegen_cuda.cpp:30-159),
 * written for gfx950 wave64: 64-bit ballots, lane groups of any power of two
 * up to 64, hardware float atomics, and errors reported to the caller instead
 * of exit().
 *
 * HIP C++ only (included by hipcc-compiled sources). */
#ifndef TACO_HIP_RT_H
#define TACO_HIP_RT_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "taco/taco_tensor_t.h"

extern "C" {
void* taco_hip_get_stream(void);
void* taco_hip_unified_alloc(size_t size);
void taco_hip_unified_free(void* ptr);
void* taco_hip_unified_realloc(void* ptr, size_t old_bytes, size_t new_bytes);
int taco_hip_unified_clear(void* ptr, size_t bytes);
int taco_hip_module_finish(int rc);
int taco_hip_module_report(int code, const char* what);
int taco_hip_host_read_bytes(void* dst, const void* src, size_t bytes);
int taco_hip_device_exclusive_scan(const int32_t* counts, int32_t m,
                                   int32_t* pos);
void* taco_hip_workspace_get(size_t bytes, int slot);
}

#define TACO_HIP_CHECK(call)                                               \
  do {                                                                     \
    int taco_rc_ = (int)(call);                                            \
    if (taco_rc_ != 0) return taco_hip_module_report(taco_rc_, #call);     \
  } while (0)
#define TACO_HIP_MIN(a, b) ((a) < (b) ? (a) : (b))
#define TACO_HIP_MAX(a, b) ((a) > (b) ? (a) : (b))

static inline hipStream_t taco_hip_stream() {
  return (hipStream_t)taco_hip_get_stream();
}

/* Host read of one element of a device array (synchronous).  Used where
 * taco's generated host code reads an index array directly, which it can do
 * only under unified memory. */
template <typename T>
static inline T taco_hip_host_read(const T* p, int64_t i) {
  T v = T();
  (void)taco_hip_host_read_bytes(&v, p + i, sizeof(T));
  return v;
}

/* Reports a non-zero status word written by a row kernel (synchronous). */
static inline int taco_hip_check_status(const int32_t* status) {
  int32_t v = taco_hip_host_read(status, 0);
  if (v & 1)
    return taco_hip_module_report(3, "a row exceeds the distinct-column "
                                     "capacity of the row kernel");
  if (v & 2)
    return taco_hip_module_report(4, "result exceeds the int32 index range");
  return 0;
}

/* ---- index searches (semantics of codegen_cuda.cpp:68-141) ---- */

/* Smallest idx in [lo, hi] with array[idx] >= target (hi if none below). */
__device__ __host__ inline int32_t taco_hip_binarySearchAfter(
    const int32_t* array, int32_t lo, int32_t hi, int32_t target) {
  if (array[lo] >= target) return lo;
  while (hi - lo > 1) { /* array[lo] < target <= array[hi] */
    int32_t mid = lo + (hi - lo) / 2;
    if (array[mid] < target) lo = mid; else hi = mid;
  }
  return hi;
}

/* Largest idx in [lo, hi] with array[idx] <= target. */
__device__ __host__ inline int32_t taco_hip_binarySearchBefore(
    const int32_t* array, int32_t lo, int32_t hi, int32_t target) {
  if (array[hi] <= target) return hi;
  while (hi - lo > 1) { /* array[lo] <= target < array[hi] */
    int32_t mid = lo + (hi - lo) / 2;
    if (array[mid] <= target) lo = mid; else hi = mid;
  }
  return lo;
}

/* Merge path over (row ends, entries) of a CSR pos array: the number of
 * rows r in [0, m) whose end item lies before merge position d, i.e. whose
 * pos[r + 1] + r + 1 <= d.  Row i = result is the row that position d falls
 * in; d - i entries precede it. */
__device__ __host__ inline int32_t taco_hip_mergePathSearch(const int32_t* pos,
                                                            int32_t m, int64_t d) {
  int32_t lo = 0, hi = m;
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)pos[mid + 1] + mid + 1 <= d) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Value of lane `idx` of the caller's team of `width` lanes (idx uniform
 * across the team).  A whole-wave team reads it with v_readlane into a
 * scalar; narrower teams use an LDS permute (__shfl). */
template <typename T>
__device__ inline T taco_hip_team_bcast(T v, int32_t idx, int32_t width) {
  if (width == 64) {
    if constexpr (sizeof(T) == 8) {
      const uint64_t b = __builtin_bit_cast(uint64_t, v);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, idx);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), idx);
      return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else if constexpr (sizeof(T) == 4) {
      return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), idx));
    }
  }
  return __shfl(v, idx, width);
}

/* results[b] = binarySearchBefore(array, lo, hi, b * per_block) for b in
 * [0, num_blocks]: the row where each block's entry chunk starts. */
__global__ __launch_bounds__(256) void taco_hip_binarySearchBeforeBlock(
    const int32_t* __restrict__ array, int32_t* __restrict__ results,
    int32_t lo, int32_t hi, int32_t per_block, int32_t num_blocks) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > num_blocks) return;
  int64_t target = b * (int64_t)per_block;
  results[b] = taco_hip_binarySearchBefore(
      array, lo, hi, target > INT_MAX ? INT_MAX : (int32_t)target);
}

static inline int taco_hip_binarySearchBeforeBlockLaunch(
    const int32_t* array, int32_t* results, int32_t lo, int32_t hi,
    int32_t per_block, int32_t block_size, int32_t num_blocks) {
  int64_t grid = ((int64_t)num_blocks + 1 + block_size - 1) / block_size;
  if (grid <= 0) return 0;
  hipLaunchKernelGGL(taco_hip_binarySearchBeforeBlock, dim3((unsigned)grid),
                     dim3(256), 0, taco_hip_stream(), array, results, lo, hi,
                     per_block, num_blocks);
  return (int)hipGetLastError();
}

/* ---- lane-group reductions and atomics ---- */

template <int G, typename T>
__device__ inline T taco_hip_group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

/* Sum over the workgroup's B threads (GPUBlockReduction,
 * codegen_cuda.cpp:875-890, where thread 0 adds the per-thread temporaries
 * serially after a barrier).  Every thread of the workgroup calls it and gets
 * the total: wave sums by a shuffle tree, then the per-wave sums through LDS,
 * added in wave order, so the result does not depend on timing. */
template <int B, typename T>
__device__ inline T taco_hip_block_sum(T v) {
  static_assert(B > 0 && B <= 1024 && (B <= 64 ? (B & (B - 1)) == 0 : B % 64 == 0),
                "block reduction: B must be a power of two <= 64 or a multiple of 64 <= 1024");
  if (B <= 64) return taco_hip_group_sum<B>(v);
  __shared__ T s_wave[B / 64 > 0 ? B / 64 : 1];
  v = taco_hip_group_sum<64>(v);
  __syncthreads(); /* earlier calls have read s_wave */
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x / 64] = v;
  __syncthreads();
  T t = s_wave[0];
#pragma unroll
  for (int w = 1; w < B / 64; ++w) t += s_wave[w];
  return t;
}

/* Float atomic adds.  By default the hardware float atomics, valid on the
 * coarse-grained device memory of hipMalloc.  Modules compiled with unified
 * memory on (TACO_HIP_COHERENT_ATOMICS, set by the runtime's default flags)
 * may write managed memory, where those may not take effect: they use an
 * integer compare-and-swap loop instead. */
#ifdef TACO_HIP_COHERENT_ATOMICS
__device__ inline void taco_hip_atomicAdd(double* a, int64_t i, double v) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a + i);
  unsigned long long old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long assumed;
  do {
    assumed = old;
    old = atomicCAS(p, assumed,
                    (unsigned long long)__double_as_longlong(__longlong_as_double((long long)assumed) + v));
  } while (assumed != old);
}
__device__ inline void taco_hip_atomicAdd(float* a, int64_t i, float v) {
  unsigned int* p = reinterpret_cast<unsigned int*>(a + i);
  unsigned int old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned int assumed;
  do {
    assumed = old;
    old = atomicCAS(p, assumed, __float_as_uint(__uint_as_float(assumed) + v));
  } while (assumed != old);
}
#else
__device__ inline void taco_hip_atomicAdd(double* a, int64_t i, double v) {
  unsafeAtomicAdd(a + i, v);
}
__device__ inline void taco_hip_atomicAdd(float* a, int64_t i, float v) {
  unsafeAtomicAdd(a + i, v);
}
#endif
__device__ inline void taco_hip_atomicAdd(int32_t* a, int64_t i, int32_t v) {
  atomicAdd(a + i, v);
}

/* Wave64 form of atomicAddWarp (codegen_cuda.cpp:142-159): when the whole
 * wave is active and every lane targets the first lane's index, the wave
 * sums with a shuffle tree and one lane issues a single atomic; otherwise
 * each lane issues its own. */
template <typename T>
__device__ inline void taco_hip_atomicAddWave(T* a, int64_t i, T v) {
  const uint64_t active = __ballot(1);
  const int64_t lead = __shfl(i, 0, 64);
  const bool same = __ballot(i == lead) == active;
  if (active == ~0ull && same) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (__lane_id() == 0) taco_hip_atomicAdd(a, i, v);
  } else {
    taco_hip_atomicAdd(a, i, v);
  }
}

/* Atomic multiply (taco's `*=` under Atomics, codegen_cuda.cpp:1354-1378,
 * where a single compare-and-swap is issued without a retry): a
 * compare-and-swap loop that retries until the product lands. */
__device__ inline void taco_hip_atomicMul(double* a, int64_t i, double v) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a + i);
  unsigned long long old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long assumed;
  do {
    assumed = old;
    old = atomicCAS(p, assumed,
                    (unsigned long long)__double_as_longlong(__longlong_as_double((long long)assumed) * v));
  } while (assumed != old);
}
__device__ inline void taco_hip_atomicMul(float* a, int64_t i, float v) {
  unsigned int* p = reinterpret_cast<unsigned int*>(a + i);
  unsigned int old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned int assumed;
  do {
    assumed = old;
    old = atomicCAS(p, assumed, __float_as_uint(__uint_as_float(assumed) * v));
  } while (assumed != old);
}
__device__ inline void taco_hip_atomicMul(int32_t* a, int64_t i, int32_t v) {
  int32_t old = __hip_atomic_load(a + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int32_t assumed;
  do {
    assumed = old;
    old = atomicCAS(a + i, assumed, assumed * v);
  } while (assumed != old);
}

/* Orders the LDS accesses of the lanes of one wave around a group
 * reduction: the group lies inside one wave, whose lanes execute together,
 * so a compiler and wavefront-scope memory fence is all it needs. */
__device__ inline void taco_hip_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* A group reduction's barrier when the group width is known only at run
 * time: wave-level ordering for groups inside one wave, else the
 * workgroup barrier. */
__device__ inline void taco_hip_group_sync(int32_t width) {
  if (width <= 64 && 64 % width == 0)
    taco_hip_wave_sync();
  else
    __syncthreads();
}

/* ---- host side of modules printed from taco's own IR ---- */

/* The host waits for the library stream before code that may touch memory
 * a queued kernel uses. */
static inline int taco_hip_stream_wait() {
  return (int)hipStreamSynchronize(taco_hip_stream());
}

/* A host scalar's device copy for a kernel that assigns it, and its value
 * back after the launch (stream-ordered, then waited for). */
static inline int taco_hip_copy_in(void* dev, const void* host, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, taco_hip_stream());
  if (e != hipSuccess) return (int)e;
  return (int)hipStreamSynchronize(taco_hip_stream());
}
static inline int taco_hip_copy_out(void* host, const void* dev, size_t bytes) {
  hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, taco_hip_stream());
  if (e != hipSuccess) return (int)e;
  return (int)hipStreamSynchronize(taco_hip_stream());
}

/* cudaMemset of taco's initValues (lowerer_impl_imperative.cpp:3357-3362):
 * stream-ordered. */
static inline int taco_hip_memsetAsync(void* p, int value, size_t bytes) {
  if (bytes == 0) return 0;
  return (int)hipMemsetAsync(p, value, bytes, taco_hip_stream());
}

static inline int taco_hip_cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Row of each workgroup's first entry (taco_binarySearchBeforeBlockLaunch,
 * codegen_cuda.cpp:112-122): results[b] = binarySearchBefore(array, lo, hi,
 * b * per_block) for b in [0, num_blocks].  Launched on the library stream;
 * returns results, which taco's IR assigns back. */
static inline int32_t* taco_hip_searchBlockStarts(const int32_t* array,
                                                  int32_t* results, int32_t lo,
                                                  int32_t hi, int32_t per_block,
                                                  int32_t block_size,
                                                  int32_t num_blocks) {
  (void)block_size;
  if (num_blocks < 0) return results;
  int64_t grid = ((int64_t)num_blocks + 1 + 255) / 256;
  hipLaunchKernelGGL(taco_hip_binarySearchBeforeBlock, dim3((unsigned)grid),
                     dim3(256), 0, taco_hip_stream(), array, results, lo, hi,
                     per_block, num_blocks);
  return results;
}

/* results[b] = binarySearchBefore(array, lo, hi, targets[b]) for b in
 * [0, num_blocks] (codegen_cuda.cpp:124-134). */
__global__ __launch_bounds__(256) void taco_hip_binarySearchIndirectBeforeBlock(
    const int32_t* __restrict__ array, int32_t* __restrict__ results,
    int32_t lo, int32_t hi, const int32_t* __restrict__ targets,
    int32_t num_blocks) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > num_blocks) return;
  results[b] = taco_hip_binarySearchBefore(array, lo, hi, targets[b]);
}

static inline int32_t* taco_hip_searchBlockStartsIndirect(
    const int32_t* array, int32_t* results, int32_t lo, int32_t hi,
    const int32_t* targets, int32_t block_size, int32_t num_blocks) {
  (void)block_size;
  if (num_blocks < 0) return results;
  int64_t grid = ((int64_t)num_blocks + 1 + 255) / 256;
  hipLaunchKernelGGL(taco_hip_binarySearchIndirectBeforeBlock,
                     dim3((unsigned)grid), dim3(256), 0, taco_hip_stream(),
                     array, results, lo, hi, targets, num_blocks);
  return results;
}

/* ---- fills (a kernel rather than a memset node, safe under capture) ---- */

template <typename T>
__global__ __launch_bounds__(256) void taco_hip_fill_kernel(T* p, int64_t n,
                                                            T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static inline int taco_hip_fill(T* p, int64_t n, T v) {
  if (n <= 0) return 0;
  int64_t grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(taco_hip_fill_kernel<T>, dim3((unsigned)grid), dim3(256),
                     0, taco_hip_stream(), p, n, v);
  return (int)hipGetLastError();
}

/* ---- sorting ---- */

__device__ inline void taco_hip_device_sort_i32(int32_t* a, int32_t n) {
  for (int32_t i = 1; i < n; ++i) {
    int32_t v = a[i], j = i - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = v;
  }
}

/* Ascending bitonic sort of one value per lane across the 64-lane wave. */
template <typename K>
__device__ inline K taco_hip_wave_sort(K v) {
  const int lane = (int)__lane_id();
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const K p = __shfl_xor(v, j, 64);
      const bool up = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      v = (lower == up) ? (p < v ? p : v) : (p > v ? p : v);
    }
  }
  return v;
}

/* The value of lane (lane ^ J), J a power of two below 64, by register
 * moves only: DPP quad permutes for 1 and 2, two row shifts for 4, a row
 * rotation by 8 for 8, and gfx950's lane-half swaps for 16 and 32 (no LDS
 * crossbar, unlike a general shuffle).  Every lane of the wave must be
 * active: take the result into a variable before any branch or ?: on it. */
template <int J>
__device__ inline uint32_t taco_hip_lane_xor_u32(uint32_t v) {
  const int lane = (int)__lane_id();
  const int s = (int)v;
  if constexpr (J == 1) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, s, 0xB1, 0xF, 0xF, false);  /* quad_perm 1,0,3,2 */
  } else if constexpr (J == 2) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, s, 0x4E, 0xF, 0xF, false);  /* quad_perm 2,3,0,1 */
  } else if constexpr (J == 4) {
    const int from_above = __builtin_amdgcn_update_dpp(0, s, 0x104, 0xF, 0xF, false);  /* row_shl:4 */
    const int from_below = __builtin_amdgcn_update_dpp(0, s, 0x114, 0xF, 0xF, false);  /* row_shr:4 */
    return (uint32_t)((lane & 4) ? from_below : from_above);
  } else if constexpr (J == 8) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, s, 0x128, 0xF, 0xF, false);  /* row_ror:8 */
  } else if constexpr (J == 16) {
    /* Swaps rows 1 and 3 of the first operand with rows 0 and 2 of the
     * second: lanes of odd rows find their partner in the first result,
     * lanes of even rows in the second. */
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    static_assert(J == 32, "lane xor distance");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
}

/* Inclusive prefix sum across the 64-lane wave by DPP moves: row shifts by
 * 1, 2, 4 and 8 within each 16-lane row, then the last lane of row 0 added
 * to row 1 and of row 2 to row 3 (row_bcast:15), and the last lane of row 1
 * to rows 2 and 3 (row_bcast:31).  Every lane of the wave must be active. */
__device__ inline int32_t taco_hip_wave_scan_add_i32(int32_t v) {
  int32_t u;
  u = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false); /* row_shr:1 */
  v += u;
  u = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false); /* row_shr:2 */
  v += u;
  u = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false); /* row_shr:4 */
  v += u;
  u = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false); /* row_shr:8 */
  v += u;
  u = __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false); /* row_bcast:15 */
  v += u;
  u = __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false); /* row_bcast:31 */
  v += u;
  return v;
}

template <int K, int J>
__device__ inline uint32_t taco_hip_bitonic_step_u32(uint32_t v) {
  const int lane = (int)__lane_id();
  const uint32_t p = taco_hip_lane_xor_u32<J>(v);
  const bool keep_min = ((lane & K) == 0) == ((lane & J) == 0);
  return keep_min ? (p < v ? p : v) : (p > v ? p : v);
}

template <int K>
__device__ inline uint32_t taco_hip_bitonic_merge_u32(uint32_t v) {
  if constexpr (K >= 64) v = taco_hip_bitonic_step_u32<K, 32>(v);
  if constexpr (K >= 32) v = taco_hip_bitonic_step_u32<K, 16>(v);
  if constexpr (K >= 16) v = taco_hip_bitonic_step_u32<K, 8>(v);
  if constexpr (K >= 8) v = taco_hip_bitonic_step_u32<K, 4>(v);
  if constexpr (K >= 4) v = taco_hip_bitonic_step_u32<K, 2>(v);
  return taco_hip_bitonic_step_u32<K, 1>(v);
}

/* The 21 compare-exchange steps' lane masks of the wave bitonic sort (all
 * ones where the lane keeps the smaller value), built once per kernel and
 * held in vector registers: as scalar lane masks they would be 42 SGPRs,
 * beyond what a kernel keeps without spilling them. */
struct TacoHipSortMasks {
  uint32_t m[21];
  __device__ TacoHipSortMasks() {
    const int lane = (int)__lane_id();
    int s = 0;
#pragma unroll
    for (int K = 2; K <= 64; K <<= 1) {
#pragma unroll
      for (int J = K >> 1; J > 0; J >>= 1) {
        uint32_t keep = (((lane & K) == 0) == ((lane & J) == 0)) ? 0xFFFFFFFFu : 0u;
        asm volatile("" : "+v"(keep)); /* keep it a per-lane value */
        m[s++] = keep;
      }
    }
  }
};

template <int K, int J>
__device__ inline uint32_t taco_hip_bitonic_step_masked(uint32_t v, uint32_t keep_min) {
  const uint32_t p = taco_hip_lane_xor_u32<J>(v);
  const uint32_t lo = p < v ? p : v, hi = p > v ? p : v;
  return hi ^ ((lo ^ hi) & keep_min); /* one three-input bit operation */
}

template <int K>
__device__ inline uint32_t taco_hip_bitonic_merge_masked(uint32_t v, const uint32_t* m) {
  /* m: this merge's steps, J = K/2 down to 1 */
  int s = 0;
  if constexpr (K >= 64) v = taco_hip_bitonic_step_masked<K, 32>(v, m[s++]);
  if constexpr (K >= 32) v = taco_hip_bitonic_step_masked<K, 16>(v, m[s++]);
  if constexpr (K >= 16) v = taco_hip_bitonic_step_masked<K, 8>(v, m[s++]);
  if constexpr (K >= 8) v = taco_hip_bitonic_step_masked<K, 4>(v, m[s++]);
  if constexpr (K >= 4) v = taco_hip_bitonic_step_masked<K, 2>(v, m[s++]);
  return taco_hip_bitonic_step_masked<K, 1>(v, m[s]);
}

/* taco_hip_wave_sort_u32 with the step masks given (same result). */
__device__ inline uint32_t taco_hip_wave_sort_u32(uint32_t v, const TacoHipSortMasks& sm) {
  v = taco_hip_bitonic_merge_masked<2>(v, sm.m + 0);
  v = taco_hip_bitonic_merge_masked<4>(v, sm.m + 1);
  v = taco_hip_bitonic_merge_masked<8>(v, sm.m + 3);
  v = taco_hip_bitonic_merge_masked<16>(v, sm.m + 6);
  v = taco_hip_bitonic_merge_masked<32>(v, sm.m + 10);
  return taco_hip_bitonic_merge_masked<64>(v, sm.m + 15);
}

/* taco_hip_wave_sort for 32-bit unsigned keys, every exchange a register
 * move (taco_hip_lane_xor_u32). */
__device__ inline uint32_t taco_hip_wave_sort_u32(uint32_t v) {
  v = taco_hip_bitonic_merge_u32<2>(v);
  v = taco_hip_bitonic_merge_u32<4>(v);
  v = taco_hip_bitonic_merge_u32<8>(v);
  v = taco_hip_bitonic_merge_u32<16>(v);
  v = taco_hip_bitonic_merge_u32<32>(v);
  return taco_hip_bitonic_merge_u32<64>(v);
}

/* Workgroups for a grid-strided row loop over m rows, 4 rows (waves) per
 * workgroup: at most 4 workgroups per CU, which every row kernel here keeps
 * resident at once. */
static inline int32_t taco_hip_row_grid(int32_t m) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    cus = n > 0 ? n : 256;
  }
  const int64_t want = ((int64_t)m + 3) / 4;
  return (int32_t)(want < 4 * (int64_t)cus ? want : 4 * (int64_t)cus);
}

/* ---- SpGEMM rows, one 64-lane wave per row of C = A * B ----
 * Pattern of row i: sorted distinct structural columns of the B rows named by
 * A's row i.  Values: products summed per column in encounter order (jA
 * outer, kB inner), the first stored, later ones added.
 * Short rows (<= 64 candidates) live in lanes; long rows use a per-wave LDS
 * hash set of at most TACO_HIP_ROW_CAP distinct columns.  Callers launch
 * 256-thread workgroups (4 waves). */

#define TACO_HIP_ROW_CAP 1024
#define TACO_HIP_ROW_WAVES 4

struct TacoHipRowSet {
  int32_t keys[TACO_HIP_ROW_CAP];
  int32_t count;
};

__device__ inline int64_t taco_hip_row_candidates(const int32_t* apos,
                                                  const int32_t* acrd,
                                                  const int32_t* bpos,
                                                  int32_t i) {
  const int lane = (int)__lane_id();
  int64_t n = 0;
  for (int32_t p = apos[i] + lane; p < apos[i + 1]; p += 64) {
    const int32_t j = acrd[p];
    n += bpos[j + 1] - bpos[j];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  return n;
}

/* Candidate q of row i in encounter order (short rows). */
__device__ inline bool taco_hip_row_candidate(
    const int32_t* apos, const int32_t* acrd, const int32_t* bpos,
    const int32_t* bcrd, int32_t i, int q, int32_t* col, int32_t* jA,
    int32_t* kB) {
  int acc = 0;
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    const int32_t b0 = bpos[j], len = bpos[j + 1] - b0;
    if (q < acc + len) {
      *jA = p;
      *kB = b0 + (q - acc);
      *col = bcrd[b0 + (q - acc)];
      return true;
    }
    acc += len;
  }
  return false;
}

/* Short-row candidates in lanes, loaded cooperatively: lane t < (row length
 * of A) reads A's entry t and its B row bounds, a wave prefix sum of the B
 * row lengths places candidate q (encounter order) in lane q, which then
 * reads its column.  Returns the row's candidate count when A's row has at
 * most 64 entries (lanes hold candidates only if it is <= 64), or -1 when A's
 * row is longer (nothing loaded).  Every lane of the wave must call it. */
__device__ inline int64_t taco_hip_row_lanes(
    const int32_t* apos, const int32_t* acrd, const int32_t* bpos,
    const int32_t* bcrd, int32_t i, int32_t* col, int32_t* jA, int32_t* kB) {
  const int lane = (int)__lane_id();
  const int32_t a0 = apos[i], la = apos[i + 1] - a0;
  if (la > 64) return -1;
  int32_t b0 = 0, len = 0;
  if (lane < la) {
    const int32_t j = acrd[a0 + lane];
    b0 = bpos[j];
    len = bpos[j + 1] - b0;
  }
  /* B row lengths clamped to 65: the prefix sums stay in int32 and a total
   * above 64 still reads as above 64. */
  const int32_t lc = len < 65 ? len : 65;
  const int32_t inc = taco_hip_wave_scan_add_i32(lc);
  const int32_t total = __builtin_amdgcn_readlane(inc, 63);
  if (total > 64) {
    int64_t n = len;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    return n;
  }
  /* Entry t holding candidate `lane`: the number of entries s < la - 1 whose
   * inclusive sum is <= lane. */
  int t = 0;
  for (int s2 = 0; s2 + 1 < la; ++s2) t += __builtin_amdgcn_readlane(inc, s2) <= lane ? 1 : 0;
  const int32_t excl = __shfl(inc - lc, t, 64);
  const int32_t bt = __shfl(b0, t, 64);
  if (lane < total) {
    *jA = a0 + t;
    *kB = bt + (lane - excl);
    *col = bcrd[*kB];
  }
  return total;
}

__device__ inline void taco_hip_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

/* Distinct columns of a long row into w (hash set).  -1 on overflow. */
__device__ inline int32_t taco_hip_row_hash(TacoHipRowSet* w,
                                            const int32_t* apos,
                                            const int32_t* acrd,
                                            const int32_t* bpos,
                                            const int32_t* bcrd, int32_t i) {
  const int lane = (int)__lane_id();
  for (int q = lane; q < TACO_HIP_ROW_CAP; q += 64) w->keys[q] = -1;
  if (lane == 0) w->count = 0;
  taco_hip_wave_lds_sync();
  bool overflow = false;
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    for (int32_t kB = bpos[j] + lane; kB < bpos[j + 1]; kB += 64) {
      const int32_t col = bcrd[kB];
      uint32_t h = ((uint32_t)col * 2654435761u) & (TACO_HIP_ROW_CAP - 1);
      for (int probe = 0; probe < TACO_HIP_ROW_CAP; ++probe) {
        const int32_t prev = atomicCAS(&w->keys[h], -1, col);
        if (prev == -1) {
          atomicAdd(&w->count, 1);
          break;
        }
        if (prev == col) break;
        h = (h + 1) & (TACO_HIP_ROW_CAP - 1);
        if (probe == TACO_HIP_ROW_CAP - 1) overflow = true;
      }
    }
  }
  taco_hip_wave_lds_sync();
  const bool any_overflow = __ballot(overflow) != 0;
  const int32_t cnt = w->count;
  return any_overflow || cnt > TACO_HIP_ROW_CAP ? -1 : cnt;
}

/* Sorts the hash set's keys ascending into w->keys[0..count). */
__device__ inline void taco_hip_row_hash_sort(TacoHipRowSet* w) {
  const int lane = (int)__lane_id();
  for (int q = lane; q < TACO_HIP_ROW_CAP; q += 64)
    if (w->keys[q] == -1) w->keys[q] = INT_MAX;
  taco_hip_wave_lds_sync();
  for (int k = 2; k <= TACO_HIP_ROW_CAP; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int q = lane; q < TACO_HIP_ROW_CAP; q += 64) {
        const int partner = q ^ j;
        if (partner > q) {
          const int32_t a = w->keys[q], b = w->keys[partner];
          const bool up = (q & k) == 0;
          if ((a > b) == up) {
            w->keys[q] = b;
            w->keys[partner] = a;
          }
        }
      }
      taco_hip_wave_lds_sync();
    }
  }
}

/* Symbolic row: returns the distinct count (count pass) or writes the sorted
 * columns at ccrd[cpos_i ...] (fill pass).  status bit 1 on overflow. */
template <bool FILL>
__device__ inline int32_t taco_hip_spgemm_row_symbolic(
    const int32_t* apos, const int32_t* acrd, const int32_t* bpos,
    const int32_t* bcrd, int32_t i, int32_t cpos_i, int32_t* ccrd,
    int32_t* status) {
  __shared__ TacoHipRowSet s_sets[TACO_HIP_ROW_WAVES];
  const int lane = (int)__lane_id();
  int32_t col = INT_MAX, jA = 0, kB = 0;
  int64_t n = taco_hip_row_lanes(apos, acrd, bpos, bcrd, i, &col, &jA, &kB);
  if (n < 0) {
    n = taco_hip_row_candidates(apos, acrd, bpos, i);
    if (n <= 64 && lane < n)
      taco_hip_row_candidate(apos, acrd, bpos, bcrd, i, lane, &col, &jA, &kB);
  }
  if (n <= 64) {
    const int32_t s = (int32_t)taco_hip_wave_sort_u32((uint32_t)col);
    const int32_t prev = __shfl_up(s, 1, 64);
    const bool head = s != INT_MAX && (lane == 0 || prev != s);
    const uint64_t heads = __ballot(head);
    if (FILL && head) {
      const int rank = __popcll(heads & ((1ull << lane) - 1));
      __builtin_nontemporal_store(s, ccrd + cpos_i + rank); /* written once */
    }
    return __popcll(heads);
  }
  TacoHipRowSet* w = &s_sets[(threadIdx.x / 64) % TACO_HIP_ROW_WAVES];
  const int32_t cnt = taco_hip_row_hash(w, apos, acrd, bpos, bcrd, i);
  if (cnt < 0) {
    if (lane == 0) atomicOr(status, 1);
    return 0;
  }
  if (FILL) {
    taco_hip_row_hash_sort(w);
    for (int q = lane; q < cnt; q += 64) __builtin_nontemporal_store(w->keys[q], ccrd + cpos_i + q);
  }
  return cnt;
}

/* Numeric row of more than TACO_HIP_ROW_CAP columns: C's values are the
 * accumulator, zeroed first.  Candidates are taken 64 at a time in
 * encounter order (jA outer, kB inner), across A's entries: lane t first
 * reads entry p+t's B row bounds, a wave prefix sum of the row lengths
 * places candidate q in lane q.  Each candidate finds its slot by binary
 * search in the row's sorted pattern; the chunk is sorted by (slot, lane),
 * and the first lane of each slot's run adds the run's products to the slot
 * one after another.  Every column so sums its products in encounter
 * order, as taco's workspace loop does. */
__device__ inline void taco_hip_spgemm_row_numeric_global(
    const int32_t* apos, const int32_t* acrd, const double* avals, const int32_t* bpos,
    const int32_t* bcrd, const double* bvals, int32_t i, int32_t c0, int32_t nc,
    const int32_t* ccrd, double* cvals) {
  const int lane = (int)__lane_id();
  const int32_t* crow = ccrd + c0;
  double* out = cvals + c0;
  for (int32_t q = lane; q < nc; q += 64) out[q] = 0.0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  const int32_t a1 = apos[i + 1];
  int32_t p = apos[i], off = 0; /* next candidate: entry p, offset off in its B row */
  while (p < a1) {
    int32_t b0 = 0, avail = 0;
    double av = 0.0;
    if (p + lane < a1) {
      const int32_t j = acrd[p + lane];
      b0 = bpos[j];
      avail = bpos[j + 1] - b0;
      av = avals[p + lane];
      if (lane == 0) {
        b0 += off;
        avail -= off;
      }
    }
    /* lengths clamped to 65: sums stay in int32, a total above 64 still
     * reads as above 64 */
    const int32_t lc = avail < 65 ? avail : 65;
    int32_t inc = lc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    const int32_t tot = __shfl(inc, 63, 64);
    const int32_t used = tot < 64 ? tot : 64;
    /* entry of candidate `lane`: the number of entries whose inclusive sum
     * is <= lane */
    int t = 0;
    for (int s2 = 0; s2 < 63; ++s2) t += __shfl(inc, s2, 64) <= lane ? 1 : 0;
    const int tq = t < 64 ? t : 63;
    const int32_t excl = __shfl(inc - lc, tq, 64);
    const int32_t bt = __shfl(b0, tq, 64);
    const double at = __shfl(av, tq, 64);
    const bool valid = lane < used;
    uint32_t slot = 0xFFFFFFFFu;
    double prod = 0.0;
    if (valid) {
      const int32_t kB = bt + (lane - excl);
      const int32_t col = bcrd[kB];
      prod = at * bvals[kB];
      int32_t lo = 0, hi = nc;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (crow[mid] < col) lo = mid + 1; else hi = mid;
      }
      if (lo < nc && crow[lo] == col) slot = (uint32_t)lo;
    }
    uint64_t key = ((uint64_t)slot << 32) | (uint32_t)lane;
    key = taco_hip_wave_sort(key);
    const uint32_t sslot = (uint32_t)(key >> 32);
    const double v = __shfl(prod, (int)(key & 63), 64);
    const uint32_t prev = __shfl_up(sslot, 1, 64);
    const bool live = sslot != 0xFFFFFFFFu;
    const bool head = live && (lane == 0 || prev != sslot);
    const uint64_t heads = __ballot(head);
    const int nlive = __popcll(__ballot(live));
    int run = 0;
    if (head) {
      const uint64_t above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1);
      const int next = above ? __ffsll((unsigned long long)above) - 1 : nlive;
      run = next - lane;
    }
    int maxrun = run;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const int x = __shfl_xor(maxrun, o, 64);
      maxrun = x > maxrun ? x : maxrun;
    }
    double acc = 0.0;
    if (head) acc = out[sslot] + v;
    for (int q = 1; q < maxrun; ++q) {
      const double u = __shfl(v, (lane + q) & 63, 64);
      if (head && q < run) acc = acc + u;
    }
    if (head) out[sslot] = acc;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    /* advance past the `used` candidates */
    if (used == 0) {
      p += 64;
      off = 0;
    } else {
      const int tl = __shfl(t, used - 1, 64);
      const int32_t done = used - __shfl(inc - lc, tl, 64); /* taken from entry tl */
      const int32_t left = __shfl(avail, tl, 64) - done;
      if (left == 0) {
        off = 0;
        p += tl + 1;
      } else {
        off = (tl == 0 ? off : 0) + done;
        p += tl;
      }
    }
  }
}

/* Numeric row on an assembled (sorted) pattern. */
/* Numeric row of at most 64 candidates held in lanes, lane c holding
 * candidate c (encounter order) with its column and product (lanes c >= n:
 * column INT_MAX): the products are sorted by (column, encounter index) and
 * summed per column in encounter order; the row's values go to
 * cvals[c0 .. c0 + nc).  Every lane of the wave must call it. */
__device__ inline void taco_hip_spgemm_row_sum_store(int64_t n, int32_t col, double prod,
                                                     int32_t c0, int32_t nc, double* cvals,
                                                     const TacoHipSortMasks* sm = nullptr) {
  const int lane = (int)__lane_id();
  (void)n;
  /* Key = column, then encounter index (which also names the lane): in
   * 32 bits when every column is below 2^26, else in 64. */
  int32_t scol, src;
  if (__ballot(col != INT_MAX && col >= (1 << 26)) == 0) {
    uint32_t key = col == INT_MAX ? 0xFFFFFFFFu : ((uint32_t)col << 6) | (uint32_t)lane;
    key = sm ? taco_hip_wave_sort_u32(key, *sm) : taco_hip_wave_sort_u32(key);
    scol = key == 0xFFFFFFFFu ? INT_MAX : (int32_t)(key >> 6);
    src = (int32_t)(key & 63);
  } else {
    uint64_t key = ((uint64_t)(uint32_t)col << 32) | (uint32_t)lane;
    key = taco_hip_wave_sort(key);
    scol = (int32_t)(key >> 32);
    src = (int32_t)(key & 63);
  }
  const double v = __shfl(prod, src, 64);
  const int32_t prev = __shfl_up(scol, 1, 64);
  const bool valid = scol != INT_MAX;
  const bool head = valid && (lane == 0 || prev != scol);
  const uint64_t heads = __ballot(head);
  const int nvalid = __popcll(__ballot(valid));
  int run = 0;
  if (head) {
    const uint64_t above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1);
    const int next = above ? __ffsll((unsigned long long)above) - 1 : nvalid;
    run = next - lane;
  }
  /* Wave maximum of the run lengths.  Each lane move is taken in full-wave
   * control flow, before the comparison: a lane move under a branch would
   * read lanes the branch has switched off. */
  uint32_t maxrun = (uint32_t)run, other;
  other = taco_hip_lane_xor_u32<32>(maxrun);
  maxrun = maxrun > other ? maxrun : other;
  other = taco_hip_lane_xor_u32<16>(maxrun);
  maxrun = maxrun > other ? maxrun : other;
  other = taco_hip_lane_xor_u32<8>(maxrun);
  maxrun = maxrun > other ? maxrun : other;
  other = taco_hip_lane_xor_u32<4>(maxrun);
  maxrun = maxrun > other ? maxrun : other;
  other = taco_hip_lane_xor_u32<2>(maxrun);
  maxrun = maxrun > other ? maxrun : other;
  other = taco_hip_lane_xor_u32<1>(maxrun);
  maxrun = maxrun > other ? maxrun : other;
  double s = v;
  for (int q = 1; q < (int)maxrun; ++q) {
    const double u = __shfl(v, (lane + q) & 63, 64);
    if (head && q < run) s = s + u;
  }
  if (head) {
    const int rank = __popcll(heads & ((1ull << lane) - 1));
    if (rank < nc) __builtin_nontemporal_store(s, cvals + c0 + rank); /* written once */
  }
}

/* Numeric row of at most 64 candidates, held in lanes (taco_hip_row_lanes). */
__device__ inline void taco_hip_spgemm_row_numeric_short(
    int64_t n, int32_t col, int32_t jA, int32_t kB, const double* avals,
    const double* bvals, int32_t c0, int32_t nc, double* cvals) {
  double prod = 0;
  if ((int)__lane_id() < n) prod = avals[jA] * bvals[kB];
  taco_hip_spgemm_row_sum_store(n, col, prod, c0, nc, cvals);
}

__device__ inline void taco_hip_spgemm_row_numeric(
    const int32_t* apos, const int32_t* acrd, const double* avals,
    const int32_t* bpos, const int32_t* bcrd, const double* bvals, int32_t i,
    const int32_t* cpos, const int32_t* ccrd, double* cvals) {
  __shared__ double s_acc[TACO_HIP_ROW_WAVES][TACO_HIP_ROW_CAP];
  const int lane = (int)__lane_id();
  const int32_t c0 = cpos[i], nc = cpos[i + 1] - c0;
  int32_t col = INT_MAX, jA = 0, kB = 0;
  int64_t n = taco_hip_row_lanes(apos, acrd, bpos, bcrd, i, &col, &jA, &kB);
  if (n < 0) {
    n = taco_hip_row_candidates(apos, acrd, bpos, i);
    if (n <= 64 && lane < n)
      taco_hip_row_candidate(apos, acrd, bpos, bcrd, i, lane, &col, &jA, &kB);
  }
  if (n <= 64) {
    taco_hip_spgemm_row_numeric_short(n, col, jA, kB, avals, bvals, c0, nc, cvals);
    return;
  }
  if (nc > TACO_HIP_ROW_CAP) {
    taco_hip_spgemm_row_numeric_global(apos, acrd, avals, bpos, bcrd, bvals, i, c0, nc, ccrd,
                                       cvals);
    return;
  }
  double* acc = s_acc[(threadIdx.x / 64) % TACO_HIP_ROW_WAVES];
  for (int q = lane; q < nc; q += 64) acc[q] = 0;
  taco_hip_wave_lds_sync();
  const int32_t* crow = ccrd + c0;
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    const double a = avals[p];
    const int32_t bend = bpos[j + 1];
    for (int32_t kb0 = bpos[j]; kb0 < bend; kb0 += 64) {
      const int32_t kB = kb0 + lane;
      int slot = -1;
      double prod = 0;
      if (kB < bend) {
        const int32_t col = bcrd[kB];
        prod = a * bvals[kB];
        int lo = 0, hi = nc;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (crow[mid] < col) lo = mid + 1; else hi = mid;
        }
        slot = lo < nc && crow[lo] == col ? lo : -1;
      }
      const int lim = bend - kb0 < 64 ? bend - kb0 : 64;
      for (int q = 0; q < lim; ++q) {
        const int sq = __shfl(slot, q, 64);
        const double vq = __shfl(prod, q, 64);
        if (lane == 0 && sq >= 0) acc[sq] = acc[sq] + vq;
        taco_hip_wave_lds_sync();
      }
    }
  }
  for (int q = lane; q < nc; q += 64) __builtin_nontemporal_store(acc[q], cvals + c0 + q);
}

#endif
