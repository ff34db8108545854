// Gather-throughput ceiling: how fast can MI355X serve the random 8-byte
// loads that dominate SpMV on scattered columns (y += a * x[col])?  The
// kernels here issue only such loads — no CSR stream, no output traffic
// beyond one value per thread — from an f64 array of 2^k entries, with the
// indices from an in-register generator.  The measured rate, against the SpMV
// kernels' own gather count, is the bound those kernels are held to; it has
// no counterpart in the reference.
//
// Paths measured:
//   vector: each lane loads its own x[index]; U independent loads in
//           flight per lane before their values are used;
//   scalar: each wave loads one wave-uniform x[index] per load (the
//           index made uniform with readfirstlane), through the scalar
//           data cache instead of the vector L1;
//   mixed:  both in the same wave, to see whether the two paths add up.
// Every index is the top log2(entries) bits of a 32-bit value, so every
// load is in bounds.

#include <climits>

#include "common.h"
#include "device.h"

namespace taco_hip {
namespace {

constexpr int kCeilThreads = 256;

__device__ __forceinline__ uint32_t mix32(uint32_t a) {
  a ^= a >> 16;
  a *= 0x7feb352du;
  a ^= a >> 15;
  a *= 0x846ca68bu;
  a ^= a >> 16;
  return a;
}

// Index streams: a 32-bit linear congruential generator per lane (vector)
// or per wave (scalar, kept in SGPRs), its top bits the index — two integer
// operations per load, so neither path is held up by index arithmetic.
__device__ __forceinline__ uint32_t lcg(uint32_t h) { return h * 1664525u + 1013904223u; }

template <int U, int S>
__global__ __launch_bounds__(kCeilThreads) void k_gather(const double* __restrict__ x,
                                                         uint32_t shift, int32_t iters,
                                                         uint32_t seed,
                                                         double* __restrict__ out) {
  const uint32_t gid = blockIdx.x * kCeilThreads + threadIdx.x;
  uint32_t h = mix32(gid ^ seed);
  // wave-uniform stream for the scalar loads
  uint32_t hs = __builtin_amdgcn_readfirstlane(mix32((gid / kWave) ^ (seed * 0x9e3779b9u)));
  double acc = 0.0, accs = 0.0;
  for (int32_t it = 0; it < iters; ++it) {
    double v[U > 0 ? U : 1];
    double w[S > 0 ? S : 1];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      h = lcg(h);
      v[u] = x[h >> shift];
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      hs = __builtin_amdgcn_readfirstlane(lcg(hs));
      w[s] = x[hs >> shift];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u];
#pragma unroll
    for (int s = 0; s < S; ++s) accs += w[s];
  }
  out[gid] = acc + accs;
}

typedef void (*GatherKernel)(const double*, uint32_t, int32_t, uint32_t, double*);

GatherKernel pick(int mode, int unroll) {
  // mode 0 vector (U = unroll), 1 scalar (S = unroll), 2 mixed (U = S = unroll)
#define TH_G(U, S) return &k_gather<U, S>
  if (mode == 0) {
    switch (unroll) {
      case 1: TH_G(1, 0); case 2: TH_G(2, 0); case 4: TH_G(4, 0);
      case 8: TH_G(8, 0); case 16: TH_G(16, 0); default: return nullptr;
    }
  }
  if (mode == 1) {
    switch (unroll) {
      case 1: TH_G(0, 1); case 2: TH_G(0, 2); case 4: TH_G(0, 4);
      case 8: TH_G(0, 8); case 16: TH_G(0, 16); default: return nullptr;
    }
  }
  if (mode == 2) {
    switch (unroll) {
      case 1: TH_G(1, 1); case 2: TH_G(2, 2); case 4: TH_G(4, 4);
      case 8: TH_G(8, 8); default: return nullptr;
    }
  }
#undef TH_G
  return nullptr;
}

}  // namespace
}  // namespace taco_hip

using namespace taco_hip;

// Times random 8-byte gathers (see the file comment).  mode: 0 vector,
// 1 scalar, 2 mixed; unroll: loads in flight per lane (vector) or per wave
// (scalar); log2_entries: the array holds 2^log2_entries f64 (an 8 KB ..
// 256 MB footprint); wg_per_cu: workgroups of 256 threads per CU;
// target_loads: about how many loads one launch makes (lane loads plus
// wave-level scalar loads).  out[0] = mean ms per launch over `reps`,
// out[1] = vector loads per launch, out[2] = scalar loads per launch.
extern "C" int taco_hip_gather_ceiling(int32_t mode, int32_t unroll, int32_t log2_entries,
                                       int32_t wg_per_cu, int64_t target_loads,
                                       int32_t reps, double* out) {
  TH_REQUIRE(out && reps > 0 && wg_per_cu > 0 && wg_per_cu <= 32 && target_loads > 0,
             "gather ceiling: bad arguments");
  TH_REQUIRE(log2_entries >= 10 && log2_entries <= 25, "gather ceiling: log2_entries in 10..25");
  GatherKernel k = pick(mode, unroll);
  TH_REQUIRE(k != nullptr, "gather ceiling: unsupported mode %d / unroll %d", mode, unroll);
  int dev = 0, cus = 256;
  TH_CHECK(hipGetDevice(&dev));
  TH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t grid = (int64_t)cus * wg_per_cu;
  const int64_t threads = grid * kCeilThreads;
  const int64_t waves = threads / kWave;
  // loads per iteration of the whole grid
  const int64_t per_iter = (mode == 1 ? 0 : threads * unroll) + (mode == 0 ? 0 : waves * unroll);
  int64_t iters = (target_loads + per_iter - 1) / per_iter;
  if (iters < 1) iters = 1;
  TH_REQUIRE(iters <= INT32_MAX, "gather ceiling: too many iterations");
  const size_t entries = (size_t)1 << log2_entries;
  hipStream_t st = current_stream();
  double* x = nullptr;
  double* o = nullptr;
  TH_CHECK(hipMalloc((void**)&x, entries * sizeof(double)));
  if (hipMalloc((void**)&o, (size_t)threads * sizeof(double)) != hipSuccess) {
    (void)hipFree(x);
    return fail("gather ceiling: allocation failed");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMemsetAsync(x, 0, entries * sizeof(double), st) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail("gather ceiling: set-up failed");
  } else {
    const uint32_t shift = 32u - (uint32_t)log2_entries;  // index = top bits, < entries
    // warm-up (also pulls the array into the caches it fits in)
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(kCeilThreads), 0, st, x, shift,
                       (int32_t)iters, 1u, o);
    (void)hipEventRecord(e0, st);
    for (int32_t r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(kCeilThreads), 0, st, x, shift,
                         (int32_t)iters, 2u + (uint32_t)r, o);
    (void)hipEventRecord(e1, st);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail("gather ceiling: launch failed");
    } else {
      out[0] = (double)ms / reps;
      out[1] = mode == 1 ? 0.0 : (double)threads * iters * unroll;
      out[2] = mode == 0 ? 0.0 : (double)waves * iters * unroll;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(x);
  (void)hipFree(o);
  return rc;
}
