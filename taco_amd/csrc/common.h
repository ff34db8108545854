// Shared helpers of libtaco_hip: error reporting, the current stream,
// descriptor views and the per-device workspace cache.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "taco_hip.h"

namespace taco_hip {

// Record a failure message for taco_hip_last_error(); returns 1.
int fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
// Record a HIP failure; returns the HIP error code (non-zero).
int hip_fail(hipError_t e, const char* what, const char* file, int line);

hipStream_t current_stream();
bool blocking();
// True when hardware float atomics are valid on `p`: always with unified
// memory off (entry points then take hipMalloc device memory), otherwise
// only for plain device allocations.  Kernels that add into an output with
// float atomics take the answer as their `coarse` argument.
bool coarse_grained(const void* p);
// Synchronise the stream when blocking mode is on; reports launch errors.
int finish(const char* what);

// Device workspace reused across calls; slot picks one of a few independent
// buffers so one entry point can hold several at once.  Grows on demand,
// never shrinks; freed at process exit.  Returns nullptr on failure (the
// error is recorded).
void* workspace(size_t bytes, int slot);
// Bytes currently held by a workspace slot on the current device (0: none).
size_t workspace_size(int slot);
// Reads n <= 64 int32 values from device memory `src` into `dst` with one
// copy through a page-locked host buffer (one per host thread), after the
// stream's queued work.  Returns 0 or an error code (recorded).
int read_back_i32(const int32_t* src, int n, int32_t* dst, hipStream_t st);
// Zeroes `bytes` (a multiple of 4) at the 4-byte-aligned p with a kernel on
// st: unlike hipMemsetAsync, safe inside a HIP graph capture.
int zero_fill(void* p, size_t bytes, hipStream_t st);

// Value types carried by descriptors (csize, in bits).
enum class VType { F64, F32, BF16, Invalid };
inline VType vtype_of(int32_t csize) {
  switch (csize) {
    case 64: return VType::F64;
    case 32: return VType::F32;
    case 16: return VType::BF16;
    default: return VType::Invalid;
  }
}

// A CSR operand: order 2, levels {dense, compressed}, identity ordering.
struct CsrView {
  int32_t m = 0, n = 0;
  int64_t nnz = 0;
  const int32_t* pos = nullptr;
  const int32_t* crd = nullptr;
  const void* vals = nullptr;
  VType vt = VType::Invalid;
};
// Fills `out` from a descriptor.  nnz comes from vals_size when positive,
// otherwise pos[m] is read from the device (a synchronising copy).
int view_csr(const taco_tensor_t* t, const char* name, CsrView* out);

// A dense operand of the given order (row-major, identity ordering).
struct DenseView {
  int32_t order = 0;
  int32_t dims[2] = {0, 0};
  void* vals = nullptr;
  VType vt = VType::Invalid;
};
int view_dense(const taco_tensor_t* t, const char* name, int order,
               DenseView* out);

}  // namespace taco_hip

#define TH_CHECK(call)                                                   \
  do {                                                                   \
    hipError_t th_e_ = (call);                                           \
    if (th_e_ != hipSuccess)                                             \
      return ::taco_hip::hip_fail(th_e_, #call, __FILE__, __LINE__);     \
  } while (0)

#define TH_LAUNCH_CHECK(what)                                            \
  do {                                                                   \
    hipError_t th_e_ = hipGetLastError();                                \
    if (th_e_ != hipSuccess)                                             \
      return ::taco_hip::hip_fail(th_e_, what, __FILE__, __LINE__);      \
  } while (0)

#define TH_REQUIRE(cond, ...)                                            \
  do {                                                                   \
    if (!(cond)) return ::taco_hip::fail(__VA_ARGS__);                   \
  } while (0)
