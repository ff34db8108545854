// CodeGen_HIP: prints IR functions as a HIP C++ module for gfx950.
//
// Counterpart of taco's CodeGen_CUDA (reference src/codegen/codegen_cuda.cpp)
// with the same module contract — a host function per IR Function with the
// signature int f(taco_tensor_t* outputs..., taco_tensor_t* inputs...), and
// extern "C" int _shim_f(void**) wrappers (codegen_cuda.cpp:1500-1540) — but
// written for wave64 CDNA:
//   * each For{GPUBlock} becomes a __global__ __launch_bounds__ kernel whose
//     free variables are passed by value (no managed-memory scalars,
//     codegen_cuda.cpp:1182-1212);
//   * GPUWarp means a lane group whose size is the thread-loop extent (taco's
//     own reading, codegen_cuda.cpp:703/723), any power of two up to 64;
//   * GPUWarpReduction temporaries reduce with a shuffle tree, not lane 0's
//     serial loop (codegen_cuda.cpp:875-890);
//   * atomics relative to GPUWarp use a 64-lane aggregated add (the wave64
//     form of atomicAddWarp, codegen_cuda.cpp:142-159);
//   * launches go to the library's current stream with no device-wide sync
//     after each launch (codegen_cuda.cpp:609); the host function syncs the
//     stream once at its end only in blocking mode;
//   * zero-fills use a fill kernel (graph-capture safe), never a memset node.
#pragma once

#include <string>
#include <vector>

#include "ir.h"

namespace taco_hip {

class CodeGenHIP {
 public:
  // Prints the module prelude followed by every function, then the shims.
  std::string compile(const std::vector<ir::Stmt>& functions);
};

}  // namespace taco_hip
