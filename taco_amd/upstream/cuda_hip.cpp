// taco's GPU runtime functions (reference include/taco/cuda.h:15-28,
// src/cuda.cpp:11-72) over libtaco_hip, and the code generator selection
// (CodeGen::init_default, src/codegen/codegen.cpp:25-32) that returns
// CodeGen_HIP when GPU code generation is on.
//
// As in taco with its GPU backend built, GPU code generation and unified
// memory start enabled: taco then allocates every tensor array with
// cuda_unified_alloc (src/storage/array.cpp:212-219) and writes it from the
// host, so the arrays must be host-accessible (managed memory).  Calls are
// synchronous at the API boundary (taco synchronises after every launch,
// codegen_cuda.cpp:609): the library's blocking mode is switched on.
//
// On a machine without a GPU the allocations fall back to host memory so
// that statements can still be lowered and their modules compiled (taco's
// CPU-only build prints GPU code the same way); nothing can be launched
// there, and a generated call reports the missing device.

#include <cstdlib>
#include <memory>
#include <mutex>

#include "codegen/codegen.h"
#include "codegen/codegen_c.h"
#include "codegen_hip.h"
#include "taco/cuda.h"
#include "taco/error.h"
#include "taco_hip.h"

namespace taco {

namespace {
bool g_codegen = true;
bool g_unified = true;
std::once_flag g_once;
int g_devices = 0;

void ensure_init() {
  std::call_once(g_once, [] {
    g_devices = taco_hip_device_count();
    taco_hip_set_unified_memory_enabled(g_unified ? 1 : 0);
    taco_hip_set_codegen_enabled(g_codegen ? 1 : 0);
    taco_hip_set_blocking(1);
  });
}
}  // namespace

bool should_use_CUDA_codegen() {
  ensure_init();
  return g_codegen;
}

bool should_use_CUDA_unified_memory() {
  ensure_init();
  return g_unified;
}

void set_CUDA_codegen_enabled(bool enabled) {
  ensure_init();
  g_codegen = enabled;
  taco_hip_set_codegen_enabled(enabled ? 1 : 0);
}

void set_CUDA_unified_memory_enabled(bool enabled) {
  ensure_init();
  g_unified = enabled;
  taco_hip_set_unified_memory_enabled(enabled ? 1 : 0);
}

std::string get_default_CUDA_compiler_flags() {
  ensure_init();
  return taco_hip_get_default_compiler_flags();
}

void* cuda_unified_alloc(size_t size) {
  ensure_init();
  if (size == 0) return nullptr;
  if (g_devices == 0) return std::malloc(size);
  void* p = taco_hip_unified_alloc(size);
  taco_uassert(p) << "unified allocation of " << size
                  << " bytes failed: " << taco_hip_last_error();
  return p;
}

void cuda_unified_free(void* ptr) {
  ensure_init();
  if (!ptr) return;
  if (g_devices == 0) {
    std::free(ptr);
    return;
  }
  taco_hip_unified_free(ptr);
}

namespace ir {

std::shared_ptr<CodeGen> CodeGen::init_default(std::ostream& dest,
                                               OutputKind outputKind) {
  if (should_use_CUDA_codegen())
    return std::make_shared<CodeGen_HIP>(dest, outputKind);
  return std::make_shared<CodeGen_C>(dest, outputKind);
}

}  // namespace ir
}  // namespace taco
