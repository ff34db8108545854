// HIP runtime of the backend: the function set of taco's CUDA runtime
// (reference include/taco/cuda.h:15-28, src/cuda.cpp:11-72) plus the
// descriptor allocation of src/taco_tensor_t.cpp, written for one MI355X with
// device-resident memory by default and managed memory as an option.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace taco_hip {

namespace {

thread_local std::string t_last_error;
thread_local hipStream_t t_stream = nullptr;
thread_local bool t_blocking = false;

// Codegen is on by default when a device is present (taco enables its CUDA
// codegen whenever CUDA is built, cuda.cpp:11-12).
int g_codegen_enabled = -1;
bool g_unified_memory = false;

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};
constexpr int kSlots = 16;
constexpr int kMaxDevices = 16;
std::mutex g_ws_mutex;
Workspace g_ws[kMaxDevices][kSlots];

struct WorkspaceReaper {
  ~WorkspaceReaper() {
    // Process exit: the runtime may already be tearing down, so errors are
    // deliberately ignored here.
    for (auto& dev : g_ws)
      for (auto& w : dev)
        if (w.ptr) (void)hipFree(w.ptr);
  }
} g_reaper;

std::string g_flags;
std::string g_arch;

}  // namespace

int fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_last_error = buf;
  return 1;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[1024];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
           hipGetErrorString(e), what, file, line);
  t_last_error = buf;
  return (int)e != 0 ? (int)e : 1;
}

hipStream_t current_stream() { return t_stream; }
bool blocking() { return t_blocking; }

bool coarse_grained(const void* p) {
  if (!g_unified_memory || !p) return true;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice && !attr.isManaged;
}

int finish(const char* what) {
  TH_LAUNCH_CHECK(what);
  if (t_blocking) TH_CHECK(hipStreamSynchronize(t_stream));
  return 0;
}

void* workspace(size_t bytes, int slot) {
  if (slot < 0 || slot >= kSlots) {
    fail("workspace slot %d out of range", slot);
    return nullptr;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fail("workspace: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Workspace& w = g_ws[dev][slot];
  if (w.bytes >= bytes && w.ptr) return w.ptr;
  if (w.ptr) {
    // The old buffer may still be in use by queued work on any stream.
    if (hipDeviceSynchronize() != hipSuccess || hipFree(w.ptr) != hipSuccess) {
      fail("workspace: failed to release old buffer");
      return nullptr;
    }
    w.ptr = nullptr;
    w.bytes = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  want = (want + 4095) & ~size_t(4095);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    hip_fail(e, "hipMalloc(workspace)", __FILE__, __LINE__);
    return nullptr;
  }
  w.ptr = p;
  w.bytes = want;
  return p;
}

size_t workspace_size(int slot) {
  int dev = 0;
  if (slot < 0 || slot >= kSlots || hipGetDevice(&dev) != hipSuccess || dev < 0 ||
      dev >= kMaxDevices)
    return 0;
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  const Workspace& w = g_ws[dev][slot];
  return w.ptr ? w.bytes : 0;
}

int read_back_i32(const int32_t* src, int n, int32_t* dst, hipStream_t st) {
  TH_REQUIRE(n >= 0 && n <= 64, "read back: %d values", n);
  // Page-locked, so the copy is one DMA of the bytes rather than a staged
  // copy through a driver buffer; kept for the thread's lifetime.
  thread_local int32_t* pinned = nullptr;
  if (!pinned) {
    void* p = nullptr;
    TH_CHECK(hipHostMalloc(&p, 64 * sizeof(int32_t), hipHostMallocDefault));
    pinned = (int32_t*)p;
  }
  TH_CHECK(hipMemcpyAsync(pinned, src, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  std::memcpy(dst, pinned, sizeof(int32_t) * (size_t)n);
  return 0;
}

int view_csr(const taco_tensor_t* t, const char* name, CsrView* out) {
  TH_REQUIRE(t, "%s: null descriptor", name);
  TH_REQUIRE(t->order == 2, "%s: expected a matrix, order is %d", name,
             t->order);
  TH_REQUIRE(t->mode_types && t->mode_types[0] == taco_mode_dense &&
                 t->mode_types[1] == taco_mode_sparse,
             "%s: expected CSR levels {dense, compressed}", name);
  TH_REQUIRE(t->mode_ordering && t->mode_ordering[0] == 0 &&
                 t->mode_ordering[1] == 1,
             "%s: expected mode ordering {0,1}", name);
  TH_REQUIRE(t->indices && t->indices[1], "%s: missing index arrays", name);
  out->m = t->dimensions[0];
  out->n = t->dimensions[1];
  TH_REQUIRE(out->m >= 0 && out->n >= 0, "%s: negative dimension", name);
  out->pos = (const int32_t*)t->indices[1][0];
  out->crd = (const int32_t*)t->indices[1][1];
  out->vals = t->vals;
  out->vt = vtype_of(t->csize);
  TH_REQUIRE(out->vt != VType::Invalid, "%s: unsupported csize %d", name,
             t->csize);
  TH_REQUIRE(out->pos != nullptr || out->m == 0, "%s: null pos", name);
  if (t->vals_size > 0) {
    out->nnz = t->vals_size;
  } else if (out->m == 0) {
    out->nnz = 0;
  } else {
    int32_t last = 0;
    TH_CHECK(hipMemcpy(&last, out->pos + out->m, sizeof(int32_t),
                       hipMemcpyDeviceToHost));
    out->nnz = last;
  }
  TH_REQUIRE(out->nnz >= 0, "%s: negative nnz", name);
  TH_REQUIRE(out->nnz == 0 || (out->crd && out->vals),
             "%s: null crd/vals with nnz %lld", name, (long long)out->nnz);
  return 0;
}

int view_dense(const taco_tensor_t* t, const char* name, int order,
               DenseView* out) {
  TH_REQUIRE(t, "%s: null descriptor", name);
  TH_REQUIRE(t->order == order, "%s: expected order %d, got %d", name, order,
             t->order);
  for (int k = 0; k < order; ++k) {
    TH_REQUIRE(t->mode_types[k] == taco_mode_dense, "%s: level %d not dense",
               name, k);
    TH_REQUIRE(t->mode_ordering[k] == k,
               "%s: only row-major (identity) ordering is supported", name);
    out->dims[k] = t->dimensions[k];
    TH_REQUIRE(out->dims[k] >= 0, "%s: negative dimension", name);
  }
  out->order = order;
  out->vals = t->vals;
  out->vt = vtype_of(t->csize);
  TH_REQUIRE(out->vt != VType::Invalid, "%s: unsupported csize %d", name,
             t->csize);
  return 0;
}

}  // namespace taco_hip

using namespace taco_hip;

extern "C" {

int taco_hip_should_use_codegen(void) {
  if (g_codegen_enabled < 0) g_codegen_enabled = taco_hip_device_count() > 0;
  return g_codegen_enabled;
}

void taco_hip_set_codegen_enabled(int enabled) {
  g_codegen_enabled = enabled ? 1 : 0;
}

int taco_hip_should_use_unified_memory(void) { return g_unified_memory; }

void taco_hip_set_unified_memory_enabled(int enabled) {
  g_unified_memory = enabled != 0;
}

const char* taco_hip_device_arch(void) {
  if (!g_arch.empty()) return g_arch.c_str();
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    return "gfx950";  // the only architecture this backend targets
  }
  std::string name = prop.gcnArchName;
  size_t colon = name.find(':');
  g_arch = name.substr(0, colon);
  return g_arch.c_str();
}

const char* taco_hip_get_default_compiler_flags(void) {
  // Counterpart of nvcc's "-w -O3 -Xcompiler -fPIC -shared -ffast-math -O3
  // --generate-code arch=compute_XY,code=sm_XY" (cuda.cpp:30-40).  Packed
  // FP32 instructions are disabled: hipcc forms them from plain float code.
  // With unified memory on, kernels may write managed memory, where the
  // hardware float atomics may not take effect: generated atomics then use
  // compare-and-swap loops (TACO_HIP_COHERENT_ATOMICS).
  g_flags = std::string("-w -O3 -fPIC -shared -std=c++17 --offload-arch=") +
            taco_hip_device_arch() +
            " -Xclang -target-feature -Xclang -packed-fp32-ops" +
            (g_unified_memory ? " -DTACO_HIP_COHERENT_ATOMICS"
                              : " -munsafe-fp-atomics");
  return g_flags.c_str();
}

// Without any GPU (code generation only: modules are printed and compiled,
// nothing can be launched), allocations are host memory, so that taco's
// host-side packing code still runs.
static bool has_device() {
  static int n = -1;
  if (n < 0) n = taco_hip_device_count();
  return n > 0;
}

// Cache of large device blocks freed through taco_hip_unified_free.  taco
// allocates every output array anew on each assemble (SpGEMM's C is 384 MB
// at 32M entries), and hipMalloc/hipFree of blocks that size cost 60-300 us
// each.  Blocks of 1 MiB and more (sizes rounded up to 2 MiB) are kept, per
// device, up to kCacheBlocks blocks and kCacheBytes bytes, and handed out
// again for an allocation of the same rounded size.  A free synchronizes
// the device first, as hipFree does, so no kernel still reads a block that
// is handed out again.  Managed allocations are never cached.
namespace {
constexpr size_t kCacheMinBytes = size_t(1) << 20;
constexpr size_t kCacheGranule = size_t(2) << 20;
constexpr size_t kCacheBlocks = 32;
constexpr size_t kCacheBytes = size_t(32) << 30;

struct DeviceBlock {
  void* p;
  size_t bytes;
  int device;
};

std::mutex g_cache_mu;
std::vector<DeviceBlock> g_live;    // cacheable blocks handed out
std::vector<DeviceBlock> g_cached;  // freed blocks kept for reuse
size_t g_cached_bytes = 0;

size_t cache_round(size_t size) {
  return (size + kCacheGranule - 1) / kCacheGranule * kCacheGranule;
}

void cache_release_all() {  // caller holds g_cache_mu
  for (const DeviceBlock& b : g_cached) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(b.device);
    (void)hipFree(b.p);
    (void)hipSetDevice(cur);
  }
  g_cached.clear();
  g_cached_bytes = 0;
}
}  // namespace

void* taco_hip_unified_alloc(size_t size) {
  if (size == 0) return nullptr;
  if (!has_device()) return std::malloc(size);
  void* p = nullptr;
  if (g_unified_memory) {
    hipError_t e = hipMallocManaged(&p, size);
    if (e != hipSuccess) {
      hip_fail(e, "taco_hip_unified_alloc", __FILE__, __LINE__);
      return nullptr;
    }
    return p;
  }
  if (size < kCacheMinBytes) {
    hipError_t e = hipMalloc(&p, size);
    if (e != hipSuccess) {
      hip_fail(e, "taco_hip_unified_alloc", __FILE__, __LINE__);
      return nullptr;
    }
    return p;
  }
  const size_t bytes = cache_round(size);
  int dev = 0;
  if (hipError_t e = hipGetDevice(&dev); e != hipSuccess) {
    hip_fail(e, "taco_hip_unified_alloc", __FILE__, __LINE__);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_cache_mu);
  for (size_t k = g_cached.size(); k-- > 0;) {  // most recently freed first
    if (g_cached[k].bytes == bytes && g_cached[k].device == dev) {
      const DeviceBlock b = g_cached[k];
      g_cached.erase(g_cached.begin() + (std::ptrdiff_t)k);
      g_cached_bytes -= b.bytes;
      g_live.push_back(b);
      return b.p;
    }
  }
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess && !g_cached.empty()) {
    (void)hipGetLastError();  // out of memory: return the cache, try again
    cache_release_all();
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    hip_fail(e, "taco_hip_unified_alloc", __FILE__, __LINE__);
    return nullptr;
  }
  g_live.push_back(DeviceBlock{p, bytes, dev});
  return p;
}

void taco_hip_unified_free(void* ptr) {
  if (!ptr) return;
  if (!has_device()) {
    std::free(ptr);
    return;
  }
  {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    for (size_t k = 0; k < g_live.size(); ++k) {
      if (g_live[k].p != ptr) continue;
      const DeviceBlock b = g_live[k];
      g_live.erase(g_live.begin() + (std::ptrdiff_t)k);
      int cur = 0;
      hipError_t e = hipGetDevice(&cur);
      if (e == hipSuccess) e = hipSetDevice(b.device);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      (void)hipSetDevice(cur);
      if (e != hipSuccess) {
        hip_fail(e, "taco_hip_unified_free", __FILE__, __LINE__);
        return;
      }
      if (g_cached.size() < kCacheBlocks && g_cached_bytes + b.bytes <= kCacheBytes) {
        g_cached.push_back(b);
        g_cached_bytes += b.bytes;
        return;
      }
      break;  // the cache is full: free the block itself
    }
  }
  hipError_t e = hipFree(ptr);
  if (e != hipSuccess) hip_fail(e, "taco_hip_unified_free", __FILE__, __LINE__);
}

void* taco_hip_unified_realloc(void* ptr, size_t old_bytes, size_t new_bytes) {
  void* q = taco_hip_unified_alloc(new_bytes);
  if (!q && new_bytes > 0) return nullptr;
  const size_t n = old_bytes < new_bytes ? old_bytes : new_bytes;
  if (ptr && n > 0) {
    if (!has_device()) {
      std::memcpy(q, ptr, n);
    } else if (hipMemcpyAsync(q, ptr, n, hipMemcpyDefault, t_stream) != hipSuccess ||
               hipStreamSynchronize(t_stream) != hipSuccess) {
      hip_fail(hipGetLastError(), "taco_hip_unified_realloc", __FILE__, __LINE__);
      taco_hip_unified_free(q);
      return nullptr;
    }
  }
  taco_hip_unified_free(ptr);
  return q;
}

int taco_hip_unified_clear(void* ptr, size_t bytes) {
  if (!ptr || bytes == 0) return 0;
  if (!has_device()) {
    std::memset(ptr, 0, bytes);
    return 0;
  }
  TH_CHECK(hipMemsetAsync(ptr, 0, bytes, t_stream));
  TH_CHECK(hipStreamSynchronize(t_stream));
  return 0;
}

const char* taco_hip_last_error(void) { return t_last_error.c_str(); }

int taco_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int taco_hip_set_stream(void* stream) {
  t_stream = (hipStream_t)stream;
  return 0;
}

void* taco_hip_get_stream(void) { return (void*)t_stream; }

void taco_hip_set_blocking(int b) { t_blocking = b != 0; }

int taco_hip_synchronize(void) {
  TH_CHECK(hipStreamSynchronize(t_stream));
  return 0;
}

int taco_hip_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  if (bytes == 0) return 0;
  hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost
                                : hipMemcpyDeviceToDevice;
  TH_CHECK(hipMemcpyAsync(dst, src, bytes, k, t_stream));
  TH_CHECK(hipStreamSynchronize(t_stream));
  return 0;
}

int taco_hip_memset(void* dst, int value, size_t bytes) {
  if (bytes == 0) return 0;
  TH_CHECK(hipMemsetAsync(dst, value, bytes, t_stream));
  TH_CHECK(hipStreamSynchronize(t_stream));
  return 0;
}

// ---- descriptor allocation (reference src/taco_tensor_t.cpp:32-79) ----
// Descriptors are host structures on this backend: kernels receive raw
// device pointers unpacked on the host, so the descriptor never needs to be
// device-visible.

taco_tensor_t* init_taco_tensor_t(int32_t order, int32_t csize,
                                  int32_t* dimensions, int32_t* modeOrdering,
                                  taco_mode_t* mode_types, void* fill_ptr) {
  taco_tensor_t* t = (taco_tensor_t*)calloc(1, sizeof(taco_tensor_t));
  if (!t) return nullptr;
  t->order = order;
  t->csize = csize;
  size_t n = order > 0 ? (size_t)order : 1;
  t->dimensions = (int32_t*)calloc(n, sizeof(int32_t));
  t->mode_ordering = (int32_t*)calloc(n, sizeof(int32_t));
  t->mode_types = (taco_mode_t*)calloc(n, sizeof(taco_mode_t));
  t->indices = (uint8_t***)calloc(n, sizeof(uint8_t**));
  int fill_bytes = csize / 8 > 0 ? csize / 8 : 1;
  t->fill_value = (uint8_t*)calloc((size_t)fill_bytes, 1);
  if (fill_ptr) memcpy(t->fill_value, fill_ptr, (size_t)(csize / 8));
  for (int32_t i = 0; i < order; i++) {
    t->dimensions[i] = dimensions[i];
    t->mode_ordering[i] = modeOrdering[i];
    t->mode_types[i] = mode_types[i];
    t->indices[i] = (uint8_t**)calloc(
        t->mode_types[i] == taco_mode_sparse ? 2 : 1, sizeof(uint8_t*));
  }
  return t;
}

void deinit_taco_tensor_t(taco_tensor_t* t) {
  if (!t) return;
  for (int i = 0; i < t->order; i++) free(t->indices[i]);
  free(t->indices);
  free(t->dimensions);
  free(t->mode_ordering);
  free(t->mode_types);
  free(t->fill_value);
  free(t);
}

}  // extern "C"

// ---- support for generated modules ----

extern "C" {

int taco_hip_module_report(int code, const char* what) {
  taco_hip::fail("generated module: %s (code %d)", what ? what : "", code);
  return code != 0 ? code : 1;
}

int taco_hip_module_finish(int rc) {
  if (rc != 0) return rc;
  return taco_hip::finish("generated module");
}

int taco_hip_host_read_bytes(void* dst, const void* src, size_t bytes) {
  hipStream_t st = taco_hip::current_stream();
  TH_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  return 0;
}

void* taco_hip_workspace_get(size_t bytes, int slot) {
  return taco_hip::workspace(bytes, slot);
}

}  // extern "C"
