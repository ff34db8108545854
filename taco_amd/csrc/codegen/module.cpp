// JIT module loader and the code-generation C ABI.
//
// Same contract as taco's ir::Module (reference src/codegen/module.cpp:
// 33-215): a module holds source text, compiles it to a shared object with an
// external compiler, dlopens it, and calls functions through their _shim_
// wrappers.  Differences, all MI355X-side: hipcc with gfx950 flags; the
// compiler runs through posix_spawn (no shell fork of a GPU process's
// address space via system()); compiled objects are cached by content hash
// so identical kernels are never rebuilt, and the cache can be filled ahead
// of time on a machine without a GPU.

#include <dlfcn.h>
#include <fcntl.h>
#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../common.h"
#include "codegen_hip.h"
#include "lower.h"

extern char** environ;

namespace taco_hip {
namespace {

std::string env_or(const char* name, const std::string& dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::string(v) : dflt;
}

// Directory holding libtaco_hip.so (the include/ tree sits at ../include).
std::string library_dir() {
  Dl_info info;
  if (dladdr((void*)&taco_hip_module_create, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t slash = p.rfind('/');
    if (slash != std::string::npos) return p.substr(0, slash);
  }
  return ".";
}

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

void mkdirs(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i < path.size(); ++i) {
    cur += path[i];
    if (path[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
  }
  mkdir(path.c_str(), 0755);
}

// Runs argv, returns its exit status (or -1), capturing nothing.
int run(const std::vector<std::string>& args, const std::string& log_path) {
  std::vector<char*> argv;
  for (auto& a : args) argv.push_back(const_cast<char*>(a.c_str()));
  argv.push_back(nullptr);
  posix_spawn_file_actions_t fa;
  posix_spawn_file_actions_init(&fa);
  posix_spawn_file_actions_addopen(&fa, 1, log_path.c_str(),
                                   O_WRONLY | O_CREAT | O_TRUNC, 0644);
  posix_spawn_file_actions_adddup2(&fa, 1, 2);
  pid_t pid;
  int rc = posix_spawnp(&pid, argv[0], &fa, nullptr, argv.data(), environ);
  posix_spawn_file_actions_destroy(&fa);
  if (rc != 0) return -1;
  int status = 0;
  if (waitpid(pid, &status, 0) < 0) return -1;
  return WIFEXITED(status) ? WEXITSTATUS(status) : -1;
}

std::vector<std::string> split_ws(const std::string& s) {
  std::istringstream is(s);
  std::vector<std::string> out;
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

std::string read_file(const std::string& p) {
  std::ifstream f(p);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::mutex g_compile_mutex;

}  // namespace
}  // namespace taco_hip

using namespace taco_hip;

struct taco_hip_module {
  std::string source;
  std::string path;
  void* handle = nullptr;
};

extern "C" {

taco_hip_module* taco_hip_module_create(void) { return new taco_hip_module(); }

int taco_hip_module_set_source(taco_hip_module* m, const char* source) {
  TH_REQUIRE(m && source, "module: null argument");
  m->source = source;
  return 0;
}

const char* taco_hip_module_get_source(const taco_hip_module* m) {
  return m ? m->source.c_str() : "";
}

const char* taco_hip_module_path(const taco_hip_module* m) {
  return m ? m->path.c_str() : "";
}

int taco_hip_module_compile(taco_hip_module* m) {
  TH_REQUIRE(m, "module: null");
  const std::string libdir = library_dir();
  const std::string root = libdir + "/..";
  const std::string cc = env_or("TACO_HIPCC", "/opt/rocm/bin/hipcc");
  const std::string flags =
      env_or("TACO_HIPCCFLAGS", taco_hip_get_default_compiler_flags());
  const std::string cache = env_or("TACO_HIP_CACHE", libdir + "/_jit_cache");
  // The key covers everything the object depends on: the source, the
  // compiler and flags, and the runtime header the source includes.
  const std::string rt = read_file(root + "/include/taco_hip_rt.h") +
                         read_file(root + "/include/taco/taco_tensor_t.h");
  char key[32];
  snprintf(key, sizeof key, "%016llx",
           (unsigned long long)fnv1a(m->source, fnv1a(cc + "\n" + flags + "\n" + rt)));
  const std::string so = cache + "/" + key + ".so";
  std::lock_guard<std::mutex> lock(g_compile_mutex);
  if (!file_exists(so)) {
    mkdirs(cache);
    char tmpl[64];
    snprintf(tmpl, sizeof tmpl, "%s.%d", key, (int)getpid());
    const std::string base = cache + "/" + tmpl;
    const std::string src = base + ".hip", tmp_so = base + ".so.tmp",
                      log = base + ".log";
    {
      std::ofstream f(src);
      f << m->source;
      if (!f) return fail("module: cannot write %s", src.c_str());
    }
    std::vector<std::string> args = {cc};
    for (auto& w : split_ws(flags)) args.push_back(w);
    args.push_back("-I" + root + "/include");
    args.push_back(src);
    args.push_back("-o");
    args.push_back(tmp_so);
    args.push_back("-L" + libdir);
    args.push_back("-ltaco_hip");
    args.push_back("-Wl,-rpath," + libdir);
    int rc = run(args, log);
    if (rc != 0) {
      std::string msg = read_file(log);
      if (msg.size() > 800) msg = msg.substr(0, 800);
      unlink(src.c_str());
      return fail("module: %s failed (exit %d): %s", cc.c_str(), rc, msg.c_str());
    }
    if (rename(tmp_so.c_str(), so.c_str()) != 0)
      return fail("module: cannot move %s into the cache", tmp_so.c_str());
    // Keep the source next to the object for inspection.
    rename(src.c_str(), (cache + "/" + key + ".hip").c_str());
    unlink(log.c_str());
  }
  if (m->handle) dlclose(m->handle);
  m->handle = dlopen(so.c_str(), RTLD_NOW | RTLD_LOCAL);
  TH_REQUIRE(m->handle, "module: dlopen failed: %s", dlerror());
  m->path = so;
  return 0;
}

void* taco_hip_module_get_func(taco_hip_module* m, const char* name) {
  if (!m || !m->handle || !name) return nullptr;
  return dlsym(m->handle, name);
}

int taco_hip_module_call_packed(taco_hip_module* m, const char* name,
                                void** args) {
  TH_REQUIRE(m && m->handle, "module: not compiled");
  std::string shim = std::string("_shim_") + name;
  void* f = dlsym(m->handle, shim.c_str());
  TH_REQUIRE(f, "module: no function %s", shim.c_str());
  typedef int (*fn_t)(void**);
  fn_t fn;
  *reinterpret_cast<void**>(&fn) = f;
  return fn(args);
}

void taco_hip_module_destroy(taco_hip_module* m) {
  if (!m) return;
  if (m->handle) dlclose(m->handle);
  delete m;
}

char* taco_hip_generate(const taco_hip_gen_spec* spec, int32_t what) {
  if (!spec) {
    fail("generate: null spec");
    return nullptr;
  }
  try {
    std::vector<ir::Stmt> fns = lower_spec(*spec);
    std::string text;
    if (what == 1) {
      for (auto& f : fns) text += ir::to_string(f);
    } else {
      text = CodeGenHIP().compile(fns);
    }
    char* out = (char*)std::malloc(text.size() + 1);
    if (!out) return nullptr;
    std::memcpy(out, text.c_str(), text.size() + 1);
    return out;
  } catch (std::exception& e) {
    fail("generate: %s", e.what());
    return nullptr;
  }
}

void taco_hip_free_string(char* s) { std::free(s); }

}  // extern "C"
