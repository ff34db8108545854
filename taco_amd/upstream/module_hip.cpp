// taco::ir::Module for the HIP backend (the class of reference
// include/taco/codegen/module.h, whose src/codegen/module.cpp:33-215 this
// file replaces).  Same contract: functions added, source printed by the
// default code generator, compiled to a shared object, dlopen'ed, called
// through _shim_<name>.  GPU modules go through libtaco_hip's module loader
// (hipcc run with posix_spawn, objects cached by a hash of source, compiler,
// flags and runtime header), so identical kernels are compiled once and the
// cache can be filled on a machine without a GPU.  CPU modules (GPU codegen
// switched off) keep taco's C compiler path.

#include <dlfcn.h>
#include <unistd.h>

#include <cstdlib>
#include <fstream>
#include <iostream>

#include "codegen/codegen.h"
#include "codegen/codegen_c.h"
#include "codegen_hip.h"
#include "taco/codegen/module.h"
#include "taco/cuda.h"
#include "taco/error.h"
#include "taco/util/env.h"
#include "taco_hip.h"

namespace taco {
namespace ir {

std::string Module::chars = "abcdefghijkmnpqrstuvwxyz0123456789";
std::default_random_engine Module::gen = std::default_random_engine();
std::uniform_int_distribution<int> Module::randint =
    std::uniform_int_distribution<int>(0, Module::chars.length() - 1);

void Module::setJITTmpdir() { tmpdir = util::getTmpdir(); }

void Module::setJITLibname() {
  libname.resize(12);
  for (int i = 0; i < 12; i++) libname[i] = chars[randint(gen)];
}

void Module::addFunction(Stmt func) { funcs.push_back(func); }

void Module::compileToSource(std::string path, std::string prefix) {
  if (!moduleFromUserSource) {
    header.str("");
    header.clear();
    source.str("");
    source.clear();
    taco_tassert(target.arch == Target::C99) << "Only C99 codegen supported currently";
    std::shared_ptr<CodeGen> sourcegen =
        CodeGen::init_default(source, CodeGen::ImplementationGen);
    std::shared_ptr<CodeGen> headergen =
        CodeGen::init_default(header, CodeGen::HeaderGen);
    bool first = true;
    for (auto& func : funcs) {
      sourcegen->compile(func, first);
      headergen->compile(func, first);
      first = false;
    }
  }
  const bool gpu = should_use_CUDA_codegen();
  std::ofstream src(path + prefix + (gpu ? ".hip" : ".c"));
  src << source.str();
  std::ofstream hdr(path + prefix + ".h");
  hdr << header.str();
}

void Module::compileToStaticLibrary(std::string, std::string) {
  taco_tassert(false) << "Compiling to a static library is not supported";
}

namespace {

std::string shims_of(const std::vector<Stmt>& funcs, bool gpu) {
  std::stringstream shims;
  for (auto& f : funcs) {
    if (gpu)
      CodeGen_HIP::generateShim(f, shims);
    else
      CodeGen_C::generateShim(f, shims);
  }
  return shims.str();
}

}  // namespace

std::string Module::compile() {
  const std::string prefix = tmpdir + libname;
  compileToSource(tmpdir, libname);
  if (should_use_CUDA_codegen()) {
    // One translation unit: definitions, then the shims that call them.
    std::string text = source.str() + "\n" + shims_of(funcs, true);
    taco_hip_module* m = taco_hip_module_create();
    taco_uassert(m) << "cannot create a HIP module";
    int rc = taco_hip_module_set_source(m, text.c_str());
    if (rc == 0) rc = taco_hip_module_compile(m);
    if (rc != 0) {
      std::string err = taco_hip_last_error();
      taco_hip_module_destroy(m);
      taco_uerror << "HIP module compilation failed: " << err;
    }
    std::string path = taco_hip_module_path(m);
    if (lib_handle) dlclose(lib_handle);
    // A handle of our own keeps the object loaded past the loader's.
    lib_handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    taco_hip_module_destroy(m);
    taco_uassert(lib_handle) << "Failed to load generated code: " << dlerror();
    return path;
  }
  // CPU code: taco's C path (reference module.cpp:138-180).
  {
    std::ofstream shims(prefix + ".c", std::ios::app);
    shims << "#include \"" << prefix << ".h\"\n" << shims_of(funcs, false);
  }
  std::string cc = util::getFromEnv(target.compiler_env, target.compiler);
  std::string cflags = util::getFromEnv("TACO_CFLAGS", "-O3 -ffast-math -std=c99") +
                       " -shared -fPIC";
#if USE_OPENMP
  cflags += " -fopenmp";
#endif
  const std::string fullpath = prefix + ".so";
  std::string cmd = cc + " " + cflags + " " + prefix + ".c -o " + fullpath + " -lm";
  int err = system(cmd.c_str());
  taco_uassert(err == 0) << "Compilation command failed:\n" << cmd << "\nreturned " << err;
  if (lib_handle) dlclose(lib_handle);
  lib_handle = dlopen(fullpath.c_str(), RTLD_NOW | RTLD_LOCAL);
  taco_uassert(lib_handle) << "Failed to load generated code, error is: " << dlerror();
  return fullpath;
}

void Module::setSource(std::string src) {
  source << src;
  moduleFromUserSource = true;
}

std::string Module::getSource() { return source.str(); }

void* Module::getFuncPtr(std::string name) {
  return lib_handle ? dlsym(lib_handle, name.c_str()) : nullptr;
}

int Module::callFuncPackedRaw(std::string name, void** args) {
  void* f = getFuncPtr(name);
  taco_uassert(f) << "no function " << name << " in the generated module";
  typedef int (*fnptr_t)(void**);
  fnptr_t fn;
  *reinterpret_cast<void**>(&fn) = f;
  return fn(args);
}

}  // namespace ir
}  // namespace taco
