// CodeGen_HIP: taco's GPU code generator for MI355X (gfx950, wave64),
// printing taco's own lowered IR (taco::ir) as a HIP module.
//
// Drop-in counterpart of CodeGen_CUDA (reference src/codegen/codegen_cuda.h,
// codegen_cuda.cpp): it plugs into taco's CodeGen::init_default
// (src/codegen/codegen.cpp:25-32) and keeps the module ABI — functions
// `int f(taco_tensor_t* results..., taco_tensor_t* operands...)` and their
// `_shim_f(void**)` wrappers — so taco's Module and TensorBase call it
// unchanged.  The design is MI355X-first rather than a translation:
//  * every tensor property is unpacked once on the host and passed to the
//    kernels by value (pointers and sizes), so neither descriptors nor host
//    scalars need device-visible memory;
//  * one kernel per GPUBlock loop, `__launch_bounds__` of the block's thread
//    count, thread/group ids for 64-lane waves;
//  * launches go to the library stream (taco_hip_get_stream) without a device
//    synchronise after each; the host waits only before host code that may
//    touch memory a kernel wrote, and at the end of the call;
//  * float atomics come from taco_hip_rt.h (hardware atomics on device
//    memory, compare-and-swap loops under unified memory); `*=` atomics are a
//    real compare-and-swap loop.
#ifndef TACO_HIP_UPSTREAM_CODEGEN_HIP_H
#define TACO_HIP_UPSTREAM_CODEGEN_HIP_H

#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "codegen/codegen.h"
#include "taco/ir/ir.h"

namespace taco {
namespace ir {

class CodeGen_HIP : public CodeGen {
 public:
  CodeGen_HIP(std::ostream& dest, OutputKind outputKind);
  ~CodeGen_HIP();

  /// Print one lowered function (the helper prelude before the first).
  void compile(Stmt stmt, bool isFirst = false) override;

  /// extern "C" int _shim_<f>(void** args): unpacks the argument array into
  /// f's parameters (results, then operands; reference codegen_cuda.cpp:
  /// 1500-1540 for the contract).
  static void generateShim(const Stmt& func, std::stringstream& ret);

  struct Kernel;

 protected:
  using IRPrinter::visit;
  void visit(const Function*) override;
  void visit(const Var*) override;
  void visit(const For*) override;
  void visit(const While*) override;
  void visit(const IfThenElse*) override;
  void visit(const GetProperty*) override;
  void visit(const Min*) override;
  void visit(const Max*) override;
  void visit(const Allocate*) override;
  void visit(const Free*) override;
  void visit(const Sqrt*) override;
  void visit(const VarDecl*) override;
  void visit(const Literal*) override;
  void visit(const Yield*) override;
  void visit(const Call*) override;
  void visit(const Store*) override;
  void visit(const Assign*) override;
  void visit(const Continue*) override;
  void visit(const Break*) override;
  void visit(const Sort*) override;
  void visit(const Print*) override;

 private:
  std::string restrictKeyword() const override { return "__restrict__"; }

  void printKernel(const Function* func, size_t index);
  void printLaunch(size_t index);
  void hostBarrier();
  std::string syncText(bool group) const;
  std::string typeOf(const Expr& e) const;
  std::string nameOf(const Expr& e);

  std::ostream& out;
  OutputKind outputKind;
  std::string funcName;

  // Names of every Var and tensor property of the function being printed.
  std::map<Expr, std::string, ExprCompare> varMap;
  std::vector<Kernel*> kernels;

  // Printing state.
  bool inKernel = false;
  Kernel* cur = nullptr;        // kernel being printed
  bool inThreadScope = false;   // inside the GPUThread loop of `cur`
  bool inGroupScope = false;    // inside the GPUWarp loop of `cur`
  int deviceLoopDepth = 0;      // serial loops open inside the kernel
  bool pendingLaunch = false;   // host code must wait before touching memory
  bool coroutine = false;       // printing a host coroutine (Yield)
  std::vector<Expr> localVars;  // its locals, kept in the context struct
  int labelCount = 0;           // its resume points
};

}  // namespace ir
}  // namespace taco

#endif
