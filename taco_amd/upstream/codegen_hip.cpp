// CodeGen_HIP over taco's lowered IR: see codegen_hip.h for the design.
//
// Reference counterparts (behaviour matched, code written anew):
//   kernel extraction per GPUBlock loop, free-variable parameters and the
//   block/group/thread id formulas      codegen_cuda.cpp:280-420, 470-559
//   launch printing                      codegen_cuda.cpp:561-611
//   thread/warp loop elision, reductions codegen_cuda.cpp:861-970
//   shared/local arrays, host alloc     codegen_cuda.cpp:1023-1150
//   atomics                             codegen_cuda.cpp:1335-1498
//   shims                               codegen_cuda.cpp:1500-1540

#include "codegen_hip.h"

#include <algorithm>
#include <functional>
#include <stdexcept>

#include "taco/error.h"
#include "taco/ir/ir_rewriter.h"
#include "taco/ir/ir_visitor.h"
#include "taco/ir/simplify.h"
#include "taco/util/strings.h"

namespace taco {
namespace ir {

namespace {

// Module prelude: the HIP runtime helpers (search, reductions, atomics,
// allocation, error reporting) live in taco_hip_rt.h.
const char* kPrelude =
    "// Emitted by CodeGen_HIP (taco_amd upstream backend) for gfx950.\n"
    "#include \"taco_hip_rt.h\"\n"
    "#include <math.h>\n"
    "#include <stdio.h>\n"
    "#include <stdlib.h>\n"
    "#include <string.h>\n"
    "#define TACO_DEREF(_a) (((___context___*)(*__ctx__))->_a)\n";

[[noreturn]] void unsupported(const std::string& what) {
  throw std::runtime_error("CodeGen_HIP: " + what);
}

int64_t literalInt(Expr e, const char* what) {
  e = ir::simplify(e);
  const Literal* lit = e.as<Literal>();
  if (!lit || !(lit->type.isInt() || lit->type.isUInt()))
    unsupported(std::string(what) + " must be a compile-time integer");
  return lit->getIntValue();
}

// True (and *v set) when e simplifies to an integer literal.
bool literalOrNone(Expr e, int64_t* v) {
  e = ir::simplify(e);
  const Literal* lit = e.as<Literal>();
  if (!lit || !(lit->type.isInt() || lit->type.isUInt())) return false;
  *v = lit->getIntValue();
  return true;
}

// A name with its digits removed.  taco numbers unnamed tensors and index
// variables from one process-wide counter (util::uniqueName), and the names
// of derived variables embed those numbers; printing names without them,
// made unique again per function in first-appearance order, gives the same
// text for the same function in every process, so the module cache (keyed by
// the text) hits across processes.
std::string stripDigits(const std::string& n) {
  std::string out;
  for (char ch : n)
    if (ch < '0' || ch > '9') out += ch;
  if (out.empty() || out == "_") out = "v";
  return out;
}

// Renames a function's tensor parameters to digit-free names.
struct TensorRenamer : public IRRewriter {
  using IRRewriter::visit;
  std::map<const Var*, Expr> repl;
  void visit(const Var* op) override {
    auto it = repl.find(op);
    expr = it != repl.end() ? it->second : Expr(op);
  }
};

Stmt canonicalTensorNames(const Function* f) {
  TensorRenamer r;
  std::set<std::string> used;
  std::vector<Expr> outs, ins;
  auto rename = [&](const Expr& e) {
    const Var* v = e.as<Var>();
    std::string base = stripDigits(v->name), n = base;
    for (int k = 0; used.count(n); ++k) n = base + "_" + std::to_string(k);
    used.insert(n);
    Expr nv = Var::make(n, v->type, v->is_ptr, v->is_tensor, v->is_parameter);
    r.repl[v] = nv;
    return nv;
  };
  for (auto& o : f->outputs) outs.push_back(rename(o));
  for (auto& in : f->inputs) ins.push_back(rename(in));
  return Function::make(f->name, outs, ins, r.rewrite(f->body));
}

bool isGpuUnit(ParallelUnit u) {
  return u == ParallelUnit::GPUBlock || u == ParallelUnit::GPUWarp ||
         u == ParallelUnit::GPUThread;
}

// Host names of every Var and tensor property in a function, as taco's
// declaration printer (CodeGen::printDecls) expects them: one canonical name
// per (tensor, property, mode, index) of an input or output tensor.
struct NameCollector : public IRVisitor {
  using IRVisitor::visit;
  std::map<Expr, std::string, ExprCompare>& names;
  std::map<Expr, std::string, ExprCompare> decls;  // properties to unpack
  std::map<std::tuple<Expr, TensorProperty, int, int>, std::string> canonical;
  std::map<std::tuple<Expr, TensorProperty, int, int>, std::string> outputProps;
  std::vector<Expr> inputs, outputs;
  std::vector<Expr> localVars;  // loop and declared variables, in order
  std::function<std::string(const std::string&)> unique;

  NameCollector(std::map<Expr, std::string, ExprCompare>& names,
                const std::vector<Expr>& ins, const std::vector<Expr>& outs,
                std::function<std::string(const std::string&)> unique)
      : names(names), inputs(ins), outputs(outs), unique(unique) {
    for (auto& v : ins) names[v] = v.as<Var>()->name;
    for (auto& v : outs) names[v] = v.as<Var>()->name;
  }

  // Arrays are matched by name (taco refers to one array through several
  // Var objects of the same name), other variables by object.
  std::map<std::string, std::string> arrayNames;
  void visit(const Var* op) override {
    if (names.count(op)) return;
    if (op->is_ptr) {
      auto it = arrayNames.find(op->name);
      if (it == arrayNames.end())
        it = arrayNames.emplace(op->name, unique(stripDigits(op->name))).first;
      names[op] = it->second;
    } else {
      names[op] = unique(stripDigits(op->name));
    }
  }
  void local(const Expr& v) {
    if (std::find(localVars.begin(), localVars.end(), v) == localVars.end())
      localVars.push_back(v);
  }
  void visit(const For* op) override {
    local(op->var);
    IRVisitor::visit(op);
  }
  void visit(const VarDecl* op) override {
    local(op->var);
    IRVisitor::visit(op);
  }
  void visit(const GetProperty* op) override {
    const bool isOut = std::find(outputs.begin(), outputs.end(), op->tensor) != outputs.end();
    const bool isIn = std::find(inputs.begin(), inputs.end(), op->tensor) != inputs.end();
    if (!isOut && !isIn) return;
    if (names.count(op)) return;
    auto key = std::make_tuple(op->tensor, op->property, op->mode, op->index);
    auto it = canonical.find(key);
    if (it != canonical.end()) {
      names[op] = it->second;
      return;
    }
    std::string n = unique(stripDigits(op->name));
    canonical[key] = n;
    names[op] = n;
    decls[op] = n;
    if (isOut) outputProps[key] = n;
  }
};

}  // namespace

// One kernel: a GPUBlock loop, its GPUWarp loop (optional) and GPUThread
// loop, the launch shape and the values it needs from the host.
struct CodeGen_HIP::Kernel {
  const For* block = nullptr;
  const For* warp = nullptr;
  const For* thread = nullptr;
  // Lanes of the thread loop (a group, taco's "warp") and iterations of the
  // warp loop.  Literal extents give a fixed workgroup size; otherwise the
  // host computes them (threadsExpr, groupsExpr) and passes them in as
  // taco_threads / taco_groups, and the workgroup size is chosen at launch.
  int64_t threads = 0;
  int64_t groups = 1;
  Expr threadsExpr, groupsExpr;
  bool dynamicThreads = false, dynamicGroups = false;
  int64_t blockSize = 0;  // threads * groups when both are literal, else 0
  std::vector<Expr> params;         // Vars and properties, first-use order
  std::vector<std::string> paramNames;
  // Host scalars the kernel assigns (taco passes them by reference,
  // codegen_cuda.cpp:1171-1229): the kernel gets a pointer to a device copy
  // and the host reads the value back after the launch.
  std::set<std::string> byRef;

  std::string threadsText() const {
    return dynamicThreads ? std::string("taco_threads") : std::to_string(threads);
  }
  std::string groupsText() const {
    return dynamicGroups ? std::string("taco_groups") : std::to_string(groups);
  }
  int64_t launchBound() const { return blockSize > 0 ? blockSize : 1024; }
};

namespace {

// Finds the kernels of a function body and their free values.
struct KernelFinder : public IRVisitor {
  using IRVisitor::visit;
  std::vector<CodeGen_HIP::Kernel*>& kernels;
  std::map<Expr, std::string, ExprCompare>& names;
  CodeGen_HIP::Kernel* cur = nullptr;
  std::set<Expr> defined;
  // taco refers to an array through several Var objects of one name (its
  // printers match pointer variables by name): arrays defined inside the
  // kernel are tracked by name as well.
  std::set<std::string> definedArrays;
  std::set<std::string> paramSet;

  void define(const Expr& v) {
    defined.insert(v);
    const Var* var = v.as<Var>();
    if (var && var->is_ptr) definedArrays.insert(var->name);
  }

  KernelFinder(std::vector<CodeGen_HIP::Kernel*>& k,
               std::map<Expr, std::string, ExprCompare>& n)
      : kernels(k), names(n) {}

  void use(const Expr& e, const std::string& name) {
    if (paramSet.count(name)) return;
    paramSet.insert(name);
    cur->params.push_back(e);
    cur->paramNames.push_back(name);
  }

  void visit(const For* op) override {
    if (!cur) {
      if (op->parallel_unit == ParallelUnit::GPUBlock) {
        auto* k = new CodeGen_HIP::Kernel();
        k->block = op;
        kernels.push_back(k);
        cur = k;
        defined.clear();
        definedArrays.clear();
        paramSet.clear();
        define(op->var);
        op->start.accept(this);
        op->increment.accept(this);
        op->contents.accept(this);
        if (!k->thread)
          unsupported("a GPUBlock loop without a GPUThread loop");
        // Extents as taco's GPU printer computes them: (end - start) / inc.
        k->threadsExpr = ir::simplify(
            Div::make(Sub::make(k->thread->end, k->thread->start), k->thread->increment));
        k->dynamicThreads = !literalOrNone(k->threadsExpr, &k->threads);
        if (k->dynamicThreads) k->threadsExpr.accept(this);
        if (k->warp) {
          k->groupsExpr = ir::simplify(
              Div::make(Sub::make(k->warp->end, k->warp->start), k->warp->increment));
          k->dynamicGroups = !literalOrNone(k->groupsExpr, &k->groups);
          if (k->dynamicGroups) k->groupsExpr.accept(this);
        }
        if (!k->dynamicThreads && !k->dynamicGroups) {
          k->blockSize = k->threads * k->groups;
          if (k->threads <= 0 || k->blockSize > 1024)
            unsupported("a workgroup of " + std::to_string(k->blockSize) + " threads");
        }
        cur = nullptr;
        return;
      }
      if (isGpuUnit(op->parallel_unit))
        unsupported("a GPU loop outside a GPUBlock loop");
      IRVisitor::visit(op);
      return;
    }
    // inside a kernel
    if (op->parallel_unit == ParallelUnit::GPUBlock)
      unsupported("nested GPUBlock loops");
    if (op->parallel_unit == ParallelUnit::GPUWarp) {
      if (cur->warp || cur->thread) unsupported("more than one GPUWarp loop in a kernel");
      cur->warp = op;
    } else if (op->parallel_unit == ParallelUnit::GPUThread) {
      if (cur->thread) unsupported("more than one GPUThread loop in a kernel");
      cur->thread = op;
    }
    define(op->var);
    op->start.accept(this);
    op->end.accept(this);
    op->increment.accept(this);
    op->contents.accept(this);
  }

  void visit(const VarDecl* op) override {
    op->rhs.accept(this);
    if (cur) define(op->var);
    else op->var.accept(this);
  }

  void visit(const Allocate* op) override {
    if (cur) define(op->var);
    op->num_elements.accept(this);
    if (op->is_realloc) op->old_elements.accept(this);
    if (!cur) op->var.accept(this);
  }

  void visit(const Assign* op) override {
    if (cur && op->lhs.as<Var>() && !defined.count(op->lhs)) {
      const Var* v = op->lhs.as<Var>();
      if (v->is_ptr) unsupported("an array variable reassigned inside a kernel");
      cur->byRef.insert(names.count(op->lhs) ? names[op->lhs] : v->name);
    }
    IRVisitor::visit(op);
  }

  void visit(const Var* op) override {
    if (!cur || defined.count(op)) return;
    if (op->is_ptr && definedArrays.count(op->name)) return;
    if (op->is_tensor) unsupported("a tensor used directly inside a kernel");
    use(op, names.count(op) ? names[op] : op->name);
  }

  void visit(const GetProperty* op) override {
    if (!cur) return;
    if (!names.count(op)) unsupported("a property of a temporary inside a kernel");
    if (op->property == TensorProperty::FillValue)
      unsupported("a fill value inside a kernel");
    use(op, names[op]);
  }

  void visit(const Yield* op) override {
    if (cur) unsupported("a yield inside a kernel");
    IRVisitor::visit(op);
  }
};

}  // namespace

CodeGen_HIP::CodeGen_HIP(std::ostream& dest, OutputKind kind)
    : CodeGen(dest, false, false, C), out(dest), outputKind(kind) {}

CodeGen_HIP::~CodeGen_HIP() {
  for (auto* k : kernels) delete k;
}

void CodeGen_HIP::compile(Stmt stmt, bool isFirst) {
  if (isFirst) out << kPrelude;
  out << "\n";
  // Simplify function bodies first, as taco's GPU printer does, so that
  // duplicate and dead declarations disappear before kernels are found.
  if (const Function* f = stmt.as<Function>()) {
    // (simplifying a coroutine would disturb its resume points)
    if (countYields(f) == 0)
      stmt = Function::make(f->name, f->outputs, f->inputs, ir::simplify(f->body));
    stmt = canonicalTensorNames(stmt.as<Function>());
  }
  stmt.accept(this);
}

std::string CodeGen_HIP::nameOf(const Expr& e) {
  auto it = varMap.find(e);
  if (it != varMap.end()) return it->second;
  if (const Var* v = e.as<Var>()) return v->name;
  if (const GetProperty* p = e.as<GetProperty>()) return p->name;
  return "?";
}

std::string CodeGen_HIP::typeOf(const Expr& e) const {
  if (const Var* v = e.as<Var>()) return printCUDAType(v->type, v->is_ptr);
  const GetProperty* p = e.as<GetProperty>();
  taco_iassert(p);
  switch (p->property) {
    case TensorProperty::Values:
      return printCUDAType(p->tensor.as<Var>()->type, true);
    case TensorProperty::Indices:
      return "int*";
    default:
      return "int";
  }
}

void CodeGen_HIP::visit(const Function* func) {
  funcName = func->name;
  if (outputKind == HeaderGen) {
    out << "#ifndef TACO_GENERATED_" << func->name << "\n";
    out << "#define TACO_GENERATED_" << func->name << "\n";
    out << printFuncName(func) << ";\n";
    out << "#endif\n";
    return;
  }
  resetUniqueNameCounters();
  varMap.clear();
  for (auto* k : kernels) delete k;
  kernels.clear();
  NameCollector names(varMap, func->inputs, func->outputs,
                      [this](const std::string& n) { return genUniqueName(n); });
  func->body.accept(&names);
  KernelFinder finder(kernels, varMap);
  func->body.accept(&finder);
  const int numYields = countYields(func);
  coroutine = numYields > 0;
  localVars = names.localVars;
  labelCount = 0;
  // Coroutines are taco's host-side iterators over a tensor's storage
  // (reference codegen.cpp:232-270, 560-640): no kernels, resumable state
  // kept in a context struct.
  if (coroutine && !kernels.empty()) unsupported("GPU loops in a coroutine");

  for (size_t i = 0; i < kernels.size(); ++i) printKernel(func, i);

  out << printFuncName(func) << " {\n";
  indent = 1;
  out << printDecls(names.decls, func->inputs, func->outputs) << "\n";
  if (coroutine)
    out << printContextDeclAndInit(varMap, localVars, numYields, funcName) << "\n";
  inKernel = false;
  pendingLaunch = false;
  print(func->body);
  if (checkForAlloc(func)) {
    hostBarrier();
    out << "\n" << printPack(names.outputProps, func->outputs);
  }
  if (coroutine) {
    out << printCoroutineFinish(numYields, funcName);
    out << "  return 0;\n";
  } else {
    // Results are complete when the call returns (taco's contract), in the
    // library's blocking mode; otherwise the stream orders later work.
    out << "  return taco_hip_module_finish(0);\n";
  }
  out << "}\n";
  indent = 0;
  coroutine = false;
}

void CodeGen_HIP::printKernel(const Function* func, size_t index) {
  (void)func;
  Kernel* k = kernels[index];
  out << "__global__ __launch_bounds__(" << k->launchBound() << ")\nvoid " << funcName
      << "_kernel" << index << "(";
  for (size_t i = 0; i < k->params.size(); ++i) {
    std::string tp = typeOf(k->params[i]);
    if (i) out << ", ";
    if (k->byRef.count(k->paramNames[i])) {
      out << tp << "* " << k->paramNames[i] << "_ref";
      continue;
    }
    out << tp;
    if (tp.back() == '*') out << " __restrict__";
    out << " " << k->paramNames[i];
  }
  if (k->dynamicThreads) out << (k->params.empty() ? "" : ", ") << "int32_t taco_threads";
  if (k->dynamicGroups)
    out << ((k->params.empty() && !k->dynamicThreads) ? "" : ", ") << "int32_t taco_groups";
  out << ") {\n";
  indent = 1;
  // Ids: block = blockIdx.x * inc + start; a thread's lane in its group of
  // `threads` lanes and its group index give the thread and warp loop values.
  const For* b = k->block;
  doIndent();
  out << printCUDAType(b->var.type(), false) << " " << nameOf(b->var) << " = (int32_t)blockIdx.x * (";
  ir::simplify(b->increment).accept(this);
  out << ") + (";
  ir::simplify(b->start).accept(this);
  out << ");\n";
  const For* t = k->thread;
  doIndent();
  out << printCUDAType(t->var.type(), false) << " " << nameOf(t->var) << " = (int32_t)(threadIdx.x % "
      << k->threadsText() << ") * (";
  ir::simplify(t->increment).accept(this);
  out << ") + (";
  ir::simplify(t->start).accept(this);
  out << ");\n";
  if (k->warp) {
    const For* w = k->warp;
    doIndent();
    out << printCUDAType(w->var.type(), false) << " " << nameOf(w->var) << " = (int32_t)(threadIdx.x / "
        << k->threadsText() << ") * (";
    ir::simplify(w->increment).accept(this);
    out << ") + (";
    ir::simplify(w->start).accept(this);
    out << ");\n";
  }
  doIndent();
  out << "if (threadIdx.x >= " << k->threadsText() << " * " << k->groupsText() << ") return;\n";
  inKernel = true;
  cur = k;
  inThreadScope = inGroupScope = false;
  deviceLoopDepth = 0;
  b->contents.accept(this);
  inKernel = false;
  cur = nullptr;
  out << "}\n\n";
  indent = 0;
}

void CodeGen_HIP::printLaunch(size_t index) {
  Kernel* k = kernels[index];
  const For* b = k->block;
  doIndent();
  out << "{\n";
  indent++;
  doIndent();
  out << "const int64_t taco_grid = ((int64_t)(";
  ir::simplify(Sub::make(b->end, b->start)).accept(this);
  out << ") + (";
  ir::simplify(b->increment).accept(this);
  out << ") - 1) / (";
  ir::simplify(b->increment).accept(this);
  out << ");\n";
  if (k->blockSize == 0) {
    doIndent();
    out << "const int32_t taco_threads = (int32_t)(";
    parentPrecedence = TOP;
    if (k->dynamicThreads) k->threadsExpr.accept(this); else out << k->threads;
    out << ");\n";
    doIndent();
    out << "const int32_t taco_groups = (int32_t)(";
    parentPrecedence = TOP;
    if (k->dynamicGroups) k->groupsExpr.accept(this); else out << k->groups;
    out << ");\n";
    doIndent();
    out << "const int64_t taco_block = (int64_t)taco_threads * taco_groups;\n";
    doIndent();
    out << "if (taco_threads <= 0 || taco_block > 1024) return taco_hip_module_report(5, "
           "\"workgroup size out of range\");\n";
  }
  // Device copies of the host scalars the kernel assigns.
  for (size_t i = 0; i < k->params.size(); ++i) {
    const std::string& n = k->paramNames[i];
    if (!k->byRef.count(n)) continue;
    const std::string tp = typeOf(k->params[i]);
    doIndent();
    out << tp << "* " << n << "_ref = (" << tp << "*)taco_hip_unified_alloc(sizeof(" << tp << "));\n";
    doIndent();
    out << "if (!" << n << "_ref) return taco_hip_module_report(2, \"allocation of " << n
        << " failed\");\n";
    doIndent();
    out << "TACO_HIP_CHECK(taco_hip_copy_in(" << n << "_ref, &" << n << ", sizeof(" << tp << ")));\n";
  }
  doIndent();
  out << "if (taco_grid > 0) {\n";
  indent++;
  doIndent();
  out << "hipLaunchKernelGGL(" << funcName << "_kernel" << index
      << ", dim3((unsigned)taco_grid), dim3(";
  if (k->blockSize > 0) out << k->blockSize; else out << "(unsigned)taco_block";
  out << "), 0, taco_hip_stream()";
  for (auto& n : k->paramNames) out << ", " << n << (k->byRef.count(n) ? "_ref" : "");
  if (k->dynamicThreads) out << ", taco_threads";
  if (k->dynamicGroups) out << ", taco_groups";
  out << ");\n";
  doIndent();
  out << "TACO_HIP_CHECK(hipGetLastError());\n";
  indent--;
  doIndent();
  out << "}\n";
  for (size_t i = 0; i < k->params.size(); ++i) {
    const std::string& n = k->paramNames[i];
    if (!k->byRef.count(n)) continue;
    doIndent();
    out << "TACO_HIP_CHECK(taco_hip_copy_out(&" << n << ", " << n << "_ref, sizeof(" << typeOf(k->params[i])
        << ")));\n";
    doIndent();
    out << "taco_hip_unified_free(" << n << "_ref);\n";
  }
  indent--;
  doIndent();
  out << "}\n";
  pendingLaunch = true;
}

// Host code after a launch waits for the stream before it may read or
// write memory a kernel uses (taco's GPU printer synchronises the device
// after every launch instead).
void CodeGen_HIP::hostBarrier() {
  if (inKernel || !pendingLaunch) return;
  doIndent();
  out << "TACO_HIP_CHECK(taco_hip_stream_wait());\n";
  pendingLaunch = false;
}

void CodeGen_HIP::visit(const Var* op) {
  const std::string n = nameOf(op);
  if (inKernel && cur && cur->byRef.count(n))
    out << "(*" << n << "_ref)";
  else
    out << n;
}

void CodeGen_HIP::visit(const GetProperty* op) { out << nameOf(op); }

void CodeGen_HIP::visit(const For* op) {
  if (!inKernel) {
    for (size_t i = 0; i < kernels.size(); ++i) {
      if (kernels[i]->block == op) {
        printLaunch(i);
        return;
      }
    }
    hostBarrier();
  } else {
    if (cur && op == cur->thread) {
      bool was = inThreadScope;
      inThreadScope = true;
      op->contents.accept(this);
      inThreadScope = was;
      return;
    }
    if (cur && op == cur->warp) {
      bool was = inGroupScope;
      inGroupScope = true;
      op->contents.accept(this);
      inGroupScope = was;
      return;
    }
    if (op->parallel_unit == ParallelUnit::GPUWarpReduction ||
        op->parallel_unit == ParallelUnit::GPUBlockReduction) {
      // Every lane has written its partial result; the first lane of the
      // group (or of the workgroup) combines them serially, as taco's
      // reduction loops specify.
      const bool group = op->parallel_unit == ParallelUnit::GPUWarpReduction;
      doIndent();
      out << syncText(group);
      doIndent();
      if (group)
        out << "if (threadIdx.x % " << cur->threadsText() << " == 0) {\n";
      else
        out << "if (threadIdx.x == 0) {\n";
      indent++;
    }
  }
  if (op->unrollFactor > 0) {
    doIndent();
    out << "#pragma unroll " << op->unrollFactor << "\n";
  }
  doIndent();
  out << "for (";
  if (!coroutine) out << printCUDAType(op->var.type(), false) << " ";
  op->var.accept(this);
  out << " = ";
  parentPrecedence = TOP;
  op->start.accept(this);
  out << "; ";
  op->var.accept(this);
  out << " < ";
  parentPrecedence = BOTTOM;
  op->end.accept(this);
  out << "; ";
  op->var.accept(this);
  const Literal* lit = op->increment.as<Literal>();
  if (lit && (lit->type.isInt() || lit->type.isUInt()) && lit->equalsScalar(1)) {
    out << "++";
  } else {
    out << " += ";
    parentPrecedence = TOP;
    op->increment.accept(this);
  }
  out << ") {\n";
  if (inKernel) deviceLoopDepth++;
  op->contents.accept(this);
  if (inKernel) deviceLoopDepth--;
  doIndent();
  out << "}\n";
  if (inKernel && (op->parallel_unit == ParallelUnit::GPUWarpReduction ||
                   op->parallel_unit == ParallelUnit::GPUBlockReduction)) {
    indent--;
    doIndent();
    out << "}\n";
    const bool group = op->parallel_unit == ParallelUnit::GPUWarpReduction;
    // The combined result and the partials' storage are reused by the
    // next iteration: all lanes wait for the combining lane.
    doIndent();
    out << syncText(group);
  }
}

// Barrier around a serial reduction over a lane group (group) or over
// the workgroup: a group inside one wave needs only wave-level ordering.
std::string CodeGen_HIP::syncText(bool group) const {
  if (!group) return "__syncthreads();\n";
  if (cur->dynamicThreads) return "taco_hip_group_sync(taco_threads);\n";
  if (cur->threads <= 64 && 64 % cur->threads == 0) return "taco_hip_wave_sync();\n";
  return "__syncthreads();\n";
}

void CodeGen_HIP::visit(const While* op) {
  hostBarrier();
  if (inKernel) deviceLoopDepth++;
  IRPrinter::visit(op);
  if (inKernel) deviceLoopDepth--;
}

void CodeGen_HIP::visit(const IfThenElse* op) {
  hostBarrier();
  IRPrinter::visit(op);
}

void CodeGen_HIP::visit(const Min* op) {
  if (op->operands.size() == 1) {
    op->operands[0].accept(this);
    return;
  }
  for (size_t i = 0; i + 1 < op->operands.size(); ++i) {
    out << "TACO_HIP_MIN(";
    parentPrecedence = TOP;
    op->operands[i].accept(this);
    out << ", ";
  }
  parentPrecedence = TOP;
  op->operands.back().accept(this);
  for (size_t i = 0; i + 1 < op->operands.size(); ++i) out << ")";
}

void CodeGen_HIP::visit(const Max* op) {
  if (op->operands.size() == 1) {
    op->operands[0].accept(this);
    return;
  }
  for (size_t i = 0; i + 1 < op->operands.size(); ++i) {
    out << "TACO_HIP_MAX(";
    parentPrecedence = TOP;
    op->operands[i].accept(this);
    out << ", ";
  }
  parentPrecedence = TOP;
  op->operands.back().accept(this);
  for (size_t i = 0; i + 1 < op->operands.size(); ++i) out << ")";
}

void CodeGen_HIP::visit(const Allocate* op) {
  const std::string et = printCUDAType(op->var.type(), false);
  const std::string name = nameOf(op->var);
  if (inKernel) {
    if (op->is_realloc) unsupported("reallocation inside a kernel");
    const int64_t n = literalInt(op->num_elements, "a kernel array's size");
    doIndent();
    if (inThreadScope) {
      // per-thread array (registers or scratch)
      out << et << " " << name << "[" << n << "];\n";
      return;
    }
    // Workgroup-level storage in LDS: one slice per lane group when the
    // array belongs to a group (allocated inside the GPUWarp loop).
    if (inGroupScope && (cur->dynamicGroups || cur->dynamicThreads))
      unsupported("a group-level array in a kernel whose workgroup shape is set at launch");
    const int64_t slices = inGroupScope ? cur->groups : 1;
    out << "__shared__ " << et << " " << name << "_lds[" << n * slices << "];\n";
    doIndent();
    out << et << "* " << name << " = " << name << "_lds";
    if (slices > 1) out << " + (threadIdx.x / " << cur->threads << ") * " << n;
    out << ";\n";
    return;
  }
  hostBarrier();
  doIndent();
  if (op->is_realloc) {
    out << name << " = (" << et << "*)taco_hip_unified_realloc(" << name << ", sizeof(" << et
        << ") * (size_t)(";
    parentPrecedence = TOP;
    op->old_elements.accept(this);
    out << "), sizeof(" << et << ") * (size_t)(";
    op->num_elements.accept(this);
    out << "));\n";
  } else {
    out << name << " = (" << et << "*)taco_hip_unified_alloc(sizeof(" << et << ") * (size_t)(";
    parentPrecedence = TOP;
    op->num_elements.accept(this);
    out << "));\n";
  }
  doIndent();
  out << "if (!" << name << " && (";
  parentPrecedence = TOP;
  op->num_elements.accept(this);
  out << ") > 0) return taco_hip_module_report(2, \"allocation of " << name << " failed\");\n";
  if (op->clear) {
    doIndent();
    out << "TACO_HIP_CHECK(taco_hip_unified_clear(" << name << ", sizeof(" << et << ") * (size_t)(";
    parentPrecedence = TOP;
    op->num_elements.accept(this);
    out << ")));\n";
  }
}

void CodeGen_HIP::visit(const Free* op) {
  if (inKernel) return;  // LDS and per-thread arrays need no release
  hostBarrier();
  doIndent();
  out << "taco_hip_unified_free(";
  parentPrecedence = TOP;
  op->var.accept(this);
  out << ");\n";
}

void CodeGen_HIP::visit(const Sqrt* op) {
  const bool f32 = op->type.isFloat() && op->type.getNumBits() == 32;
  out << (f32 ? "sqrtf(" : "sqrt(");
  parentPrecedence = TOP;
  op->a.accept(this);
  out << ")";
}

void CodeGen_HIP::visit(const VarDecl* op) {
  hostBarrier();
  if (coroutine) {  // declared in the context
    doIndent();
    op->var.accept(this);
    parentPrecedence = TOP;
    out << " = ";
    op->rhs.accept(this);
    out << ";\n";
    return;
  }
  const Var* v = op->var.as<Var>();
  const bool isPtr = v && v->is_ptr;
  doIndent();
  out << printCUDAType(op->var.type(), isPtr);
  if (isPtr) out << " __restrict__";
  out << " ";
  op->var.accept(this);
  parentPrecedence = TOP;
  out << " = ";
  op->rhs.accept(this);
  out << ";\n";
}

void CodeGen_HIP::visit(const Literal* op) {
  if (op->type.isComplex()) unsupported("complex values");
  IRPrinter::visit(op);
}

void CodeGen_HIP::visit(const Yield* op) {
  if (inKernel || !coroutine) unsupported("a yield outside a host coroutine");
  printYield(op, localVars, varMap, labelCount, funcName);
  labelCount++;
}

void CodeGen_HIP::visit(const Call* op) {
  // taco's runtime helper names map onto the HIP runtime's.
  static const std::map<std::string, std::string> kNames = {
      {"taco_binarySearchAfter", "taco_hip_binarySearchAfter"},
      {"taco_binarySearchBefore", "taco_hip_binarySearchBefore"},
      {"taco_binarySearchBeforeBlockLaunch", "taco_hip_searchBlockStarts"},
      {"taco_binarySearchIndirectBeforeBlockLaunch", "taco_hip_searchBlockStartsIndirect"},
      {"cudaMemset", "taco_hip_memsetAsync"},
  };
  if (inKernel && op->func == "cudaMemset") unsupported("cudaMemset inside a kernel");
  if (op->func == "omp_get_max_threads") {  // host code here runs on one thread
    out << "1";
    return;
  }
  if (op->func == "omp_get_thread_num") {
    out << "0";
    return;
  }
  auto it = kNames.find(op->func);
  // The search helpers and the memset enqueue work on the stream: host
  // code after them waits like after a launch.
  if (!inKernel && it != kNames.end() && it->first != "taco_binarySearchAfter" &&
      it->first != "taco_binarySearchBefore")
    pendingLaunch = true;
  out << (it != kNames.end() ? it->second : op->func) << "(";
  for (size_t i = 0; i < op->args.size(); ++i) {
    if (i) out << ", ";
    parentPrecedence = CALL;
    op->args[i].accept(this);
  }
  out << ")";
}

void CodeGen_HIP::visit(const Store* op) {
  hostBarrier();
  if (!op->use_atomics || !inKernel) {
    IRPrinter::visit(op);
    return;
  }
  // The increment's left operand is the stored element itself (taco builds
  // compound stores as Store(a, i, Op(Load(a, i), b)), possibly with equal
  // but distinct index expressions).
  auto sameLoad = [&](const Expr& e) { return e.as<Load>() != nullptr; };
  doIndent();
  parentPrecedence = TOP;
  if (const Add* add = op->data.as<Add>()) {
    if (!sameLoad(add->a)) unsupported("atomic store that is not a compound add");
    // A store at the kernel's top level (every lane of a wave arrives
    // together) or one taco marks as a warp-level race may aggregate over
    // the wave when all lanes target the same element.
    const bool wave = deviceLoopDepth == 0 || op->atomic_parallel_unit == ParallelUnit::GPUWarp;
    out << (wave ? "taco_hip_atomicAddWave(" : "taco_hip_atomicAdd(");
    op->arr.accept(this);
    out << ", (int64_t)(";
    parentPrecedence = TOP;
    op->loc.accept(this);
    out << "), ";
    parentPrecedence = TOP;
    add->b.accept(this);
    out << ");\n";
  } else if (const Mul* mul = op->data.as<Mul>()) {
    if (!sameLoad(mul->a)) unsupported("atomic store that is not a compound multiply");
    out << "taco_hip_atomicMul(";
    op->arr.accept(this);
    out << ", (int64_t)(";
    parentPrecedence = TOP;
    op->loc.accept(this);
    out << "), ";
    parentPrecedence = TOP;
    mul->b.accept(this);
    out << ");\n";
  } else if (const BitOr* bor = op->data.as<BitOr>()) {
    if (!sameLoad(bor->a)) unsupported("atomic store that is not a compound or");
    out << "atomicOr(&";
    op->arr.accept(this);
    out << "[";
    parentPrecedence = TOP;
    op->loc.accept(this);
    out << "], ";
    parentPrecedence = TOP;
    bor->b.accept(this);
    out << ");\n";
  } else {
    unsupported("atomic store of an unsupported form");
  }
}

void CodeGen_HIP::visit(const Assign* op) {
  hostBarrier();
  if (op->use_atomics && inKernel) {
    const std::string n = nameOf(op->lhs);
    if (cur->byRef.count(n)) {
      // a host scalar every thread adds into: an atomic on its device copy
      doIndent();
      parentPrecedence = TOP;
      if (const Add* add = op->rhs.as<Add>()) {
        out << "taco_hip_atomicAdd(" << n << "_ref, 0, ";
        add->b.accept(this);
      } else if (const Mul* mul = op->rhs.as<Mul>()) {
        out << "taco_hip_atomicMul(" << n << "_ref, 0, ";
        mul->b.accept(this);
      } else {
        unsupported("atomic assignment of an unsupported form");
      }
      out << ");\n";
      return;
    }
    // A scalar declared inside the kernel is private to each thread:
    // the update needs no atomic.
  }
  IRPrinter::visit(op);
}

void CodeGen_HIP::visit(const Continue*) {
  doIndent();
  // At the kernel's top level the enclosing loop is the thread's own
  // iteration: skipping it ends the thread.
  out << ((inKernel && deviceLoopDepth == 0) ? "return;\n" : "continue;\n");
}

void CodeGen_HIP::visit(const Break*) {
  doIndent();
  out << ((inKernel && deviceLoopDepth == 0) ? "return;\n" : "break;\n");
}

void CodeGen_HIP::visit(const Sort* op) {
  if (inKernel) unsupported("sorting inside a kernel");
  hostBarrier();
  doIndent();
  out << "qsort(";
  for (size_t i = 0; i < op->args.size(); ++i) {
    if (i) out << ", ";
    parentPrecedence = CALL;
    op->args[i].accept(this);
  }
  out << ", taco_hip_cmp_i32);\n";
}

void CodeGen_HIP::visit(const Print* op) {
  hostBarrier();
  IRPrinter::visit(op);
}

void CodeGen_HIP::generateShim(const Stmt& func, std::stringstream& ret) {
  const Function* f = func.as<Function>();
  ret << "extern \"C\" int _shim_" << f->name << "(void** args) {\n";
  ret << "  return " << f->name << "(";
  size_t i = 0;
  const Datatype rt = f->getReturnType().second;
  if (rt != Datatype()) {  // coroutine: context, coordinates, values, capacity
    ret << "(void**)(args[0]), (char*)(args[1]), (" << printCUDAType(rt, true)
        << ")(args[2]), (int32_t*)(args[3])";
    i = 4;
  }
  auto arg = [&](const Expr& e) {
    const Var* v = e.as<Var>();
    std::string tp = v->is_tensor ? "taco_tensor_t*" : printCUDAType(v->type, v->is_ptr);
    ret << (i ? ", " : "") << "(" << tp << ")(args[" << i << "])";
    ++i;
  };
  for (auto& o : f->outputs) arg(o);
  for (auto& in : f->inputs) arg(in);
  ret << ");\n}\n";
}

}  // namespace ir
}  // namespace taco
