// IR node builders and a text dump.
#include "ir.h"

#include <sstream>

namespace taco_hip {
namespace ir {

const char* hip_type(DType t) {
  switch (t) {
    case DType::Bool: return "bool";
    case DType::Int32: return "int32_t";
    case DType::Int64: return "int64_t";
    case DType::UInt64: return "uint64_t";
    case DType::Float32: return "float";
    case DType::Float64: return "double";
    case DType::BFloat16: return "uint16_t";
  }
  return "void";
}

int dtype_bits(DType t) {
  switch (t) {
    case DType::Bool: return 8;
    case DType::Int32: return 32;
    case DType::Int64: return 64;
    case DType::UInt64: return 64;
    case DType::Float32: return 32;
    case DType::Float64: return 64;
    case DType::BFloat16: return 16;
  }
  return 0;
}

namespace {
std::shared_ptr<ExprNode> mk(EK k, DType t) {
  auto e = std::make_shared<ExprNode>();
  e->kind = k;
  e->type = t;
  return e;
}
std::shared_ptr<StmtNode> ms(SK k) {
  auto s = std::make_shared<StmtNode>();
  s->kind = k;
  return s;
}
DType join(DType a, DType b) {
  auto rank = [](DType t) {
    switch (t) {
      case DType::Bool: return 0;
      case DType::Int32: return 1;
      case DType::Int64: return 2;
      case DType::UInt64: return 3;
      case DType::BFloat16: return 4;
      case DType::Float32: return 5;
      case DType::Float64: return 6;
    }
    return 0;
  };
  return rank(a) >= rank(b) ? a : b;
}
}  // namespace

Expr lit(int64_t v, DType t) {
  auto e = mk(EK::Literal, t);
  e->ival = v;
  e->fval = (double)v;
  return e;
}
Expr flit(double v, DType t) {
  auto e = mk(EK::Literal, t);
  e->fval = v;
  e->ival = (int64_t)v;
  return e;
}
Expr var(const std::string& name, DType t, bool is_ptr, bool is_tensor) {
  auto e = mk(EK::Var, t);
  e->name = name;
  e->is_ptr = is_ptr;
  e->is_tensor = is_tensor;
  return e;
}
Expr tensor_var(const std::string& name) {
  return var(name, DType::Int32, true, true);
}
Expr bin(EK k, Expr a, Expr b) {
  DType t = join(a->type, b->type);
  if (k == EK::Eq || k == EK::Neq || k == EK::Gt || k == EK::Lt ||
      k == EK::Gte || k == EK::Lte || k == EK::And || k == EK::Or)
    t = DType::Bool;
  auto e = mk(k, t);
  e->a = a;
  e->b = b;
  return e;
}
Expr lnot(Expr a) {
  auto e = mk(EK::Not, DType::Bool);
  e->a = a;
  return e;
}
Expr cast(DType t, Expr a) {
  auto e = mk(EK::Cast, t);
  e->a = a;
  return e;
}
Expr load(Expr arr, Expr idx) {
  auto e = mk(EK::Load, arr->type);
  e->a = arr;
  e->b = idx;
  return e;
}
Expr call(const std::string& fn, std::vector<Expr> args, DType t) {
  auto e = mk(EK::Call, t);
  e->name = fn;
  e->args = std::move(args);
  return e;
}
Expr ternary(Expr c, Expr a, Expr b) {
  auto e = mk(EK::Ternary, join(a->type, b->type));
  e->a = c;
  e->b = a;
  e->c = b;
  return e;
}
Expr get_property(Expr tensor, TensorProperty p, int mode, int index,
                  const std::string& name, DType t, bool is_ptr) {
  auto e = mk(EK::GetProperty, t);
  e->tensor = tensor;
  e->property = p;
  e->mode = mode;
  e->index = index;
  e->name = name;
  e->is_ptr = is_ptr;
  return e;
}

Stmt block(std::vector<Stmt> s) {
  auto b = ms(SK::Block);
  for (auto& x : s)
    if (x) b->stmts.push_back(x);
  return b;
}
Stmt scope(Stmt s) {
  auto b = ms(SK::Scope);
  b->body = s;
  return b;
}
Stmt decl(Expr v, Expr init) {
  auto s = ms(SK::VarDecl);
  s->var = v;
  s->rhs = init;
  return s;
}
Stmt assign(Expr v, Expr value, bool atomics, ParallelUnit unit) {
  auto s = ms(SK::Assign);
  s->var = v;
  s->rhs = value;
  s->use_atomics = atomics;
  s->atomic_unit = unit;
  return s;
}
Stmt compound_add(Expr v, Expr value, bool atomics, ParallelUnit unit) {
  return assign(v, add(v, value), atomics, unit);
}
Stmt store(Expr arr, Expr idx, Expr value, bool atomics, ParallelUnit unit) {
  auto s = ms(SK::Store);
  s->arr = arr;
  s->loc = idx;
  s->rhs = value;
  s->use_atomics = atomics;
  s->atomic_unit = unit;
  return s;
}
Stmt store_add(Expr arr, Expr idx, Expr value, bool atomics,
               ParallelUnit unit) {
  return store(arr, idx, add(load(arr, idx), value), atomics, unit);
}
Stmt if_then(Expr c, Stmt t, Stmt e) {
  auto s = ms(SK::IfThenElse);
  s->cond = c;
  s->then_s = t;
  s->else_s = e;
  return s;
}
Stmt for_loop(Expr v, Expr start, Expr end, Expr inc, Stmt body, LoopKind k,
              ParallelUnit u, OutputRaceStrategy r, int unroll) {
  auto s = ms(SK::For);
  s->var = v;
  s->start = start;
  s->end = end;
  s->inc = inc;
  s->body = body;
  s->loop_kind = k;
  s->unit = u;
  s->race = r;
  s->unroll = unroll;
  return s;
}
Stmt while_loop(Expr c, Stmt body) {
  auto s = ms(SK::While);
  s->cond = c;
  s->body = body;
  return s;
}
Stmt allocate(Expr v, Expr n, bool clear, bool on_device) {
  auto s = ms(SK::Allocate);
  s->var = v;
  s->num_elements = n;
  s->clear = clear;
  s->on_device = on_device;
  return s;
}
Stmt allocate_scratch(Expr v, Expr n, int slot, bool clear) {
  auto s = ms(SK::Allocate);
  s->var = v;
  s->num_elements = n;
  s->clear = clear;
  s->on_device = true;
  s->scratch_slot = slot;
  return s;
}
Stmt free_stmt(Expr v) {
  auto s = ms(SK::Free);
  s->var = v;
  return s;
}
Stmt comment(const std::string& text) {
  auto s = ms(SK::Comment);
  s->text = text;
  return s;
}
Stmt ret(Expr v) {
  auto s = ms(SK::Return);
  s->rhs = v;
  return s;
}
Stmt brk() { return ms(SK::Break); }
Stmt allocate_shared(Expr v, Expr n) {
  auto s = ms(SK::Allocate);
  s->var = v;
  s->num_elements = n;
  s->shared = true;
  return s;
}
Stmt sync_threads() { return ms(SK::SyncThreads); }
Stmt sort(Expr arr, Expr size) {
  auto s = ms(SK::Sort);
  s->arr = arr;
  s->loc = size;
  return s;
}
Stmt launch(const std::string& launcher, std::vector<Expr> args) {
  auto s = ms(SK::Launch);
  s->name = launcher;
  s->outputs = std::move(args);
  return s;
}
Stmt evaluate(Expr e) {
  auto s = ms(SK::Evaluate);
  s->rhs = e;
  return s;
}
Stmt function(const std::string& name, Stmt body, std::vector<Expr> outputs,
              std::vector<Expr> inputs) {
  auto s = ms(SK::Function);
  s->name = name;
  s->body = body;
  s->outputs = std::move(outputs);
  s->inputs = std::move(inputs);
  return s;
}

// ---- dump ----
namespace {
const char* unit_name(ParallelUnit u) {
  switch (u) {
    case ParallelUnit::NotParallel: return "NotParallel";
    case ParallelUnit::DefaultUnit: return "DefaultUnit";
    case ParallelUnit::GPUBlock: return "GPUBlock";
    case ParallelUnit::GPUWarp: return "GPUWarp";
    case ParallelUnit::GPUThread: return "GPUThread";
    case ParallelUnit::CPUThread: return "CPUThread";
    case ParallelUnit::CPUVector: return "CPUVector";
    case ParallelUnit::CPUThreadGroupReduction: return "CPUThreadGroupReduction";
    case ParallelUnit::GPUBlockReduction: return "GPUBlockReduction";
    case ParallelUnit::GPUWarpReduction: return "GPUWarpReduction";
  }
  return "?";
}
const char* op_str(EK k) {
  switch (k) {
    case EK::Add: return " + ";
    case EK::Sub: return " - ";
    case EK::Mul: return " * ";
    case EK::Div: return " / ";
    case EK::Rem: return " % ";
    case EK::Eq: return " == ";
    case EK::Neq: return " != ";
    case EK::Gt: return " > ";
    case EK::Lt: return " < ";
    case EK::Gte: return " >= ";
    case EK::Lte: return " <= ";
    case EK::And: return " && ";
    case EK::Or: return " || ";
    default: return " ? ";
  }
}
void dump(std::ostream& os, const Stmt& s, int ind) {
  std::string pad(ind * 2, ' ');
  if (!s) return;
  switch (s->kind) {
    case SK::Block:
      for (auto& x : s->stmts) dump(os, x, ind);
      break;
    case SK::Scope:
      os << pad << "{\n";
      dump(os, s->body, ind + 1);
      os << pad << "}\n";
      break;
    case SK::VarDecl:
      os << pad << hip_type(s->var->type) << (s->var->is_ptr ? "* " : " ")
         << s->var->name << " = " << to_string(s->rhs) << "\n";
      break;
    case SK::Assign:
      os << pad << s->var->name << " = " << to_string(s->rhs)
         << (s->use_atomics ? std::string("  [atomic ") + unit_name(s->atomic_unit) + "]" : "")
         << "\n";
      break;
    case SK::Store:
      os << pad << to_string(s->arr) << "[" << to_string(s->loc)
         << "] = " << to_string(s->rhs)
         << (s->use_atomics ? std::string("  [atomic ") + unit_name(s->atomic_unit) + "]" : "")
         << "\n";
      break;
    case SK::IfThenElse:
      os << pad << "if " << to_string(s->cond) << "\n";
      dump(os, s->then_s, ind + 1);
      if (s->else_s) {
        os << pad << "else\n";
        dump(os, s->else_s, ind + 1);
      }
      break;
    case SK::For:
      os << pad << "for " << s->var->name << " in [" << to_string(s->start)
         << ", " << to_string(s->end) << ") step " << to_string(s->inc) << " "
         << unit_name(s->unit) << (s->unroll ? " unroll" : "") << "\n";
      dump(os, s->body, ind + 1);
      break;
    case SK::While:
      os << pad << "while " << to_string(s->cond) << "\n";
      dump(os, s->body, ind + 1);
      break;
    case SK::Allocate:
      os << pad << "allocate " << s->var->name << "[" << to_string(s->num_elements)
         << "]" << (s->clear ? " clear" : "")
         << (s->shared ? " shared" : s->on_device ? " device" : " host") << "\n";
      break;
    case SK::Free:
      os << pad << "free " << s->var->name << "\n";
      break;
    case SK::Function:
      os << "function " << s->name << "\n";
      dump(os, s->body, ind + 1);
      break;
    case SK::Comment:
      os << pad << "// " << s->text << "\n";
      break;
    case SK::Return:
      os << pad << "return " << to_string(s->rhs) << "\n";
      break;
    case SK::Break:
      os << pad << "break\n";
      break;
    case SK::Sort:
      os << pad << "sort " << to_string(s->arr) << "[0.." << to_string(s->loc) << ")\n";
      break;
    case SK::Launch:
      os << pad << "launch " << s->name << "\n";
      break;
    case SK::SyncThreads:
      os << pad << "syncthreads\n";
      break;
    case SK::Evaluate:
      os << pad << to_string(s->rhs) << "\n";
      break;
  }
}
}  // namespace

std::string to_string(const Expr& e) {
  if (!e) return "<null>";
  std::ostringstream os;
  switch (e->kind) {
    case EK::Literal:
      if (e->type == DType::Float64 || e->type == DType::Float32)
        os << e->fval;
      else
        os << e->ival;
      break;
    case EK::Var: os << e->name; break;
    case EK::Neg: os << "-(" << to_string(e->a) << ")"; break;
    case EK::Not: os << "!(" << to_string(e->a) << ")"; break;
    case EK::Min: os << "min(" << to_string(e->a) << ", " << to_string(e->b) << ")"; break;
    case EK::Max: os << "max(" << to_string(e->a) << ", " << to_string(e->b) << ")"; break;
    case EK::Cast: os << "(" << hip_type(e->type) << ")(" << to_string(e->a) << ")"; break;
    case EK::Load: os << to_string(e->a) << "[" << to_string(e->b) << "]"; break;
    case EK::Call: {
      os << e->name << "(";
      for (size_t i = 0; i < e->args.size(); ++i)
        os << (i ? ", " : "") << to_string(e->args[i]);
      os << ")";
      break;
    }
    case EK::GetProperty: os << e->name; break;
    case EK::Ternary:
      os << "(" << to_string(e->a) << " ? " << to_string(e->b) << " : "
         << to_string(e->c) << ")";
      break;
    default:
      os << "(" << to_string(e->a) << op_str(e->kind) << to_string(e->b) << ")";
  }
  return os.str();
}

std::string to_string(const Stmt& s) {
  std::ostringstream os;
  dump(os, s, 0);
  return os.str();
}

}  // namespace ir
}  // namespace taco_hip
