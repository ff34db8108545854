// Imperative IR consumed by CodeGen_HIP.
//
// The node vocabulary and the parallel/race tags are those of taco's IR
// (reference include/taco/ir/ir.h:22-797, include/taco/ir_tags.h:11-26), so a
// lowered taco function maps onto these nodes one to one: For carries its
// LoopKind, ParallelUnit and unroll factor; Assign/Store carry the atomics
// flag and the parallel unit the atomic is relative to; Allocate carries
// clear/realloc.  The implementation is this backend's own: small
// shared_ptr nodes with an enum tag, built by the lowering builders in
// lower.cpp and printed by codegen_hip.cpp.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace taco_hip {
namespace ir {

enum class ParallelUnit {
  NotParallel, DefaultUnit, GPUBlock, GPUWarp, GPUThread, CPUThread, CPUVector,
  CPUThreadGroupReduction, GPUBlockReduction, GPUWarpReduction
};
enum class OutputRaceStrategy { IgnoreRaces, NoRaces, Atomics, Temporary,
                                ParallelReduction };
enum class LoopKind { Serial, Static, Dynamic, Runtime, Vectorized,
                      Static_Chunked };

enum class DType { Bool, Int32, Int64, UInt64, Float32, Float64, BFloat16 };
const char* hip_type(DType t);
int dtype_bits(DType t);

enum class TensorProperty { Order, Dimension, ComponentSize, ModeOrdering,
                            ModeTypes, Indices, Values, ValuesSize };

struct ExprNode;
struct StmtNode;
using Expr = std::shared_ptr<const ExprNode>;
using Stmt = std::shared_ptr<const StmtNode>;

enum class EK {
  Literal, Var, Neg, Add, Sub, Mul, Div, Rem, Min, Max, Eq, Neq, Gt, Lt, Gte,
  Lte, And, Or, Not, Cast, Call, Load, GetProperty, Ternary
};

struct ExprNode {
  EK kind;
  DType type = DType::Int32;
  // Literal
  int64_t ival = 0;
  double fval = 0;
  // Var
  std::string name;
  bool is_ptr = false;
  bool is_tensor = false;
  // operands (binary: a, b; unary/cast: a; load: a = array, b = index;
  // call: args; ternary: a ? b : c)
  Expr a, b, c;
  std::vector<Expr> args;
  // GetProperty
  Expr tensor;
  TensorProperty property = TensorProperty::Dimension;
  int mode = 0;
  int index = 0;
};

enum class SK {
  Block, Scope, VarDecl, Assign, Store, IfThenElse, For, While, Allocate, Free,
  Function, Comment, Return, Break, Sort, Launch, SyncThreads, Evaluate
};

struct StmtNode {
  SK kind;
  std::vector<Stmt> stmts;  // Block / Function body pieces
  Expr var;                 // VarDecl/Assign target, For variable, Allocate var
  Expr rhs;                 // VarDecl init, Assign value, Store data
  Expr arr, loc;            // Store array/location
  Expr cond;                // If / While
  Stmt then_s, else_s, body;
  // Assign/Store
  bool use_atomics = false;
  ParallelUnit atomic_unit = ParallelUnit::NotParallel;
  // For
  Expr start, end, inc;
  LoopKind loop_kind = LoopKind::Serial;
  ParallelUnit unit = ParallelUnit::NotParallel;
  OutputRaceStrategy race = OutputRaceStrategy::IgnoreRaces;
  int unroll = 0;
  // Allocate
  Expr num_elements;
  bool clear = false;
  bool is_realloc = false;
  bool on_device = true;
  int scratch_slot = -1;  // >= 0: library workspace slot, not freed
  bool shared = false;    // in a kernel: one LDS array per workgroup
  // Function
  std::string name;
  std::vector<Expr> outputs, inputs;
  // Comment
  std::string text;
  // Sort: sort arr[0..loc) ascending, wave-cooperative inside a GPUWarp loop
  // Launch: name = runtime helper kernel launcher, args in `outputs`
};

// ---- expression builders ----
Expr lit(int64_t v, DType t = DType::Int32);
Expr flit(double v, DType t = DType::Float64);
Expr var(const std::string& name, DType t, bool is_ptr = false,
         bool is_tensor = false);
Expr tensor_var(const std::string& name);
Expr bin(EK k, Expr a, Expr b);
inline Expr add(Expr a, Expr b) { return bin(EK::Add, a, b); }
inline Expr sub(Expr a, Expr b) { return bin(EK::Sub, a, b); }
inline Expr mul(Expr a, Expr b) { return bin(EK::Mul, a, b); }
inline Expr div(Expr a, Expr b) { return bin(EK::Div, a, b); }
inline Expr rem(Expr a, Expr b) { return bin(EK::Rem, a, b); }
inline Expr min(Expr a, Expr b) { return bin(EK::Min, a, b); }
inline Expr max(Expr a, Expr b) { return bin(EK::Max, a, b); }
inline Expr lt(Expr a, Expr b) { return bin(EK::Lt, a, b); }
inline Expr lte(Expr a, Expr b) { return bin(EK::Lte, a, b); }
inline Expr gt(Expr a, Expr b) { return bin(EK::Gt, a, b); }
inline Expr gte(Expr a, Expr b) { return bin(EK::Gte, a, b); }
inline Expr eq(Expr a, Expr b) { return bin(EK::Eq, a, b); }
inline Expr neq(Expr a, Expr b) { return bin(EK::Neq, a, b); }
inline Expr land(Expr a, Expr b) { return bin(EK::And, a, b); }
Expr lnot(Expr a);
Expr cast(DType t, Expr a);
Expr load(Expr arr, Expr idx);
Expr call(const std::string& fn, std::vector<Expr> args, DType t);
Expr ternary(Expr c, Expr a, Expr b);
Expr get_property(Expr tensor, TensorProperty p, int mode = 0, int index = 0,
                  const std::string& name = "", DType t = DType::Int32,
                  bool is_ptr = false);

// ---- statement builders ----
Stmt block(std::vector<Stmt> s);
Stmt scope(Stmt s);
Stmt decl(Expr v, Expr init);
Stmt assign(Expr v, Expr value, bool atomics = false,
            ParallelUnit unit = ParallelUnit::NotParallel);
Stmt compound_add(Expr v, Expr value, bool atomics = false,
                  ParallelUnit unit = ParallelUnit::NotParallel);
Stmt store(Expr arr, Expr idx, Expr value, bool atomics = false,
           ParallelUnit unit = ParallelUnit::NotParallel);
Stmt store_add(Expr arr, Expr idx, Expr value, bool atomics = false,
               ParallelUnit unit = ParallelUnit::NotParallel);
Stmt if_then(Expr c, Stmt t, Stmt e = nullptr);
Stmt for_loop(Expr v, Expr start, Expr end, Expr inc, Stmt body,
              LoopKind k = LoopKind::Serial,
              ParallelUnit u = ParallelUnit::NotParallel,
              OutputRaceStrategy r = OutputRaceStrategy::IgnoreRaces,
              int unroll = 0);
Stmt while_loop(Expr c, Stmt body);
Stmt allocate(Expr v, Expr n, bool clear = false, bool on_device = true);
// Device scratch from the library's reusable workspace (slot 4..7).
Stmt allocate_scratch(Expr v, Expr n, int slot, bool clear = false);
// A workgroup-shared (LDS) array of literal size inside a GPUBlock loop
// (the reference's __shared__ temporaries, codegen_cuda.cpp:1023-1070).
Stmt allocate_shared(Expr v, Expr n);
// Workgroup barrier.
Stmt sync_threads();
Stmt free_stmt(Expr v);
Stmt comment(const std::string& text);
Stmt ret(Expr v);
Stmt brk();
Stmt sort(Expr arr, Expr size);
Stmt launch(const std::string& launcher, std::vector<Expr> args);
Stmt evaluate(Expr e);
Stmt function(const std::string& name, Stmt body, std::vector<Expr> outputs,
              std::vector<Expr> inputs);

// Human-readable dump (tests and debugging).
std::string to_string(const Stmt& s);
std::string to_string(const Expr& e);

}  // namespace ir
}  // namespace taco_hip
