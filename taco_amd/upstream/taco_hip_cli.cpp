// taco-hip: taco's command-line front end (reference tools/taco.cpp) over
// taco with this package's GPU backend.  Expressions, formats, dimensions,
// data types, inputs, generated fills and scheduling commands are taken in
// taco's own option syntax and handed to taco's parser, scheduling language
// and lowering; -hip (or taco's spelling, -cuda) selects CodeGen_HIP and the
// HIP runtime instead of taco's C code generator.  -time=<repeat> times
// compilation, assembly and <repeat> computations; with -hip the computations
// are timed by HIP events on the library's stream, and the host clock is
// reported beside them.
//
// Options (taco's names and meanings): <expr>, -f=<tensor>:<format>[:<ordering>],
// -t=<tensor>:<float|double>, -d=<var|tensor>:<size>[,<size>...],
// -i=<tensor>:<file>, -o=<tensor>:<file>, -g=<tensor>:<fill>, -s=<commands>,
// -c, -time[=<repeat>], -print-compute, -print-assembly, -print-evaluate,
// -print-kernels, -print-concrete, -print-nocolor, -prefix=<name>,
// -nthreads=<n>, -hip / -cuda, -help.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "codegen/codegen.h"
#include "taco.h"
#include "taco/cuda.h"
#include "taco/index_notation/index_notation_nodes.h"
#include "taco/index_notation/index_notation_visitor.h"
#include "taco/index_notation/provenance_graph.h"
#include "taco/index_notation/transformations.h"
#include "taco/ir/ir.h"
#include "taco/lower/lower.h"
#include "taco/parser/parser.h"
#include "taco/parser/schedule_parser.h"
#include "taco/util/fill.h"
#include "taco/util/strings.h"
#include "taco_hip.h"

using namespace taco;

namespace {

int fail_with(const std::string& msg, int code) {
  std::cerr << "taco-hip: " << msg << std::endl;
  return code;
}

void usage() {
  std::cout <<
      "Usage: taco-hip <index expression> [options]\n"
      "  -f=<tensor>:<format>[:<ordering>]  per-mode formats: d dense, s sparse,\n"
      "                                     u sparse not unique, q singleton\n"
      "  -t=<tensor>:<type>                 float or double (default)\n"
      "  -d=<var|tensor>:<size>[,...]       dimension of an index variable or tensor\n"
      "  -i=<tensor>:<file>                 read a tensor (.mtx, .ttx, .tns, .rb)\n"
      "  -o=<tensor>:<file>                 write a tensor after computing\n"
      "  -g=<tensor>:<fill>                 generate data: vectors d r s h,\n"
      "                                     matrices d s h l f b\n"
      "  -s=\"<command>(<params>),...\"       scheduling commands: split, divide,\n"
      "                                     pos, fuse, reorder, bound, unroll,\n"
      "                                     parallelize, precompute\n"
      "  -c                                 compute kernel that also assembles\n"
      "  -time[=<repeat>]                   time compile, assemble, compute\n"
      "  -print-compute (default), -print-assembly, -print-evaluate,\n"
      "  -print-kernels, -print-concrete, -print-nocolor\n"
      "  -prefix=<name>                     prefix of the generated functions\n"
      "  -nthreads=<n>                      CPU threads (taco's C code)\n"
      "  -hip (also -cuda)                  CodeGen_HIP and the HIP runtime\n";
}

std::vector<std::string> split_once(const std::string& s, char sep) {
  const size_t k = s.find(sep);
  if (k == std::string::npos) return {s};
  return {s.substr(0, k), s.substr(k + 1)};
}

std::vector<std::string> split_all(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == sep) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  out.push_back(cur);
  return out;
}

bool parse_size(const std::string& s, size_t* out) {
  char* end = nullptr;
  const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
  if (s.empty() || *end) return false;
  *out = (size_t)v;
  return true;
}

struct Timing {
  double mean = 0, stdev = 0, median = 0;
  int n = 0;
};

Timing summarize(std::vector<double> ts) {
  Timing t;
  t.n = (int)ts.size();
  if (ts.empty()) return t;
  std::sort(ts.begin(), ts.end());
  // taco's TimeResults: mean and stdev over the middle 80% when there are
  // ten or more samples (include/taco/util/timers.h:52-77)
  size_t lo = 0, hi = ts.size();
  if (ts.size() >= 10) {
    lo = ts.size() / 10;
    hi = ts.size() - lo;
  }
  double s = 0;
  for (size_t i = lo; i < hi; ++i) s += ts[i];
  t.mean = s / (double)(hi - lo);
  double v = 0;
  for (size_t i = lo; i < hi; ++i) v += (ts[i] - t.mean) * (ts[i] - t.mean);
  t.stdev = hi - lo > 1 ? std::sqrt(v / (double)(hi - lo - 1)) : 0.0;
  t.median = ts.size() % 2 ? ts[ts.size() / 2]
                           : 0.5 * (ts[ts.size() / 2 - 1] + ts[ts.size() / 2]);
  return t;
}

void print_timing(const char* what, const Timing& t) {
  if (t.n == 1) {
    std::printf("%s %.4f ms\n", what, t.mean);
  } else {
    std::printf("%s\n  mean:   %.4f ms\n  stdev:  %.4f ms\n  median: %.4f ms\n", what, t.mean,
                t.stdev, t.median);
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Index variable of the statement named `name` (an original or a derived
// variable of an earlier command).
IndexVar find_var(const IndexStmt& stmt, const std::string& name) {
  ProvenanceGraph graph(stmt);
  for (const IndexVar& v : graph.getAllIndexVars())
    if (v.getName() == name) return v;
  taco_uerror << "index variable '" << name << "' is not in " << stmt;
  return IndexVar();
}

// The subexpression of stmt whose printed form, blanks removed, is `text`.
struct FindExpr : public IndexNotationVisitor {
  using IndexNotationVisitor::visit;
  std::string want;
  IndexExpr found;
  static std::string squeeze(std::string s) {
    s.erase(std::remove(s.begin(), s.end(), ' '), s.end());
    return s;
  }
  void check(const IndexExpr& e) {
    if (!found.defined() && squeeze(util::toString(e)) == want) found = e;
  }
  void visit(const AccessNode* n) override { check(IndexExpr(n)); }
  void visit(const UnaryExprNode* n) override {
    check(IndexExpr(n));
    if (!found.defined()) IndexNotationVisitor::visit(n);
  }
  void visit(const BinaryExprNode* n) override {
    check(IndexExpr(n));
    if (!found.defined()) IndexNotationVisitor::visit(n);
  }
};

ParallelUnit parse_unit(const std::string& s) {
  static const std::map<std::string, ParallelUnit> units = {
      {"NotParallel", ParallelUnit::NotParallel},
      {"DefaultUnit", ParallelUnit::DefaultUnit},
      {"GPUBlock", ParallelUnit::GPUBlock},
      {"GPUWarp", ParallelUnit::GPUWarp},
      {"GPUThread", ParallelUnit::GPUThread},
      {"CPUThread", ParallelUnit::CPUThread},
      {"CPUVector", ParallelUnit::CPUVector}};
  auto it = units.find(s);
  taco_uassert(it != units.end()) << "unknown parallel unit " << s;
  return it->second;
}

OutputRaceStrategy parse_race(const std::string& s) {
  static const std::map<std::string, OutputRaceStrategy> races = {
      {"IgnoreRaces", OutputRaceStrategy::IgnoreRaces},
      {"NoRaces", OutputRaceStrategy::NoRaces},
      {"Atomics", OutputRaceStrategy::Atomics},
      {"Temporary", OutputRaceStrategy::Temporary},
      {"ParallelReduction", OutputRaceStrategy::ParallelReduction}};
  auto it = races.find(s);
  taco_uassert(it != races.end()) << "unknown output race strategy " << s;
  return it->second;
}

BoundType parse_bound(const std::string& s) {
  static const std::map<std::string, BoundType> kinds = {
      {"MinExact", BoundType::MinExact},
      {"MinConstraint", BoundType::MinConstraint},
      {"MaxExact", BoundType::MaxExact},
      {"MaxConstraint", BoundType::MaxConstraint}};
  auto it = kinds.find(s);
  taco_uassert(it != kinds.end()) << "unknown bound type " << s;
  return it->second;
}

size_t size_arg(const std::string& s, const char* cmd) {
  size_t v = 0;
  taco_uassert(parse_size(s, &v)) << cmd << ": '" << s << "' is not a size";
  return v;
}

// taco's scheduling commands applied in order (the command set of taco's
// CLI, tools/taco.cpp:311-623).  Returns true when a GPU unit is used.
bool apply_schedule(const std::vector<std::vector<std::string>>& cmds, IndexStmt& stmt) {
  bool gpu = false;
  for (const std::vector<std::string>& c : cmds) {
    const std::string& op = c[0];
    const std::vector<std::string> a(c.begin() + 1, c.end());
    auto need = [&](size_t n, const char* form) {
      taco_uassert(a.size() == n) << "'" << op << "' takes " << n << " parameters: " << form;
    };
    if (op == "split" || op == "divide") {
      need(4, "(i, i1, i2, factor)");
      IndexVar outer(a[1]), inner(a[2]);
      const size_t f = size_arg(a[3], op.c_str());
      stmt = op == "split" ? stmt.split(find_var(stmt, a[0]), outer, inner, f)
                           : stmt.divide(find_var(stmt, a[0]), outer, inner, f);
    } else if (op == "pos") {
      need(3, "(i, ipos, tensor)");
      bool done = false;
      for (const Access& acc : getArgumentAccesses(stmt)) {
        if (acc.getTensorVar().getName() == a[2]) {
          IndexVar p(a[1]);
          stmt = stmt.pos(find_var(stmt, a[0]), p, acc);
          done = true;
          break;
        }
      }
      taco_uassert(done) << "pos: no operand named " << a[2];
    } else if (op == "fuse") {
      need(3, "(i, j, f)");
      IndexVar f(a[2]);
      stmt = stmt.fuse(find_var(stmt, a[0]), find_var(stmt, a[1]), f);
    } else if (op == "reorder") {
      taco_uassert(a.size() >= 2) << "'reorder' takes the variables outermost first";
      std::vector<IndexVar> order;
      for (const std::string& v : a) order.push_back(find_var(stmt, v));
      stmt = stmt.reorder(order);
    } else if (op == "bound") {
      need(4, "(i, ibound, bound, type)");
      IndexVar b(a[1]);
      stmt = stmt.bound(find_var(stmt, a[0]), b, size_arg(a[2], "bound"), parse_bound(a[3]));
    } else if (op == "unroll") {
      need(2, "(i, factor)");
      stmt = stmt.unroll(find_var(stmt, a[0]), size_arg(a[1], "unroll"));
    } else if (op == "parallelize") {
      need(3, "(i, unit, strategy)");
      const ParallelUnit unit = parse_unit(a[1]);
      gpu |= unit == ParallelUnit::GPUBlock || unit == ParallelUnit::GPUWarp ||
             unit == ParallelUnit::GPUThread;
      stmt = stmt.parallelize(find_var(stmt, a[0]), unit, parse_race(a[2]));
    } else if (op == "precompute") {
      taco_uassert(a.size() == 3 || a.size() == 4)
          << "'precompute' takes (expr, i, iw [, workspace])";
      FindExpr fe;
      fe.want = FindExpr::squeeze(a[0]);
      stmt.accept(&fe);
      taco_uassert(fe.found.defined()) << "precompute: no subexpression " << a[0];
      std::vector<IndexVar> is, iws;
      for (const std::string& v : parser::varListParser(a[1])) is.push_back(find_var(stmt, v));
      for (const std::string& v : parser::varListParser(a[2])) {
        IndexVar w(v);
        ProvenanceGraph graph(stmt);
        for (const IndexVar& u : graph.getAllIndexVars())
          if (u.getName() == v) w = u;
        iws.push_back(w);
      }
      std::vector<Dimension> dims;
      const auto domains = stmt.getIndexVarDomains();
      for (const IndexVar& v : is) {
        auto it = domains.find(v);
        dims.push_back(it != domains.end() ? it->second : Dimension(v));
      }
      TensorVar ws(a.size() == 4 ? a[3] : "workspace", Type(Float64, dims),
                   Format(std::vector<ModeFormatPack>(dims.size(), Dense)));
      stmt = stmt.precompute(fe.found, is, iws, ws);
    } else {
      taco_uerror << "unknown scheduling command " << op;
    }
  }
  return gpu;
}

bool parse_format(const std::string& spec, Format* out) {
  std::vector<ModeFormatPack> packs;
  for (char ch : spec) {
    switch (ch) {
      case 'd': packs.push_back(ModeFormat::Dense); break;
      case 's': packs.push_back(ModeFormat::Sparse); break;
      case 'u': packs.push_back(ModeFormat::Sparse(ModeFormat::NOT_UNIQUE)); break;
      case 'q': packs.push_back(ModeFormat::Singleton); break;
      case 'c': packs.push_back(ModeFormat::Singleton(ModeFormat::NOT_UNIQUE)); break;
      case 'p': packs.push_back(ModeFormat::Singleton(ModeFormat::PADDED)); break;
      default: return false;
    }
  }
  *out = Format(packs);
  return true;
}

bool parse_fill(const std::string& s, util::FillMethod* out) {
  static const std::map<std::string, util::FillMethod> fills = {
      {"d", util::FillMethod::Dense},       {"u", util::FillMethod::Uniform},
      {"r", util::FillMethod::Random},      {"s", util::FillMethod::Sparse},
      {"h", util::FillMethod::HyperSparse}, {"v", util::FillMethod::SlicingV},
      {"l", util::FillMethod::SlicingH},    {"f", util::FillMethod::FEM},
      {"b", util::FillMethod::Blocked}};
  auto it = fills.find(s);
  if (it == fills.end()) return false;
  *out = it->second;
  return true;
}

// The result's stored components as a coordinate file, values with 17
// significant digits (taco's write uses the stream default of 6).
template <typename T>
void gather(const TensorBase& t, std::vector<int32_t>* coords, std::vector<double>* vals) {
  for (auto& v : iterate<T>(t)) {
    for (int k = 0; k < t.getOrder(); ++k) coords->push_back(v.first[k]);
    vals->push_back((double)v.second);
  }
}

int write_result(const std::string& path, const TensorBase& t) {
  std::vector<int32_t> coords;
  std::vector<double> vals;
  if (t.getComponentType() == Float64) gather<double>(t, &coords, &vals);
  else gather<float>(t, &coords, &vals);
  const std::vector<int>& d = t.getDimensions();
  std::vector<int32_t> dims(d.begin(), d.end());
  if (taco_hip_write_tensor_file(path.c_str(), t.getOrder(), dims.data(), (int64_t)vals.size(),
                                 coords.data(), vals.data(), 0) != 0)
    return fail_with(taco_hip_last_error(), 11);
  return 0;
}

int run(int argc, char** argv) {
  std::string expr, prefix;
  std::map<std::string, Format> formats;
  std::map<std::string, Datatype> types;
  std::map<std::string, std::vector<int>> dims;
  std::map<std::string, std::string> inputs, outputs;
  std::map<std::string, util::FillMethod> fills;
  std::vector<std::vector<std::string>> schedule;
  bool hip = false, with_assembly = false, timing = false;
  bool p_compute = false, p_assembly = false, p_evaluate = false, p_kernels = false,
       p_concrete = false;
  int repeat = 1, nthreads = 0;
  for (int i = 1; i < argc; ++i) {
    std::string arg = argv[i];
    if (arg.rfind("--", 0) == 0) arg = arg.substr(1);
    if (arg.empty() || arg[0] != '-') {
      if (!expr.empty()) return fail_with("more than one expression", 3);
      expr = arg;
      continue;
    }
    const std::vector<std::string> kv = split_once(arg, '=');
    const std::string& name = kv[0];
    const std::string value = kv.size() > 1 ? kv[1] : "";
    if (name == "-help") {
      usage();
      return 0;
    } else if (name == "-f") {
      const std::vector<std::string> d = split_all(value, ':');
      if (d.size() < 2 || d.size() > 3) return fail_with("incorrect format descriptor", 4);
      Format f;
      if (!parse_format(d[1], &f)) return fail_with("unknown mode format in " + d[1], 4);
      if (d.size() == 3) {
        std::vector<int> ordering;
        for (const std::string& o : split_all(d[2], ',')) ordering.push_back(std::atoi(o.c_str()));
        if (ordering.size() != d[1].size()) return fail_with("ordering does not match format", 4);
        f = Format(f.getModeFormatPacks(), ordering);
      }
      formats.insert({d[0], f});
    } else if (name == "-t") {
      const std::vector<std::string> d = split_all(value, ':');
      if (d.size() != 2) return fail_with("incorrect type descriptor", 4);
      if (d[1] == "double") types.insert({d[0], Float64});
      else if (d[1] == "float") types.insert({d[0], Float32});
      else return fail_with("data type " + d[1] + " is not on the HIP backend's path", 4);
    } else if (name == "-d") {
      const std::vector<std::string> d = split_all(value, ':');
      if (d.size() != 2) return fail_with("incorrect dimension descriptor", 4);
      std::vector<int> sizes;
      for (const std::string& s : split_all(d[1], ',')) {
        size_t v = 0;
        if (!parse_size(s, &v) || v == 0 || v > 0x7fffffff) return fail_with("bad dimension " + s, 4);
        sizes.push_back((int)v);
      }
      dims.insert({d[0], sizes});
    } else if (name == "-i" || name == "-o") {
      const std::vector<std::string> d = split_once(value, ':');
      if (d.size() != 2) return fail_with("incorrect file descriptor", 4);
      (name == "-i" ? inputs : outputs).insert({d[0], d[1]});
    } else if (name == "-g") {
      const std::vector<std::string> d = split_all(value, ':');
      util::FillMethod m;
      if (d.size() != 2 || !parse_fill(d[1], &m)) return fail_with("incorrect fill descriptor", 4);
      fills.insert({d[0], m});
    } else if (name == "-s") {
      for (const std::vector<std::string>& c : parser::ScheduleParser(value)) schedule.push_back(c);
    } else if (name == "-c") {
      with_assembly = true;
    } else if (name == "-time") {
      timing = true;
      if (!value.empty()) {
        size_t v = 0;
        if (!parse_size(value, &v) || v == 0) return fail_with("bad -time repeat", 4);
        repeat = (int)v;
      }
    } else if (name == "-print-compute") {
      p_compute = true;
    } else if (name == "-print-assembly") {
      p_assembly = true;
    } else if (name == "-print-evaluate") {
      p_evaluate = true;
    } else if (name == "-print-kernels") {
      p_kernels = true;
    } else if (name == "-print-concrete") {
      p_concrete = true;
    } else if (name == "-print-nocolor") {
    } else if (name == "-prefix") {
      prefix = value;
    } else if (name == "-nthreads") {
      nthreads = std::atoi(value.c_str());
    } else if (name == "-hip" || name == "-cuda") {
      hip = true;
    } else {
      return fail_with("unknown option " + arg, 2);
    }
  }
  if (expr.empty()) {
    usage();
    return 1;
  }
  if (!p_assembly && !p_evaluate && !p_kernels && !timing) p_compute = true;

  std::map<std::string, TensorBase> loaded;
  for (const auto& in : inputs) {
    Format f = formats.count(in.first) ? formats.at(in.first) : Format();
    TensorBase t = formats.count(in.first) ? read(in.second, f, false)
                                          : read(in.second, ModeFormat::Dense, false);
    t.setName(in.first);
    t.pack();
    loaded.insert({in.first, t});
    std::cout << in.first << " size: (" << util::join(t.getDimensions(), " x ") << ")"
              << std::endl;
  }
  parser::Parser parser(expr, formats, types, dims, loaded, 42);
  try {
    parser.parse();
  } catch (parser::ParseError& e) {
    return fail_with(e.getMessage(), 6);
  }
  TensorBase result = parser.getResultTensor();
  for (const auto& g : fills) {
    TensorBase t = parser.getTensor(g.first);
    util::fillTensor(t, g.second);
    loaded.insert({g.first, t});
  }
  bool benchmark = true;
  for (const auto& t : parser.getTensors())
    if (!(t.second == result) && !loaded.count(t.second.getName())) benchmark = false;

  if (nthreads > 0) taco_set_num_threads(nthreads);
  IndexStmt stmt = makeConcreteNotation(makeReductionNotation(result.getAssignment()));
  stmt = reorderLoopsTopologically(stmt);
  if (!schedule.empty()) {
    hip |= apply_schedule(schedule, stmt);
  } else {
    stmt = insertTemporaries(stmt);
    stmt = parallelizeOuterLoop(stmt);
  }
  set_CUDA_codegen_enabled(hip);
  stmt = scalarPromote(stmt);
  if (p_concrete) std::cout << stmt << std::endl;

  ir::Stmt compute = lower(stmt, prefix + "compute", with_assembly, true);
  ir::Stmt assemble = lower(stmt, prefix + "assemble", true, false);
  ir::Stmt evaluate = lower(stmt, prefix + "evaluate", true, true);
  auto print = [&](const ir::Stmt& f) {
    std::shared_ptr<ir::CodeGen> cg =
        ir::CodeGen::init_default(std::cout, ir::CodeGen::ImplementationGen);
    cg->compile(f, false);
  };
  if (p_kernels) {
    std::shared_ptr<ir::CodeGen> cg =
        ir::CodeGen::init_default(std::cout, ir::CodeGen::ImplementationGen);
    cg->compile(ir::Block::make({compute, assemble, evaluate}), true);
  } else {
    if (p_compute) print(compute);
    if (p_assembly) print(assemble);
    if (p_evaluate) print(evaluate);
  }
  if (!timing) return 0;
  if (!benchmark) return fail_with("-time needs every operand (-i or -g)", 9);

  const double c0 = now_ms();
  result.compile(stmt, with_assembly);
  const double compile_ms = now_ms() - c0;
  const double a0 = now_ms();
  result.assemble();
  const double assemble_ms = now_ms() - a0;
  std::vector<double> host, dev;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hip && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    return fail_with("hipEventCreate failed", 10);
  hipStream_t st = (hipStream_t)taco_hip_get_stream();
  for (int r = 0; r < repeat; ++r) {
    const double t0 = now_ms();
    if (hip) (void)hipEventRecord(e0, st);
    result.compute();
    if (hip) {
      (void)hipEventRecord(e1, st);
      float ms = 0.f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        return fail_with("HIP event timing failed", 10);
      dev.push_back(ms);
    }
    host.push_back(now_ms() - t0);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  std::printf("Compile:  %.4f ms\nAssemble: %.4f ms\n", compile_ms, assemble_ms);
  print_timing("Compute (host clock):", summarize(host));
  if (hip) print_timing("Compute (HIP events):", summarize(dev));
  for (const auto& o : outputs) {
    if (o.first != result.getName()) return fail_with("-o names the result tensor only", 4);
    if (int rc = write_result(o.second, result)) return rc;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (TacoException& e) {
    return fail_with(e.what(), 5);
  } catch (std::exception& e) {
    return fail_with(e.what(), 5);
  }
}
