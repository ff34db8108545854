// The reference's GPU test statements run through taco itself — its
// Tensor/IndexVar API, scheduling language, lowering and TensorBase
// compile/assemble/compute — with CodeGen_HIP, the HIP Module and the HIP
// runtime plugged in (libtaco_upstream.so).  Inputs are regenerated exactly
// as the reference's tests draw them (srand/rand with the same seeds, sizes
// and float arithmetic); expected results come from dense brute force over
// the same inserted values, which are small integers, so every correct
// summation order gives them exactly.
//
// Statements (reference file:line):
//   spmv_rows_gpu        scheduleSpMVRowsGPU        tests-scheduling-eval.cpp:211-223, inputs :1210-1256
//   spmv_splitpos_gpu    scheduleSpMVSplitPosGPU    :233-247
//   spmv_thread_row_gpu  scheduleSpMVThreadPerRowGPU :225-231
//   spmv_default_gpu     parallelizeOuterLoop       transformations.cpp:1562-1606
//   spmv_precompute_gpu  scheduleSpMVGPU            :193-209 (see below)
//   spmm_gpu             scheduleSpMMGPU            :249-268, inputs :1258-1307
//   spmm_rows_gpu        scheduleSpMMRowsGPU        :345-355
//   spmm_dcsr_gpu        scheduleSpMMNZRowsGPU      :357-369, inputs :1309-1356
//   sddmm_gpu            scheduleSDDMMGPU           :270-287, inputs :1358-1412
//   ttm_gpu              scheduleTTMGPU             :289-306, inputs :1414-1466
//   ttv_gpu              scheduleTTVGPU             :308-325, inputs :1468-1516 (see below)
//   mttkrp_gpu           scheduleMTTKRPGPU          :327-342, inputs :1518-1580
//   atomic_reduction_gpu tests-scheduling.cpp:444-490
//   temporary_reduction_gpu tests-scheduling.cpp:492-536
//   tiling_gpu           tests-scheduling.cpp:393-440
//   windowing_gpu_XY     tests-windowing.cpp:342-386 (X, Y in {d, s})
//
// scheduleSpMVGPU and scheduleTTVGPU precompute the product into a
// per-thread array; taco's precompute rewrite then makes the consumer a
// plain assignment (transformations.cpp:418-445), so taco's own GPU code
// stores `y[i] = precomputed[k]` for every entry of a row in turn (the
// reference's CUDA printer emits the same store).  The result of such a row
// is whichever entry's product landed last.  Those two statements are
// checked for exactly that: each output holds one of its row's products.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "codegen_hip.h"
#include "taco.h"
#include "taco/cuda.h"
#include "taco/index_notation/transformations.h"
#include "taco/lower/lower.h"
#include "taco_hip.h"

using namespace taco;

namespace {

const int WARP_SIZE = 32;
IndexVar i("i"), j("j"), k("k"), l("l");

// ---- the reference's GPU schedules (tests-scheduling-eval.cpp:193-369,
// restated as test inputs) ----

IndexStmt scheduleSpMVGPU(IndexStmt stmt, Tensor<double> A, IndexExpr pre,
                          int NNZ_PER_THREAD = 8, int BLOCK_SIZE = 256) {
  int NNZ_PER_WARP = NNZ_PER_THREAD * WARP_SIZE;
  int NNZ_PER_TB = NNZ_PER_THREAD * BLOCK_SIZE;
  IndexVar f("f"), fpos("fpos"), fpos1("fpos1"), fpos2("fpos2"), block("block"),
      warp("warp"), thread("thread"), thread_nz("thread_nz"), thread_nz_pre("thread_nz_pre");
  TensorVar precomputed("precomputed", Type(Float64, {Dimension(thread_nz)}), taco::dense);
  return stmt.fuse(i, j, f)
      .pos(f, fpos, A(i, j))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, fpos2, NNZ_PER_WARP)
      .split(fpos2, thread, thread_nz, NNZ_PER_THREAD)
      .reorder({block, warp, thread, thread_nz})
      .precompute(pre, thread_nz, thread_nz_pre, precomputed)
      .unroll(thread_nz_pre, NNZ_PER_THREAD)
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleSpMVRowsGPU(IndexStmt stmt, Tensor<double> A, int ROWS_PER_WARP = 1,
                              int BLOCK_SIZE = 256) {
  int ROWS_PER_TB = ROWS_PER_WARP * BLOCK_SIZE;
  IndexVar block("block"), warp("warp"), thread("thread"), thread_nz("thread_nz"),
      jpos("jpos"), block_row("block_row"), warp_row("warp_row");
  return stmt.split(i, block, block_row, ROWS_PER_TB)
      .split(block_row, warp_row, warp, BLOCK_SIZE / WARP_SIZE)
      .pos(j, jpos, A(i, j))
      .split(jpos, thread_nz, thread, WARP_SIZE)
      .reorder({block, warp, warp_row, thread, thread_nz})
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Temporary);
}

IndexStmt scheduleSpMVThreadPerRowGPU(IndexStmt stmt, int BLOCK_SIZE = 256) {
  IndexVar block("block"), thread("thread");
  return stmt.split(i, block, thread, BLOCK_SIZE)
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::NoRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::NoRaces);
}

IndexStmt scheduleSpMVSplitPosGPU(IndexStmt stmt, Tensor<double> A, int NNZ_PER_THREAD = 8,
                                  int BLOCK_SIZE = 256) {
  int NNZ_PER_WARP = NNZ_PER_THREAD * WARP_SIZE;
  int NNZ_PER_TB = NNZ_PER_THREAD * BLOCK_SIZE;
  IndexVar f("f"), fpos("fpos"), fpos1("fpos1"), fpos2("fpos2"), block("block"),
      warp("warp"), thread("thread"), thread_nz("thread_nz");
  return stmt.fuse(i, j, f)
      .pos(f, fpos, A(i, j))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, fpos2, NNZ_PER_WARP)
      .split(fpos2, thread, thread_nz, NNZ_PER_THREAD)
      .reorder({block, warp, thread, thread_nz})
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleSpMMGPU(IndexStmt stmt, Tensor<double> A, int NNZ_PER_WARP = 8,
                          int BLOCK_SIZE = 256) {
  int NNZ_PER_TB = NNZ_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar f("f"), fpos("fpos"), block("block"), fpos1("fpos1"), warp("warp"), nnz("nnz");
  IndexVar dense_val_unbounded("dense_val_unbounded"), dense_val("dense_val"), thread("thread");
  return stmt.reorder({i, j, k})
      .fuse(i, j, f)
      .pos(f, fpos, A(i, j))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, nnz, NNZ_PER_WARP)
      .split(k, dense_val_unbounded, thread, WARP_SIZE)
      .reorder({block, warp, thread, dense_val_unbounded, nnz})
      .bound(dense_val_unbounded, dense_val, 4, BoundType::MaxExact)
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleSpMMRowsGPU(IndexStmt stmt, int ROWS_PER_WARP = 4, int BLOCK_SIZE = 256) {
  int ROWS_PER_TB = ROWS_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar i1("i1"), block("block"), warp("warp"), warp_row("warp_row"), thread("thread"),
      thread_col("thread_col");
  return stmt.split(i, block, i1, ROWS_PER_TB)
      .split(i1, warp, warp_row, ROWS_PER_WARP)
      .split(k, thread, thread_col, 32)
      .reorder({block, warp, warp_row, thread, thread_col, j})
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleSpMMNZRowsGPU(IndexStmt stmt, Tensor<double> A, int NZ_ROWS_PER_WARP = 4,
                                int BLOCK_SIZE = 256) {
  int NZ_ROWS_PER_TB = NZ_ROWS_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar ip("ip"), ip1("ip1"), block("block"), warp("warp"), warp_row("warp_row"),
      thread("thread"), thread_col("thread_col");
  return stmt.pos(i, ip, A(i, j))
      .split(ip, block, ip1, NZ_ROWS_PER_TB)
      .split(ip1, warp, warp_row, NZ_ROWS_PER_WARP)
      .split(k, thread, thread_col, 32)
      .reorder({block, warp, warp_row, thread, thread_col, j})
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleSDDMMGPU(IndexStmt stmt, Tensor<double> B, int NNZ_PER_WARP = 8 * 32,
                           int BLOCK_SIZE = 256, int CO_FACTOR = 4) {
  int NNZ_PER_TB = NNZ_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar f("f"), fpos("fpos"), block("block"), fpos1("fpos1"), warp("warp"), nnz("nnz");
  IndexVar dense_val_unbounded("dense_val_unbounded"), dense_val("dense_val"), thread("thread");
  return stmt.reorder({i, k, j})
      .fuse(i, k, f)
      .pos(f, fpos, B(i, k))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, nnz, NNZ_PER_WARP)
      .split(j, dense_val_unbounded, thread, WARP_SIZE)
      .bound(dense_val_unbounded, dense_val, CO_FACTOR, BoundType::MaxExact)
      .reorder({block, warp, nnz, thread, dense_val})
      .unroll(dense_val, CO_FACTOR)
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::Atomics)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::ParallelReduction);
}

IndexStmt scheduleTTMGPU(IndexStmt stmt, Tensor<double> B, int NNZ_PER_WARP = 8 * 32,
                         int BLOCK_SIZE = 256, int CO_FACTOR = 4) {
  int NNZ_PER_TB = NNZ_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar jk("jk"), f("f"), fpos("fpos"), block("block"), fpos1("fpos1"), warp("warp"),
      nnz("nnz"), dense_val_unbounded("dense_val_unbounded"), dense_val("dense_val"),
      thread("thread");
  return stmt.reorder({i, j, k, l})
      .fuse(j, k, jk)
      .fuse(i, jk, f)
      .pos(f, fpos, B(i, j, k))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, nnz, NNZ_PER_WARP)
      .split(l, dense_val_unbounded, thread, WARP_SIZE)
      .bound(dense_val_unbounded, dense_val, CO_FACTOR, BoundType::MaxExact)
      .reorder({block, warp, nnz, thread, dense_val})
      .unroll(dense_val, CO_FACTOR)
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleTTVGPU(IndexStmt stmt, Tensor<double> B, IndexExpr pre,
                         int NNZ_PER_WARP = 8 * 32, int BLOCK_SIZE = 256) {
  int NNZ_PER_TB = NNZ_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar jk("jk"), f("f"), fpos("fpos"), block("block"), fpos1("fpos1"), warp("warp"),
      fpos2("fpos2"), thread("thread"), thread_nz("thread_nz"), thread_nz_pre("thread_nz_pre");
  TensorVar precomputed("precomputed", Type(Float64, {Dimension(thread_nz)}), taco::dense);
  return stmt.fuse(j, k, jk)
      .fuse(i, jk, f)
      .pos(f, fpos, B(i, j, k))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, fpos2, NNZ_PER_WARP)
      .split(fpos2, thread, thread_nz, NNZ_PER_WARP / WARP_SIZE)
      .reorder({block, warp, thread, thread_nz})
      .precompute(pre, thread_nz, thread_nz_pre, precomputed)
      .unroll(thread_nz_pre, NNZ_PER_WARP / WARP_SIZE)
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

IndexStmt scheduleMTTKRPGPU(IndexStmt stmt, Tensor<double> B, int NNZ_PER_WARP = 16,
                            int BLOCK_SIZE = 256) {
  int NNZ_PER_TB = NNZ_PER_WARP * (BLOCK_SIZE / WARP_SIZE);
  IndexVar kl("kl"), f("f"), fpos("fpos"), block("block"), fpos1("fpos1"), warp("warp"),
      nnz("nnz"), dense_val_unbounded("dense_val_unbounded"), dense_val("dense_val"),
      thread("thread");
  return stmt.reorder({i, k, l, j})
      .fuse(k, l, kl)
      .fuse(i, kl, f)
      .pos(f, fpos, B(i, k, l))
      .split(fpos, block, fpos1, NNZ_PER_TB)
      .split(fpos1, warp, nnz, NNZ_PER_WARP)
      .split(j, dense_val_unbounded, thread, WARP_SIZE)
      .bound(dense_val_unbounded, dense_val, 1, BoundType::MaxExact)
      .reorder({block, warp, dense_val, thread, nnz})
      .parallelize(block, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
      .parallelize(warp, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces)
      .parallelize(thread, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
}

// ---- inputs: the reference tests' rand() draws ----

float rand_float() { return (float)rand() / (float)(RAND_MAX); }
double tval(float rf, float scale) { return (double)((int)(rf * 3 / scale)); }

// Dense copy of a tensor whose modes are all dense (row-major over the
// tensor's dimensions, whatever its mode ordering).
std::vector<double> dense_values(TensorBase& t) {
  taco_tensor_t* d = t.getTacoTensorT();
  const int order = d->order;
  std::vector<int> dims(order);
  size_t n = 1;
  for (int q = 0; q < order; ++q) {
    dims[q] = d->dimensions[q];
    n *= (size_t)dims[q];
  }
  std::vector<double> out(n, 0.0);
  if (order == 0) {
    out.assign(1, t.getComponentType() == Int32 ? (double)((int*)d->vals)[0]
                                                 : ((double*)d->vals)[0]);
    return out;
  }
  // storage position p over modes in storage order mode_ordering
  std::vector<int> idx(order, 0);
  for (size_t p = 0; p < n; ++p) {
    size_t rem = p;
    for (int q = order - 1; q >= 0; --q) {
      const int mode = d->mode_ordering[q];
      idx[mode] = (int)(rem % (size_t)dims[mode]);
      rem /= (size_t)dims[mode];
    }
    size_t r = 0;
    for (int q = 0; q < order; ++q) r = r * (size_t)dims[q] + (size_t)idx[q];
    out[r] = t.getComponentType() == Int32 ? (double)((int*)d->vals)[p] : ((double*)d->vals)[p];
  }
  return out;
}

struct Ctx {
  bool compileOnly = false;
  std::string msg;
  std::string source;  // emit mode: the generated module text
  bool emit = false;
};

// compile (and run) `out` under `stmt`; false when compile-only
bool run(Ctx& c, TensorBase& out, IndexStmt stmt, bool evaluate = false) {
  if (c.emit) {
    if (getenv("TACO_DROPIN_TRACE")) fprintf(stderr, "emit: concretize\n");
    IndexStmt s = scalarPromote(stmt.concretize());
    if (getenv("TACO_DROPIN_TRACE")) fprintf(stderr, "emit: lower+print\n");
    std::stringstream ss;
    ir::CodeGen_HIP src(ss, ir::CodeGen::ImplementationGen);
    src.compile(lower(s, "assemble", true, false), true);
    src.compile(lower(s, "compute", false, true), false);
    std::stringstream sh;
    ir::CodeGen_HIP::generateShim(lower(s, "compute", false, true), sh);
    c.source = ss.str() + "\n" + sh.str();
    return false;
  }
  out.compile(stmt);
  if (c.compileOnly) return false;
  if (evaluate) {
    out.compute();
  } else {
    out.assemble();
    out.compute();
  }
  return true;
}

bool check_equal(Ctx& c, const std::vector<double>& got, const std::vector<double>& want) {
  if (got.size() != want.size()) {
    c.msg = "size mismatch";
    return false;
  }
  size_t bad = 0, first = 0;
  for (size_t q = 0; q < got.size(); ++q)
    if (!(got[q] == want[q])) {
      if (!bad) first = q;
      ++bad;
    }
  if (bad) {
    std::ostringstream os;
    os << bad << " of " << got.size() << " values differ; first at " << first << ": got "
       << got[first] << " want " << want[first] << "; last error: " << taco_hip_last_error();
    c.msg = os.str();
    return false;
  }
  std::ostringstream os;
  os << got.size() << " values exact";
  c.msg = os.str();
  return true;
}

// ---- statements ----

struct SpmvInputs {
  int NI = 1021 / 10, NJ = 1039 / 10;
  std::vector<std::map<int, double>> rows;
  std::vector<double> x;
};

void fill_spmv(Tensor<double>& A, Tensor<double>& x, SpmvInputs& in, unsigned seed, float sparsity) {
  srand(seed);
  in.rows.assign(in.NI, {});
  for (int a = 0; a < in.NI; a++)
    for (int b = 0; b < in.NJ; b++) {
      float rf = rand_float();
      if (rf < sparsity) {
        double v = tval(rf, sparsity);
        A.insert({a, b}, v);
        in.rows[a][b] = v;
      }
    }
  in.x.assign(in.NJ, 0.0);
  for (int b = 0; b < in.NJ; b++) {
    float rf = rand_float();
    in.x[b] = tval(rf, sparsity);
    x.insert({b}, in.x[b]);
  }
  x.pack();
  A.pack();
}

bool test_spmv(Ctx& c, const std::string& which) {
  SpmvInputs in;
  Tensor<double> A("A", {in.NI, in.NJ}, CSR);
  Tensor<double> x("x", {in.NJ}, Format({Dense}));
  Tensor<double> y("y", {in.NI}, Format({Dense}));
  fill_spmv(A, x, in, 94353, .01f);
  IndexExpr pre = A(i, j) * x(j);
  y(i) = pre;
  IndexStmt stmt = y.getAssignment().concretize();
  if (which == "spmv_rows_gpu") stmt = scheduleSpMVRowsGPU(stmt, A);
  else if (which == "spmv_splitpos_gpu") stmt = scheduleSpMVSplitPosGPU(stmt, A);
  else if (which == "spmv_thread_row_gpu") stmt = scheduleSpMVThreadPerRowGPU(stmt);
  else if (which == "spmv_precompute_gpu") stmt = scheduleSpMVGPU(stmt, A, pre);
  else if (which == "spmv_default_gpu")
    stmt = parallelizeOuterLoop(insertTemporaries(reorderLoopsTopologically(
        makeConcreteNotation(makeReductionNotation(y.getAssignment())))));
  if (!run(c, y, stmt)) return true;
  std::vector<double> got = dense_values(y);
  std::vector<double> want(in.NI, 0.0);
  for (int a = 0; a < in.NI; a++)
    for (auto& e : in.rows[a]) want[a] += e.second * in.x[e.first];
  if (which == "spmv_precompute_gpu") {
    // each non-empty row holds one of its entries' products
    int rowsChecked = 0;
    for (int a = 0; a < in.NI; a++) {
      if (in.rows[a].empty()) continue;
      bool ok = false;
      for (auto& e : in.rows[a]) ok |= got[a] == e.second * in.x[e.first];
      if (!ok) {
        std::ostringstream os;
        os << "row " << a << " holds " << got[a] << ", not one of its products";
        c.msg = os.str();
        return false;
      }
      ++rowsChecked;
    }
    c.msg = std::to_string(rowsChecked) + " rows hold one of their products";
    return true;
  }
  return check_equal(c, got, want);
}

bool test_spmm(Ctx& c, const std::string& which) {
  const int NI = 1021 / 10, NJ = 1039 / 10, NK = 128;
  const float SP = .3f;
  const bool dcsr = which == "spmm_dcsr_gpu";
  Tensor<double> A("A", {NI, NJ}, dcsr ? Format({Sparse, Sparse}) : CSR);
  Tensor<double> B("B", {NJ, NK}, {Dense, Dense});
  std::vector<std::vector<double>> Ad(NI, std::vector<double>(NJ, 0.0));
  std::vector<std::vector<double>> Bd(NJ, std::vector<double>(NK, 0.0));
  srand(dcsr ? 25643 : 434321);
  for (int a = 0; a < NI; a++)
    for (int b = 0; b < NJ; b++) {
      float rf = rand_float();
      if (rf < SP) {
        Ad[a][b] = tval(rf, SP);
        A.insert({a, b}, Ad[a][b]);
      }
    }
  for (int b = 0; b < NJ; b++)
    for (int q = 0; q < NK; q++) {
      float rf = rand_float();
      Bd[b][q] = tval(rf, SP);
      B.insert({b, q}, Bd[b][q]);
    }
  A.pack();
  B.pack();
  Format cf = which == "spmm_gpu" ? Format({{Dense, Dense}, {1, 0}}) : Format({Dense, Dense});
  Tensor<double> C("C", {NI, NK}, cf);
  IndexStmt stmt;
  if (which == "spmm_gpu") {
    IndexExpr pre = A(i, j);
    C(i, k) = B(j, k) * pre;
    stmt = scheduleSpMMGPU(C.getAssignment().concretize(), A);
  } else {
    C(i, k) = A(i, j) * B(j, k);
    stmt = C.getAssignment().concretize();
    stmt = dcsr ? scheduleSpMMNZRowsGPU(stmt, A) : scheduleSpMMRowsGPU(stmt);
  }
  if (!run(c, C, stmt)) return true;
  std::vector<double> want((size_t)NI * NK, 0.0);
  for (int a = 0; a < NI; a++)
    for (int b = 0; b < NJ; b++)
      for (int q = 0; q < NK; q++) want[(size_t)a * NK + q] += Ad[a][b] * Bd[b][q];
  return check_equal(c, dense_values(C), want);
}

bool test_sddmm(Ctx& c) {
  const int NI = 1021 / 10, NK = 1039 / 10, NJ = 128;
  const float SP = .3f;
  Tensor<double> A("A", {NI, NK}, {Dense, Dense});
  Tensor<double> B("B", {NI, NK}, CSR);
  Tensor<double> C("C", {NI, NJ}, {Dense, Dense});
  Tensor<double> D("D", {NJ, NK}, {Dense, Dense});
  std::vector<double> Bd((size_t)NI * NK, 0.0), Cd((size_t)NI * NJ), Dd((size_t)NJ * NK);
  srand(535366);
  for (int a = 0; a < NI; a++)
    for (int b = 0; b < NJ; b++) {
      float rf = rand_float();
      Cd[(size_t)a * NJ + b] = tval(rf, SP);
      C.insert({a, b}, Cd[(size_t)a * NJ + b]);
    }
  for (int a = 0; a < NI; a++)
    for (int q = 0; q < NK; q++) {
      float rf = rand_float();
      if (rf < SP) {
        Bd[(size_t)a * NK + q] = tval(rf, SP);
        B.insert({a, q}, Bd[(size_t)a * NK + q]);
      }
    }
  for (int b = 0; b < NJ; b++)
    for (int q = 0; q < NK; q++) {
      float rf = rand_float();
      Dd[(size_t)b * NK + q] = tval(rf, SP);
      D.insert({b, q}, Dd[(size_t)b * NK + q]);
    }
  B.pack();
  C.pack();
  D.pack();
  A(i, k) = B(i, k) * C(i, j) * D(j, k);
  IndexStmt stmt = scheduleSDDMMGPU(A.getAssignment().concretize(), B);
  if (!run(c, A, stmt)) return true;
  std::vector<double> want((size_t)NI * NK, 0.0);
  for (int a = 0; a < NI; a++)
    for (int q = 0; q < NK; q++) {
      double s = 0;
      for (int b = 0; b < NJ; b++)
        s += Bd[(size_t)a * NK + q] * Cd[(size_t)a * NJ + b] * Dd[(size_t)b * NK + q];
      want[(size_t)a * NK + q] = s;
    }
  return check_equal(c, dense_values(A), want);
}

struct Sparse3 {
  std::vector<std::tuple<int, int, int, double>> e;
};

bool test_ttm(Ctx& c) {
  const int NI = 1021 / 40, NJ = 1039 / 40, NK = 1232 / 40, NL = 128;
  const float SP = .1f;
  Tensor<double> A("A", {NI, NJ, NL}, {Dense, Dense, Dense});
  Tensor<double> B("B", {NI, NJ, NK}, {Sparse, Sparse, Sparse});
  Tensor<double> C("C", {NK, NL}, {Dense, Dense});
  Sparse3 Bs;
  std::vector<double> Cd((size_t)NK * NL);
  srand(34644);
  for (int a = 0; a < NI; a++)
    for (int b = 0; b < NJ; b++)
      for (int q = 0; q < NK; q++) {
        float rf = rand_float();
        if (rf < SP) {
          double v = tval(rf, SP);
          B.insert({a, b, q}, v);
          Bs.e.emplace_back(a, b, q, v);
        }
      }
  for (int q = 0; q < NK; q++)
    for (int r = 0; r < NL; r++) {
      float rf = rand_float();
      Cd[(size_t)q * NL + r] = (double)((int)(rf * 3));
      C.insert({q, r}, Cd[(size_t)q * NL + r]);
    }
  B.pack();
  C.pack();
  A(i, j, l) = B(i, j, k) * C(k, l);
  IndexStmt stmt = scheduleTTMGPU(A.getAssignment().concretize(), B);
  if (!run(c, A, stmt)) return true;
  std::vector<double> want((size_t)NI * NJ * NL, 0.0);
  for (auto& t : Bs.e) {
    int a, b, q;
    double v;
    std::tie(a, b, q, v) = t;
    for (int r = 0; r < NL; r++) want[((size_t)a * NJ + b) * NL + r] += v * Cd[(size_t)q * NL + r];
  }
  return check_equal(c, dense_values(A), want);
}

bool test_ttv(Ctx& c) {
  const int NI = 1021 / 10, NJ = 1039 / 10, NK = 1057 / 10;
  const float SP = .3f;
  Tensor<double> A("A", {NI, NJ}, {Dense, Dense});
  Tensor<double> B("B", {NI, NJ, NK}, {Sparse, Sparse, Sparse});
  Tensor<double> cv("c", {NK}, Format({Dense}));
  Sparse3 Bs;
  std::vector<double> cd(NK);
  srand(35325);
  for (int a = 0; a < NI; a++)
    for (int b = 0; b < NJ; b++)
      for (int q = 0; q < NK; q++) {
        float rf = rand_float();
        if (rf < SP) {
          double v = tval(rf, SP);
          B.insert({a, b, q}, v);
          Bs.e.emplace_back(a, b, q, v);
        }
      }
  for (int q = 0; q < NK; q++) {
    float rf = rand_float();
    cd[q] = (double)((int)(rf * 3));
    cv.insert({q}, cd[q]);
  }
  B.pack();
  cv.pack();
  IndexExpr pre = B(i, j, k) * cv(k);
  A(i, j) = pre;
  IndexStmt stmt = scheduleTTVGPU(A.getAssignment().concretize(), B, pre);
  if (!run(c, A, stmt)) return true;
  // precompute consumer is a plain store: each (i, j) fiber holds one of its
  // products (see the file header)
  std::map<std::pair<int, int>, std::vector<double>> prods;
  for (auto& t : Bs.e) {
    int a, b, q;
    double v;
    std::tie(a, b, q, v) = t;
    prods[{a, b}].push_back(v * cd[q]);
  }
  std::vector<double> got = dense_values(A);
  for (auto& p : prods) {
    double g = got[(size_t)p.first.first * NJ + p.first.second];
    bool ok = false;
    for (double v : p.second) ok |= g == v;
    if (!ok) {
      std::ostringstream os;
      os << "fiber (" << p.first.first << "," << p.first.second << ") holds " << g
         << ", not one of its products";
      c.msg = os.str();
      return false;
    }
  }
  c.msg = std::to_string(prods.size()) + " fibers hold one of their products";
  return true;
}

bool test_mttkrp(Ctx& c) {
  const int NI = 1021 / 40, NJ = 32, NK = 1039 / 40, NL = 1232 / 40;
  const float SP = .1f;
  Tensor<double> A("A", {NI, NJ}, {Dense, Dense});
  Tensor<double> B("B", {NI, NK, NL}, {Sparse, Sparse, Sparse});
  Tensor<double> C("C", {NK, NJ}, {Dense, Dense});
  Tensor<double> D("D", {NL, NJ}, {Dense, Dense});
  Sparse3 Bs;
  std::vector<double> Cd((size_t)NK * NJ), Dd((size_t)NL * NJ);
  srand(5464164);
  for (int a = 0; a < NI; a++)
    for (int q = 0; q < NK; q++)
      for (int r = 0; r < NL; r++) {
        float rf = rand_float();
        if (rf < SP) {
          double v = tval(rf, SP);
          B.insert({a, q, r}, v);
          Bs.e.emplace_back(a, q, r, v);
        }
      }
  for (int q = 0; q < NK; q++)
    for (int b = 0; b < NJ; b++) {
      float rf = rand_float();
      Cd[(size_t)q * NJ + b] = (double)((int)(rf * 3));
      C.insert({q, b}, Cd[(size_t)q * NJ + b]);
    }
  for (int r = 0; r < NL; r++)
    for (int b = 0; b < NJ; b++) {
      float rf = rand_float();
      Dd[(size_t)r * NJ + b] = (double)((int)(rf * 3));
      D.insert({r, b}, Dd[(size_t)r * NJ + b]);
    }
  B.pack();
  C.pack();
  D.pack();
  A(i, j) = B(i, k, l) * C(k, j) * D(l, j);
  IndexStmt stmt = scheduleMTTKRPGPU(A.getAssignment().concretize(), B);
  if (!run(c, A, stmt)) return true;
  std::vector<double> want((size_t)NI * NJ, 0.0);
  for (auto& t : Bs.e) {
    int a, q, r;
    double v;
    std::tie(a, q, r, v) = t;
    for (int b = 0; b < NJ; b++)
      want[(size_t)a * NJ + b] += v * Cd[(size_t)q * NJ + b] * Dd[(size_t)r * NJ + b];
  }
  return check_equal(c, dense_values(A), want);
}

bool test_reduction(Ctx& c, OutputRaceStrategy strategy) {
  Tensor<double> A("A", {8}, Format({Sparse}));
  Tensor<double> B("B", {8}, Format({Dense}));
  Tensor<double> C("C");
  double want = 0;
  for (int q = 0; q < 8; q++) {
    if (q % 2 == 0) A.insert({q}, (double)q);
    B.insert({q}, (double)q);
    if (q % 2 == 0) want += (double)q * q;
  }
  A.pack();
  B.pack();
  IndexVar block("block"), thread("thread"), i0("i0"), i1("i1");
  C = A(i) * B(i);
  IndexStmt stmt = C.getAssignment().concretize();
  stmt = stmt.split(i, i0, i1, 2)
             .split(i0, block, thread, 2)
             .parallelize(block, ParallelUnit::GPUBlock, strategy)
             .parallelize(thread, ParallelUnit::GPUThread, strategy);
  if (!run(c, C, stmt)) return true;
  return check_equal(c, dense_values(C), {want});
}

bool test_tiling(Ctx& c) {
  Tensor<double> A("A", {8, 8}, CSR);
  Tensor<double> B("B", {8, 8}, CSC);
  Tensor<double> C("C", {8, 8}, Format({Dense, Dense}));
  std::vector<double> Ad(64, 0.0), Bd(64, 0.0);
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) {
      if ((a + b) % 2 == 0) {
        A.insert({a, b}, (double)(a + b));
        Ad[a * 8 + b] = a + b;
      }
      if ((a + b) != 2 && (a + b) != 3 && (a + b) != 4) {
        B.insert({a, b}, (double)(a + b));
        Bd[a * 8 + b] = a + b;
      }
    }
  A.pack();
  B.pack();
  IndexVar i0("i0"), i1("i1"), j0("j0"), j1("j1"), k0("k0"), k1("k1");
  IndexVar ii("i"), jj("j"), kk("k");
  C(ii, jj) = A(ii, kk) * B(kk, jj);
  IndexStmt stmt = C.getAssignment().concretize();
  stmt = stmt.split(ii, i0, i1, 2)
             .split(jj, j0, j1, 2)
             .split(kk, k0, k1, 2)
             .reorder({i0, j0, k0, i1, j1, k1})
             .parallelize(i0, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces)
             .parallelize(j0, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
  if (!run(c, C, stmt)) return true;
  std::vector<double> want(64, 0.0);
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++)
      for (int q = 0; q < 8; q++) want[a * 8 + b] += Ad[a * 8 + q] * Bd[q * 8 + b];
  return check_equal(c, dense_values(C), want);
}

bool test_windowing(Ctx& c, ModeFormat x, ModeFormat y) {
  const int dim = 10;
  Tensor<int> a("a", {dim, dim}, Format{Dense, x});
  Tensor<int> b("b", {dim, dim}, Format{Dense, y});
  Tensor<int> cc("c", {2, 2}, Format{Dense, Dense});
  for (int p = 0; p < dim; p++)
    for (int q = 0; q < dim; q++) {
      a.insert({p, q}, p + q);
      b.insert({p, q}, p + q);
    }
  a.pack();
  b.pack();
  IndexVar ii("i"), jj("j"), i1("i1"), i2("i2"), i3("i3"), i4("i4");
  cc(ii, jj) = a(ii(4, 6), jj(4, 6)) + b(ii(2, 4), jj(2, 4));
  IndexStmt stmt = cc.getAssignment().concretize();
  stmt = stmt.split(ii, i1, i2, 512)
             .split(i2, i3, i4, 32)
             .parallelize(i1, ParallelUnit::GPUBlock, OutputRaceStrategy::NoRaces)
             .parallelize(i3, ParallelUnit::GPUWarp, OutputRaceStrategy::NoRaces)
             .parallelize(i4, ParallelUnit::GPUThread, OutputRaceStrategy::NoRaces);
  if (c.emit) return run(c, cc, stmt), true;
  cc.compile(stmt);
  if (c.compileOnly) return true;
  cc.evaluate();
  return check_equal(c, dense_values(cc), {12, 14, 14, 16});
}

const std::vector<std::string> kTests = {
    "spmv_rows_gpu", "spmv_splitpos_gpu", "spmv_thread_row_gpu", "spmv_default_gpu",
    "spmv_precompute_gpu", "spmm_gpu", "spmm_rows_gpu", "spmm_dcsr_gpu", "sddmm_gpu",
    "ttm_gpu", "ttv_gpu", "mttkrp_gpu", "atomic_reduction_gpu", "temporary_reduction_gpu",
    "tiling_gpu", "windowing_gpu_dd", "windowing_gpu_ds", "windowing_gpu_sd",
    "windowing_gpu_ss"};

bool dispatch(Ctx& c, const std::string& name) {
  if (name.rfind("spmv_", 0) == 0) return test_spmv(c, name);
  if (name.rfind("spmm_", 0) == 0) return test_spmm(c, name);
  if (name == "sddmm_gpu") return test_sddmm(c);
  if (name == "ttm_gpu") return test_ttm(c);
  if (name == "ttv_gpu") return test_ttv(c);
  if (name == "mttkrp_gpu") return test_mttkrp(c);
  if (name == "atomic_reduction_gpu") return test_reduction(c, OutputRaceStrategy::Atomics);
  if (name == "temporary_reduction_gpu") return test_reduction(c, OutputRaceStrategy::Temporary);
  if (name == "tiling_gpu") return test_tiling(c);
  if (name.rfind("windowing_gpu_", 0) == 0) {
    ModeFormat x = name[14] == 'd' ? Dense : Sparse;
    ModeFormat y = name[15] == 'd' ? Dense : Sparse;
    return test_windowing(c, x, y);
  }
  c.msg = "unknown test " + name;
  return false;
}

int finish(const Ctx& c, bool ok, char* msg, int msglen) {
  if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", c.msg.c_str());
  return ok ? 0 : 1;
}

}  // namespace

extern "C" {

/* Comma-separated statement names. */
const char* taco_dropin_list(void) {
  static std::string s;
  if (s.empty())
    for (auto& n : kTests) s += (s.empty() ? "" : ",") + n;
  return s.c_str();
}

/* mode 0: compile, assemble, compute and check; 1: compile only (fills the
 * module cache, needs no GPU).  Returns 0 on success; msg gets the detail. */
int taco_dropin_run(const char* name, int mode, char* msg, int msglen) {
  Ctx c;
  c.compileOnly = mode == 1;
  try {
    bool ok = dispatch(c, name);
    if (c.compileOnly && ok) c.msg = "compiled";
    return finish(c, ok, msg, msglen);
  } catch (std::exception& e) {
    c.msg = std::string("exception: ") + e.what();
    return finish(c, false, msg, msglen);
  }
}

/* The HIP module text CodeGen_HIP prints for a statement (assemble and
 * compute, plus the compute shim), without compiling it. */
int taco_dropin_emit(const char* name, char* buf, int buflen) {
  Ctx c;
  c.emit = true;
  try {
    dispatch(c, name);
  } catch (std::exception& e) {
    c.source = std::string("exception: ") + e.what();
    snprintf(buf, (size_t)buflen, "%s", c.source.c_str());
    return 1;
  }
  if ((int)c.source.size() + 1 > buflen) return -(int)(c.source.size() + 1);
  memcpy(buf, c.source.c_str(), c.source.size() + 1);
  return 0;
}

}  // extern "C"
