// Driver over the real taco library, compiled from the reference's own
// sources into oracle/_ref/libtaco_ref.so (see oracle/Makefile, target `ref`).
//
// TEST INFRASTRUCTURE ONLY: this is the strongest form of the CPU oracle —
// taco's own CodeGen_C output, JIT-compiled by taco's own ir::Module with the
// system C compiler — used by tests to pin oracle/taco_cpu.c and by bench.py's
// cpu_baseline leg as the reference CPU timing.  The product never links it.
//
// Statement construction follows the reference's own API:
//   y(i) = A(i,j) * x(j)          tensor.h Tensor/IndexVar, tensor.cpp:605-674
//   C(i,k) = A(i,j) * B(j,k)      dense C row-major ({Dense,Dense})
//   SpGEMM two-phase schedule     test/tests-scheduling-eval.cpp:68-94
// Kernels are built with taco::compile(IndexStmt) (index_notation/kernel.cpp:
// 111-128), which applies parallelizeOuterLoop (transformations.cpp:1562-1606)
// to the compute function, i.e. taco's default OpenMP CPU parallelisation.

#include "taco/tensor.h"
#include "taco/index_notation/index_notation.h"
#include "taco/index_notation/transformations.h"
#include "taco/index_notation/kernel.h"
#include "taco/storage/index.h"
#include "taco/storage/array.h"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <sstream>
#include <vector>

using namespace taco;

namespace {

std::string g_last_error;
std::string g_last_source;

double now_ms() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double, std::milli>(
             clk::now().time_since_epoch()).count();
}

template <typename T>
Tensor<T> dense_vector(const char* name, int n, T* data) {
  Tensor<T> t(name, {n}, Format({Dense}));
  TensorStorage st = t.getStorage();
  st.setValues(makeArray(data, (size_t)n, Array::UserOwns));
  t.setStorage(st);
  return t;
}

template <typename T>
Tensor<T> dense_matrix(const char* name, int r, int c, T* data) {
  Tensor<T> t(name, {r, c}, Format({Dense, Dense}));
  TensorStorage st = t.getStorage();
  st.setValues(makeArray(data, (size_t)r * (size_t)c, Array::UserOwns));
  t.setStorage(st);
  return t;
}

template <typename T>
Tensor<T> csr_matrix(const char* name, int m, int n, int32_t* pos,
                     int32_t* crd, T* vals) {
  TensorBase b = makeCSR<T>(name, {m, n}, (int*)pos, (int*)crd, vals,
                            Array::UserOwns);
  return Tensor<T>(b);
}

// Lower, JIT and run `stmt` through taco's Kernel API.  `args` are result
// storages first, then operands in getArguments order (tensor.cpp:779-806).
int run_kernel(IndexStmt stmt, std::vector<TensorStorage> args, int reps,
               double* times_ms) {
  Kernel k = compile(stmt);
  {
    std::stringstream ss;
    ss << k;
    g_last_source = ss.str();
  }
  if (!k.assemble(args)) return 2;
  if (!k.compute(args)) return 3;
  for (int r = 0; r < reps; ++r) {
    double t0 = now_ms();
    if (!k.compute(args)) return 3;
    times_ms[r] = now_ms() - t0;
  }
  return 0;
}

template <typename T>
int spmv_impl(int m, int n, int32_t* pos, int32_t* crd, T* vals, T* x, T* y,
              int reps, double* times_ms) {
  Tensor<T> A = csr_matrix<T>("A", m, n, pos, crd, vals);
  Tensor<T> xv = dense_vector<T>("x", n, x);
  Tensor<T> yv("y", {m}, Format({Dense}));
  IndexVar i("i"), j("j");
  yv(i) = A(i, j) * xv(j);
  IndexStmt stmt = makeConcreteNotation(makeReductionNotation(yv.getAssignment()));
  stmt = reorderLoopsTopologically(stmt);
  stmt = insertTemporaries(stmt);
  std::vector<TensorStorage> args = {yv.getStorage(), A.getStorage(),
                                     xv.getStorage()};
  int rc = run_kernel(stmt, args, reps, times_ms);
  if (rc) return rc;
  const T* out = (const T*)args[0].getValues().getData();
  std::memcpy(y, out, sizeof(T) * (size_t)m);
  return 0;
}

template <typename T>
int spmm_impl(int m, int n, int kdim, int32_t* pos, int32_t* crd, T* vals,
              T* B, T* C, int reps, double* times_ms) {
  Tensor<T> A = csr_matrix<T>("A", m, n, pos, crd, vals);
  Tensor<T> Bt = dense_matrix<T>("B", n, kdim, B);
  Tensor<T> Ct("C", {m, kdim}, Format({Dense, Dense}));
  IndexVar i("i"), j("j"), k("k");
  Ct(i, k) = A(i, j) * Bt(j, k);
  IndexStmt stmt = makeConcreteNotation(makeReductionNotation(Ct.getAssignment()));
  stmt = reorderLoopsTopologically(stmt);
  stmt = insertTemporaries(stmt);
  std::vector<TensorStorage> args = {Ct.getStorage(), A.getStorage(),
                                     Bt.getStorage()};
  int rc = run_kernel(stmt, args, reps, times_ms);
  if (rc) return rc;
  const T* out = (const T*)args[0].getValues().getData();
  std::memcpy(C, out, sizeof(T) * (size_t)m * (size_t)kdim);
  return 0;
}

}  // namespace

extern "C" {

const char* taco_ref_last_error() { return g_last_error.c_str(); }
const char* taco_ref_last_source() { return g_last_source.c_str(); }

void taco_ref_set_num_threads(int n) {
  taco_set_num_threads(n);
  taco_set_parallel_schedule(ParallelSchedule::Static, 0);
}

#define GUARD(...)                                   \
  try {                                              \
    __VA_ARGS__;                                     \
  } catch (std::exception & e) {                     \
    g_last_error = e.what();                         \
    return 1;                                        \
  } catch (...) {                                    \
    g_last_error = "unknown exception";              \
    return 1;                                        \
  }

int taco_ref_spmv_f64(int m, int n, int32_t* pos, int32_t* crd, double* vals,
                      double* x, double* y, int reps, double* times_ms) {
  GUARD(return spmv_impl<double>(m, n, pos, crd, vals, x, y, reps, times_ms));
}

int taco_ref_spmv_f32(int m, int n, int32_t* pos, int32_t* crd, float* vals,
                      float* x, float* y, int reps, double* times_ms) {
  GUARD(return spmv_impl<float>(m, n, pos, crd, vals, x, y, reps, times_ms));
}

int taco_ref_spmm_f64(int m, int n, int k, int32_t* pos, int32_t* crd,
                      double* vals, double* B, double* C, int reps,
                      double* times_ms) {
  GUARD(return spmm_impl<double>(m, n, k, pos, crd, vals, B, C, reps,
                                 times_ms));
}

int taco_ref_spmm_f32(int m, int n, int k, int32_t* pos, int32_t* crd,
                      float* vals, float* B, float* C, int reps,
                      double* times_ms) {
  GUARD(return spmm_impl<float>(m, n, k, pos, crd, vals, B, C, reps,
                                times_ms));
}

// Two-phase SpGEMM C = A*B with C in CSR, under the reference's CPU SpGEMM
// schedule (reorderLoopsTopologically, precompute into a dense workspace w,
// assemble(C, Insert, separately schedulable), parallelize i and the query
// loop over CPU threads).  Outputs are malloc'ed copies owned by the caller
// (free with taco_ref_free).  asm_ms/cmp_ms receive `reps` timings each.
int taco_ref_spgemm_f64(int m, int n, int kdim, int32_t* apos, int32_t* acrd,
                        double* avals, int32_t* bpos, int32_t* bcrd,
                        double* bvals, int32_t** cpos_out, int32_t** ccrd_out,
                        double** cvals_out, int reps, double* asm_ms,
                        double* cmp_ms) {
  GUARD({
    Tensor<double> A = csr_matrix<double>("A", m, n, apos, acrd, avals);
    Tensor<double> B = csr_matrix<double>("B", n, kdim, bpos, bcrd, bvals);
    Tensor<double> C("C", {m, kdim}, CSR);
    IndexVar i("i"), j("j"), k("k");
    C(i, k) = A(i, j) * B(j, k);
    IndexStmt stmt = C.getAssignment().concretize();
    Assignment assign = stmt.as<Forall>().getStmt().as<Forall>().getStmt()
                            .as<Forall>().getStmt().as<Assignment>();
    TensorVar result = assign.getLhs().getTensorVar();
    stmt = reorderLoopsTopologically(stmt);
    IndexVar jw = assign.getLhs().getIndexVars()[1];
    TensorVar w("w", Type(result.getType().getDataType(),
                          {result.getType().getShape().getDimension(1)}),
                taco::dense);
    stmt = stmt.precompute(assign.getRhs(), jw, jw, w);
    stmt = stmt.assemble(result, AssembleStrategy::Insert, true);
    auto qi_stmt = stmt.as<Assemble>().getQueries();
    IndexVar qi;
    if (isa<Where>(qi_stmt)) {
      qi = qi_stmt.as<Where>().getConsumer().as<Forall>().getIndexVar();
    } else {
      qi = qi_stmt.as<Forall>().getIndexVar();
    }
    stmt = stmt.parallelize(i, ParallelUnit::CPUThread,
                            OutputRaceStrategy::NoRaces)
               .parallelize(qi, ParallelUnit::CPUThread,
                            OutputRaceStrategy::NoRaces);

    Kernel kern = compile(stmt);
    {
      std::stringstream ss;
      ss << kern;
      g_last_source = ss.str();
    }
    std::vector<TensorStorage> args = {C.getStorage(), A.getStorage(),
                                       B.getStorage()};
    for (int r = 0; r < (reps > 0 ? reps : 1); ++r) {
      double t0 = now_ms();
      if (!kern.assemble(args)) return 2;
      double t1 = now_ms();
      if (!kern.compute(args)) return 3;
      double t2 = now_ms();
      if (reps > 0) {
        asm_ms[r] = t1 - t0;
        cmp_ms[r] = t2 - t1;
      }
    }
    TensorStorage cs = args[0];
    Index idx = cs.getIndex();
    Array posA = idx.getModeIndex(1).getIndexArray(0);
    Array crdA = idx.getModeIndex(1).getIndexArray(1);
    const int32_t* cp = (const int32_t*)posA.getData();
    int32_t nnz = cp[m];
    *cpos_out = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(m + 1));
    *ccrd_out = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    *cvals_out = (double*)std::malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    std::memcpy(*cpos_out, cp, sizeof(int32_t) * (size_t)(m + 1));
    std::memcpy(*ccrd_out, crdA.getData(), sizeof(int32_t) * (size_t)nnz);
    std::memcpy(*cvals_out, cs.getValues().getData(),
                sizeof(double) * (size_t)nnz);
    return 0;
  });
}

void taco_ref_free(void* p) { std::free(p); }

}  // extern "C"
