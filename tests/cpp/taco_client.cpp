// A C++ caller of the C ABI, doing for y(i) = A(i,j) * x(j) and the CSR
// product C(i,k) = A(i,j) * B(j,k) what taco's TensorBase does on its GPU path
// (src/tensor.cpp:605-858): build taco_tensor_t descriptors, ask CodeGen_HIP
// for the module source, compile it with the module loader, and call the
// _shim_ entry points with taco's argument order (results first).
//
//   taco_client generate   lower + compile both modules (no GPU needed)
//   taco_client run        also run them on the GPU and check the results
//                          against plain loops (integer-valued data: exact)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "taco/taco_tensor_t.h"
#include "taco_hip.h"

namespace {

int fail(const char* what) {
  std::fprintf(stderr, "FAIL %s: %s\n", what, taco_hip_last_error());
  return 1;
}

struct Csr {
  int32_t m, n;
  std::vector<int32_t> pos, crd;
  std::vector<double> vals;
};

// Rows of 0..span entries at distinct sorted columns, small integer values.
Csr make_csr(int32_t m, int32_t n, int span, uint32_t seed) {
  Csr a{m, n, {0}, {}, {}};
  uint32_t s = seed;
  auto next = [&]() { s = s * 1664525u + 1013904223u; return s >> 8; };
  std::vector<char> used((size_t)n, 0);
  for (int32_t i = 0; i < m; ++i) {
    const int k = (int)(next() % (uint32_t)(span + 1));
    std::vector<int32_t> cols;
    for (int t = 0; t < k; ++t) {
      const int32_t c = (int32_t)(next() % (uint32_t)n);
      if (!used[(size_t)c]) { used[(size_t)c] = 1; cols.push_back(c); }
    }
    std::sort(cols.begin(), cols.end());
    for (int32_t c : cols) {
      used[(size_t)c] = 0;
      a.crd.push_back(c);
      a.vals.push_back((double)((int)(next() % 7) - 3));
    }
    a.pos.push_back((int32_t)a.crd.size());
  }
  return a;
}

template <typename T>
T* to_device(const std::vector<T>& h) {
  T* d = (T*)taco_hip_unified_alloc(sizeof(T) * h.size());
  if (!h.empty() && (!d || taco_hip_memcpy(d, h.data(), sizeof(T) * h.size(), 0))) return nullptr;
  return d;
}

template <typename T>
std::vector<T> to_host(const void* d, size_t n) {
  std::vector<T> h(n);
  if (n && taco_hip_memcpy(h.data(), d, sizeof(T) * n, 1)) h.clear();
  return h;
}

// CSR descriptor (dense rows, compressed columns) over device arrays.
taco_tensor_t* csr_desc(int32_t m, int32_t n, int32_t* pos, int32_t* crd, double* vals,
                        int32_t nnz) {
  int32_t dims[2] = {m, n}, order[2] = {0, 1};
  taco_mode_t types[2] = {taco_mode_dense, taco_mode_sparse};
  double fill = 0;
  taco_tensor_t* t = init_taco_tensor_t(2, 64, dims, order, types, &fill);
  t->indices[0][0] = (uint8_t*)&t->dimensions[0];
  t->indices[1][0] = (uint8_t*)pos;
  t->indices[1][1] = (uint8_t*)crd;
  t->vals = (uint8_t*)vals;
  t->vals_size = nnz;
  return t;
}

taco_tensor_t* vec_desc(int32_t n, double* vals) {
  int32_t dims[1] = {n}, order[1] = {0};
  taco_mode_t types[1] = {taco_mode_dense};
  double fill = 0;
  taco_tensor_t* t = init_taco_tensor_t(1, 64, dims, order, types, &fill);
  t->indices[0][0] = (uint8_t*)&t->dimensions[0];
  t->vals = (uint8_t*)vals;
  t->vals_size = n;
  return t;
}

taco_hip_module* build(const taco_hip_gen_spec& spec, const char* what) {
  char* src = taco_hip_generate(&spec, 0);
  if (!src) { fail(what); return nullptr; }
  taco_hip_module* mod = taco_hip_module_create();
  const int rc = taco_hip_module_set_source(mod, src) || taco_hip_module_compile(mod);
  taco_hip_free_string(src);
  if (rc) { fail(what); taco_hip_module_destroy(mod); return nullptr; }
  std::printf("%s: %s\n", what, taco_hip_module_path(mod));
  return mod;
}

}  // namespace

int main(int argc, char** argv) {
  const bool run = argc > 1 && std::strcmp(argv[1], "run") == 0;
  // scheduleSpMVGPU (NNZ_PER_THREAD 8, BLOCK_SIZE 256, WARP_SIZE 32) and the
  // two-phase SpGEMM.
  const taco_hip_gen_spec spmv_spec = {TACO_HIP_OP_SPMV, TACO_HIP_GEN_NNZ_SPLIT, 64, 256, 8, 32, 0, 0};
  const taco_hip_gen_spec spgemm_spec = {TACO_HIP_OP_SPGEMM, TACO_HIP_GEN_TWO_PHASE, 64, 256, 0, 0, 0, 0};
  taco_hip_module* spmv = build(spmv_spec, "spmv module");
  taco_hip_module* spgemm = build(spgemm_spec, "spgemm module");
  if (!spmv || !spgemm) return 1;
  if (!run) {
    taco_hip_module_destroy(spmv);
    taco_hip_module_destroy(spgemm);
    std::printf("generate ok\n");
    return 0;
  }
  if (taco_hip_device_count() < 1) return fail("no GPU");

  const Csr A = make_csr(3000, 2000, 40, 7), B = make_csr(2000, 2500, 6, 11);
  std::vector<double> x((size_t)A.n);
  for (int32_t j = 0; j < A.n; ++j) x[(size_t)j] = (double)(j % 5 - 2);

  int32_t* apos = to_device(A.pos);
  int32_t* acrd = to_device(A.crd);
  double* avals = to_device(A.vals);
  double* xd = to_device(x);
  double* yd = (double*)taco_hip_unified_alloc(sizeof(double) * (size_t)A.m);
  if (!apos || !acrd || !avals || !xd || !yd) return fail("upload");
  taco_tensor_t* tA = csr_desc(A.m, A.n, apos, acrd, avals, (int32_t)A.crd.size());
  taco_tensor_t* tx = vec_desc(A.n, xd);
  taco_tensor_t* ty = vec_desc(A.m, yd);

  // y = A x: results, then operands (tensor.cpp:779-806 packArguments)
  void* spmv_args[3] = {ty, tA, tx};
  if (taco_hip_module_call_packed(spmv, "compute", spmv_args)) return fail("spmv compute");
  if (taco_hip_synchronize()) return fail("synchronize");
  const std::vector<double> y = to_host<double>(yd, (size_t)A.m);
  for (int32_t i = 0; i < A.m; ++i) {
    double s = 0;
    for (int32_t p = A.pos[(size_t)i]; p < A.pos[(size_t)i + 1]; ++p)
      s += A.vals[(size_t)p] * x[(size_t)A.crd[(size_t)p]];
    if (y.size() != (size_t)A.m || y[(size_t)i] != s) {
      std::fprintf(stderr, "FAIL spmv row %d\n", i);
      return 1;
    }
  }
  std::printf("spmv ok: %d rows\n", A.m);

  // C = A B with CSR output: assemble allocates C's pos/crd/vals
  int32_t* bpos = to_device(B.pos);
  int32_t* bcrd = to_device(B.crd);
  double* bvals = to_device(B.vals);
  if (!bpos || !bcrd || !bvals) return fail("upload B");
  taco_tensor_t* tB = csr_desc(B.m, B.n, bpos, bcrd, bvals, (int32_t)B.crd.size());
  taco_tensor_t* tC = csr_desc(A.m, B.n, nullptr, nullptr, nullptr, 0);
  void* gemm_args[3] = {tC, tA, tB};
  if (taco_hip_module_call_packed(spgemm, "assemble", gemm_args)) return fail("spgemm assemble");
  if (taco_hip_module_call_packed(spgemm, "compute", gemm_args)) return fail("spgemm compute");
  if (taco_hip_synchronize()) return fail("synchronize");
  const std::vector<int32_t> cpos = to_host<int32_t>(tC->indices[1][0], (size_t)A.m + 1);
  const int32_t cnnz = cpos.empty() ? -1 : cpos.back();
  const std::vector<int32_t> ccrd = to_host<int32_t>(tC->indices[1][1], (size_t)(cnnz > 0 ? cnnz : 0));
  const std::vector<double> cvals = to_host<double>(tC->vals, (size_t)(cnnz > 0 ? cnnz : 0));
  // taco's workspace order: jA outer, kB inner, first product stored
  int32_t q = 0;
  for (int32_t i = 0; i < A.m; ++i) {
    std::map<int32_t, double> row;
    for (int32_t p = A.pos[(size_t)i]; p < A.pos[(size_t)i + 1]; ++p) {
      const int32_t j = A.crd[(size_t)p];
      for (int32_t r = B.pos[(size_t)j]; r < B.pos[(size_t)j + 1]; ++r) {
        const double v = A.vals[(size_t)p] * B.vals[(size_t)r];
        auto it = row.find(B.crd[(size_t)r]);
        if (it == row.end()) row.emplace(B.crd[(size_t)r], v); else it->second += v;
      }
    }
    if (cpos.empty() || cpos[(size_t)i] != q) { std::fprintf(stderr, "FAIL spgemm pos %d\n", i); return 1; }
    for (const auto& kv : row) {
      if (q >= cnnz || ccrd[(size_t)q] != kv.first || cvals[(size_t)q] != kv.second) {
        std::fprintf(stderr, "FAIL spgemm row %d\n", i);
        return 1;
      }
      ++q;
    }
  }
  if (q != cnnz) { std::fprintf(stderr, "FAIL spgemm nnz\n"); return 1; }
  std::printf("spgemm ok: nnz %d\n", cnnz);

  // outputs of assemble belong to the caller (freed with the runtime's free)
  taco_hip_unified_free(tC->indices[1][0]);
  taco_hip_unified_free(tC->indices[1][1]);
  taco_hip_unified_free(tC->vals);
  for (void* p : {(void*)apos, (void*)acrd, (void*)avals, (void*)xd, (void*)yd, (void*)bpos,
                  (void*)bcrd, (void*)bvals})
    taco_hip_unified_free(p);
  for (taco_tensor_t* t : {tA, tx, ty, tB, tC}) deinit_taco_tensor_t(t);
  taco_hip_module_destroy(spmv);
  taco_hip_module_destroy(spgemm);
  std::printf("run ok\n");
  return 0;
}
