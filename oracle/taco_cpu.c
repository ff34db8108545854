/*
 * CPU oracle: C/OpenMP restatement of the code taco's CodeGen_C emits for the
 * sparse hot path.  TEST INFRASTRUCTURE ONLY (see taco_cpu.h).
 *
 * Each function keeps the emitted loop nest, the result initialisation and
 * the sequential per-row summation order of the generated C:
 *   - parallel outer loop `#pragma omp parallel for schedule(runtime)`
 *     (reference src/codegen/codegen_c.cpp:393-417, LoopKind::Runtime from
 *     src/lower/lowerer_impl_imperative.cpp:1223-1226), with the runtime
 *     schedule set to static, chunk 0 (src/tensor.cpp:1387-1388,
 *     src/codegen/module.cpp:182-215);
 *   - row loop over the compressed level `for (jA = pos[i]; jA < pos[i+1];
 *     jA++)` (src/lower/lowerer_impl_imperative.cpp:1304-1416);
 *   - reduction results zero-filled before the parallel loop
 *     (src/lower/lowerer_impl_imperative.cpp:3348-3364);
 *   - two-phase SpGEMM with a per-thread dense workspace, an index list
 *     sorted with qsort, and the pos-cursor / shift-back protocol of the
 *     compressed mode (src/lower/lowerer_impl_imperative.cpp:2197-2271,
 *     2554-2561, 2616-2779; src/lower/mode_format_compressed.cpp:217-271).
 */
#include "taco_cpu.h"

#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

static int g_threads = 1; /* taco's default thread count (src/tensor.cpp:1389) */

void oracle_set_num_threads(int n) {
  if (n > 0) g_threads = n;
}
int oracle_get_num_threads(void) { return g_threads; }

static void enter_parallel(void) {
  omp_set_schedule(omp_sched_static, 0);
  omp_set_num_threads(g_threads);
}

/* y(i) = A(i,j) * x(j): generated compute() of the default schedule. */
#define SPMV_BODY(T)                                                    \
  enter_parallel();                                                     \
  for (int32_t py = 0; py < m; py++) y[py] = 0;                         \
  _Pragma("omp parallel for schedule(runtime)")                         \
  for (int32_t i = 0; i < m; i++) {                                     \
    for (int32_t jA = pos[i]; jA < pos[i + 1]; jA++) {                  \
      int32_t j = crd[jA];                                              \
      y[i] = y[i] + vals[jA] * x[j];                                    \
    }                                                                   \
  }

void oracle_spmv_f64(int m, int n, const int32_t* pos, const int32_t* crd,
                     const double* vals, const double* x, double* y) {
  (void)n;
  SPMV_BODY(double)
}

void oracle_spmv_f32(int m, int n, const int32_t* pos, const int32_t* crd,
                     const float* vals, const float* x, float* y) {
  (void)n;
  SPMV_BODY(float)
}

/* C(i,k) = A(i,j) * B(j,k), dense row-major C and B: loops i, jA, k. */
#define SPMM_BODY(T)                                                    \
  enter_parallel();                                                     \
  int64_t total = (int64_t)m * (int64_t)k;                              \
  for (int64_t pC = 0; pC < total; pC++) C[pC] = 0;                     \
  _Pragma("omp parallel for schedule(runtime)")                         \
  for (int32_t i = 0; i < m; i++) {                                     \
    for (int32_t jA = pos[i]; jA < pos[i + 1]; jA++) {                  \
      int32_t j = crd[jA];                                              \
      for (int32_t kk = 0; kk < k; kk++) {                              \
        int64_t kC = (int64_t)i * k + kk;                               \
        int64_t kB = (int64_t)j * k + kk;                               \
        C[kC] = C[kC] + vals[jA] * B[kB];                               \
      }                                                                 \
    }                                                                   \
  }

void oracle_spmm_f64(int m, int n, int k, const int32_t* pos,
                     const int32_t* crd, const double* vals, const double* B,
                     double* C) {
  (void)n;
  SPMM_BODY(double)
}

void oracle_spmm_f32(int m, int n, int k, const int32_t* pos,
                     const int32_t* crd, const float* vals, const float* B,
                     float* C) {
  (void)n;
  SPMM_BODY(float)
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Assemble phase.  Query pass (ccrd == NULL): structural nnz per row via the
 * Bool "already set" dedup, then the serial prefix sum into cpos.  Insert pass
 * (ccrd != NULL, cpos already holding the prefix sums): scatter, qsort the
 * index list, append through the pos cursor, then shift pos back. */
int64_t oracle_spgemm_assemble(int m, int kdim, const int32_t* apos,
                               const int32_t* acrd, const int32_t* bpos,
                               const int32_t* bcrd, int32_t* cpos,
                               int32_t* ccrd) {
  enter_parallel();
  int nt = g_threads;
  int32_t* list_all = (int32_t*)malloc(sizeof(int32_t) * (size_t)kdim * nt);
  bool* set_all = (bool*)calloc((size_t)kdim * nt, sizeof(bool));
  if (!list_all || !set_all) {
    free(list_all);
    free(set_all);
    return -1;
  }
  if (ccrd == NULL) {
    int32_t* nnz = (int32_t*)calloc((size_t)(m > 0 ? m : 1), sizeof(int32_t));
#pragma omp parallel for schedule(runtime)
    for (int32_t qi = 0; qi < m; qi++) {
      int32_t size = 0;
      int32_t* list = list_all + (size_t)kdim * omp_get_thread_num();
      bool* set = set_all + (size_t)kdim * omp_get_thread_num();
      for (int32_t qjA = apos[qi]; qjA < apos[qi + 1]; qjA++) {
        int32_t qj = acrd[qjA];
        for (int32_t qkB = bpos[qj]; qkB < bpos[qj + 1]; qkB++) {
          int32_t qk = bcrd[qkB];
          if (!set[qk]) {
            list[size++] = qk;
            set[qk] = 1;
          }
        }
      }
      for (int32_t l = 0; l < size; l++) {
        nnz[qi] = nnz[qi] + 1;
        set[list[l]] = 0;
      }
    }
    int64_t acc = 0;
    cpos[0] = 0;
    for (int32_t i = 0; i < m; i++) {
      acc += nnz[i];
      cpos[i + 1] = (int32_t)acc;
    }
    free(nnz);
    free(list_all);
    free(set_all);
    return acc;
  }
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m + 1));
  memcpy(cur, cpos, sizeof(int32_t) * (size_t)(m + 1));
#pragma omp parallel for schedule(runtime)
  for (int32_t i = 0; i < m; i++) {
    int32_t size = 0;
    int32_t* list = list_all + (size_t)kdim * omp_get_thread_num();
    bool* set = set_all + (size_t)kdim * omp_get_thread_num();
    for (int32_t jA = apos[i]; jA < apos[i + 1]; jA++) {
      int32_t j = acrd[jA];
      for (int32_t kB = bpos[j]; kB < bpos[j + 1]; kB++) {
        int32_t k = bcrd[kB];
        if (!set[k]) {
          list[size++] = k;
          set[k] = 1;
        }
      }
    }
    qsort(list, (size_t)size, sizeof(int32_t), cmp_i32);
    for (int32_t l = 0; l < size; l++) {
      int32_t k = list[l];
      int32_t p = cur[i];
      cur[i] = cur[i] + 1;
      ccrd[p] = k;
      set[k] = 0;
    }
  }
  int64_t total = cpos[m];
  free(cur);
  free(list_all);
  free(set_all);
  return total;
}

/* Compute phase: first product stored, later ones accumulated, in (jA, kB)
 * encounter order; values emitted in sorted column order. */
void oracle_spgemm_compute_f64(int m, int kdim, const int32_t* apos,
                               const int32_t* acrd, const double* avals,
                               const int32_t* bpos, const int32_t* bcrd,
                               const double* bvals, const int32_t* cpos,
                               const int32_t* ccrd, double* cvals) {
  (void)ccrd;
  enter_parallel();
  int nt = g_threads;
  double* w_all = (double*)malloc(sizeof(double) * (size_t)kdim * nt);
  int32_t* list_all = (int32_t*)malloc(sizeof(int32_t) * (size_t)kdim * nt);
  bool* set_all = (bool*)calloc((size_t)kdim * nt, sizeof(bool));
#pragma omp parallel for schedule(runtime)
  for (int32_t i = 0; i < m; i++) {
    int32_t size = 0;
    double* w = w_all + (size_t)kdim * omp_get_thread_num();
    int32_t* list = list_all + (size_t)kdim * omp_get_thread_num();
    bool* set = set_all + (size_t)kdim * omp_get_thread_num();
    for (int32_t jA = apos[i]; jA < apos[i + 1]; jA++) {
      int32_t j = acrd[jA];
      for (int32_t kB = bpos[j]; kB < bpos[j + 1]; kB++) {
        int32_t k = bcrd[kB];
        if (!set[k]) {
          w[k] = avals[jA] * bvals[kB];
          list[size++] = k;
          set[k] = 1;
        } else {
          w[k] = w[k] + avals[jA] * bvals[kB];
        }
      }
    }
    qsort(list, (size_t)size, sizeof(int32_t), cmp_i32);
    int32_t p = cpos[i];
    for (int32_t l = 0; l < size; l++) {
      int32_t k = list[l];
      cvals[p++] = w[k];
      set[k] = 0;
    }
  }
  free(w_all);
  free(list_all);
  free(set_all);
}
