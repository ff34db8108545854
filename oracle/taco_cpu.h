/*
 * CPU oracle for the taco sparse hot path (SpMV, SpMM, two-phase SpGEMM).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls these functions; tests/, the smoke check and bench.py's CPU baseline
 * use them as the checker.  The product path fails loudly without its HIP
 * library and never falls back to this code.
 *
 * Two groups of functions:
 *   oracle_*      plain-C/OpenMP restatement of the C code that taco's
 *                 CodeGen_C emits for these statements (loop nests, summation
 *                 order and result initialisation kept as emitted).  Compiled
 *                 like taco's JIT: -O3 -ffast-math -std=c99 -fopenmp
 *                 (reference src/codegen/module.cpp:126-139).
 *   exact_*       extended-precision oracle: long double with Neumaier
 *                 compensation per output element, plus the condition sums
 *                 sum_j |a_ij * x_j| that scale the tolerance gate.  Compiled
 *                 without -ffast-math (oracle_exact.c).
 *
 * Index arrays are int32 CSR: pos[m+1], crd[nnz], as taco's compressed mode
 * stores them (reference src/lower/mode_format_compressed.cpp:217-271).
 */
#ifndef TACO_ORACLE_CPU_H
#define TACO_ORACLE_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* OpenMP thread count used by the oracle_* loops (taco_set_num_threads). */
void oracle_set_num_threads(int n);
int oracle_get_num_threads(void);

/* y = A x, A CSR m x n. */
void oracle_spmv_f64(int m, int n, const int32_t* pos, const int32_t* crd,
                     const double* vals, const double* x, double* y);
void oracle_spmv_f32(int m, int n, const int32_t* pos, const int32_t* crd,
                     const float* vals, const float* x, float* y);

/* C = A B, A CSR m x n, B dense n x k row-major, C dense m x k row-major. */
void oracle_spmm_f64(int m, int n, int k, const int32_t* pos,
                     const int32_t* crd, const double* vals, const double* B,
                     double* C);
void oracle_spmm_f32(int m, int n, int k, const int32_t* pos,
                     const int32_t* crd, const float* vals, const float* B,
                     float* C);

/* Two-phase SpGEMM C = A B, all CSR.  assemble: fills cpos[m+1] and returns
 * nnz(C) when ccrd == NULL (query pass), else fills ccrd[nnz(C)] with the
 * sorted structural column indices.  compute: fills cvals on that pattern.
 * `kdim` is the column count of B (and C).  Returns nnz(C), or -1 on error. */
int64_t oracle_spgemm_assemble(int m, int kdim, const int32_t* apos,
                               const int32_t* acrd, const int32_t* bpos,
                               const int32_t* bcrd, int32_t* cpos,
                               int32_t* ccrd);
void oracle_spgemm_compute_f64(int m, int kdim, const int32_t* apos,
                               const int32_t* acrd, const double* avals,
                               const int32_t* bpos, const int32_t* bcrd,
                               const double* bvals, const int32_t* cpos,
                               const int32_t* ccrd, double* cvals);

/* ---- extended-precision oracle (oracle_exact.c) ---- */

/* y_exact[i] = sum_j a_ij x_j (long double, compensated), rounded to double;
 * cond[i] = sum_j |a_ij x_j|. Either output may be NULL. */
void exact_spmv(int m, const int32_t* pos, const int32_t* crd,
                const double* vals, const double* x, double* y_exact,
                double* cond);
/* Same for C = A B (row-major dense B, C). */
void exact_spmm(int m, int k, const int32_t* pos, const int32_t* crd,
                const double* vals, const double* B, double* C_exact,
                double* cond);
/* Values of C = A B on a given pattern (cpos/ccrd), plus condition sums. */
void exact_spgemm_values(int m, int kdim, const int32_t* apos,
                         const int32_t* acrd, const double* avals,
                         const int32_t* bpos, const int32_t* bcrd,
                         const double* bvals, const int32_t* cpos,
                         const int32_t* ccrd, double* cvals_exact,
                         double* cond);

#ifdef __cplusplus
}
#endif
#endif
