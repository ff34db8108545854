/*
 * Extended-precision oracle ("the FP64 oracle").  TEST INFRASTRUCTURE ONLY.
 *
 * Every output element is summed in long double (x87 80-bit, 64-bit
 * mantissa) with Neumaier's compensated summation, then rounded once to
 * double.  Each product a*b of two doubles is first split exactly into two
 * doubles by the FMA two-product (Ogita, Rump and Oishi, "Accurate sum and
 * dot product", 2005): p = fl(a*b), e = fma(a, b, -p), with p + e = a*b
 * exactly (no underflow occurs for the value ranges used here); both terms
 * enter the compensated sum.  The result is accurate to well below one
 * double ulp of sum_j |a_ij b_j| for all sizes used here, which is what the
 * tolerance gate |got - exact| <= tol * sum_j |a_ij b_j| needs.
 *
 * Compiled without -ffast-math and in ISO C mode, so without floating-point
 * contraction (see oracle/Makefile): fast-math would allow the compiler to
 * cancel the compensation terms algebraically.
 */
#include "taco_cpu.h"

#include <math.h>
#include <stdlib.h>

typedef struct {
  long double s, c;
} nsum_t;

static inline void nadd(nsum_t* a, long double v) {
  long double t = a->s + v;
  if (fabsl(a->s) >= fabsl(v))
    a->c += (a->s - t) + v;
  else
    a->c += (v - t) + a->s;
  a->s = t;
}

/* Add the exact product a*b as its two-product terms p + e. */
static inline void nadd_prod(nsum_t* acc, double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  nadd(acc, p);
  nadd(acc, e);
}

static inline double nresult(const nsum_t* a) { return (double)(a->s + a->c); }

void exact_spmv(int m, const int32_t* pos, const int32_t* crd,
                const double* vals, const double* x, double* y_exact,
                double* cond) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < m; i++) {
    nsum_t acc = {0, 0};
    long double c = 0;
    for (int32_t p = pos[i]; p < pos[i + 1]; p++) {
      double a = vals[p], b = x[crd[p]];
      nadd_prod(&acc, a, b);
      c += fabsl((long double)a * (long double)b);
    }
    if (y_exact) y_exact[i] = nresult(&acc);
    if (cond) cond[i] = (double)c;
  }
}

/* Row by row: each nonzero's B row is read once and added into k running
 * sums.  The terms of every output (i, kk) are added in entry order, as in
 * exact_spmv, so column kk equals exact_spmv(A, B[:, kk]) bit for bit. */
void exact_spmm(int m, int k, const int32_t* pos, const int32_t* crd,
                const double* vals, const double* B, double* C_exact,
                double* cond) {
#pragma omp parallel
  {
    nsum_t* acc = (nsum_t*)malloc(sizeof(nsum_t) * (size_t)(k > 0 ? k : 1));
    long double* cs = (long double*)malloc(sizeof(long double) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < m; i++) {
      for (int32_t kk = 0; kk < k; kk++) {
        acc[kk].s = 0;
        acc[kk].c = 0;
        cs[kk] = 0;
      }
      for (int32_t p = pos[i]; p < pos[i + 1]; p++) {
        const double a = vals[p];
        const double* brow = B + (int64_t)crd[p] * k;
        for (int32_t kk = 0; kk < k; kk++) {
          nadd_prod(&acc[kk], a, brow[kk]);
          cs[kk] += fabsl((long double)a * (long double)brow[kk]);
        }
      }
      for (int32_t kk = 0; kk < k; kk++) {
        if (C_exact) C_exact[(int64_t)i * k + kk] = nresult(&acc[kk]);
        if (cond) cond[(int64_t)i * k + kk] = (double)cs[kk];
      }
    }
    free(acc);
    free(cs);
  }
}

/* Values on a given sorted pattern: for each output (i, k), sum over the
 * pairs (jA, kB) with crd_B[kB] == k.  Each row uses a per-thread dense map
 * from column to output slot. */
void exact_spgemm_values(int m, int kdim, const int32_t* apos,
                         const int32_t* acrd, const double* avals,
                         const int32_t* bpos, const int32_t* bcrd,
                         const double* bvals, const int32_t* cpos,
                         const int32_t* ccrd, double* cvals_exact,
                         double* cond) {
#pragma omp parallel
  {
    int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)kdim);
    nsum_t* acc = (nsum_t*)malloc(sizeof(nsum_t) * (size_t)kdim);
    long double* cs = (long double*)malloc(sizeof(long double) * (size_t)kdim);
    for (int32_t q = 0; q < kdim; q++) slot[q] = -1;
#pragma omp for schedule(static)
    for (int32_t i = 0; i < m; i++) {
      for (int32_t p = cpos[i]; p < cpos[i + 1]; p++) {
        slot[ccrd[p]] = p;
        acc[ccrd[p]].s = 0;
        acc[ccrd[p]].c = 0;
        cs[ccrd[p]] = 0;
      }
      for (int32_t jA = apos[i]; jA < apos[i + 1]; jA++) {
        int32_t j = acrd[jA];
        for (int32_t kB = bpos[j]; kB < bpos[j + 1]; kB++) {
          int32_t k = bcrd[kB];
          if (slot[k] < 0) continue; /* not in pattern: caller's bug */
          nadd_prod(&acc[k], avals[jA], bvals[kB]);
          cs[k] += fabsl((long double)avals[jA] * (long double)bvals[kB]);
        }
      }
      for (int32_t p = cpos[i]; p < cpos[i + 1]; p++) {
        int32_t k = ccrd[p];
        if (cvals_exact) cvals_exact[p] = nresult(&acc[k]);
        if (cond) cond[p] = (double)cs[k];
        slot[k] = -1;
      }
    }
    free(slot);
    free(acc);
    free(cs);
  }
}
