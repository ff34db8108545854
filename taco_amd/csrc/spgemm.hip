// Two-phase SpGEMM on the GPU, C(i,k) = A(i,j) * B(j,k), all CSR, f64.
//
// taco has no GPU SpGEMM (its default GPU schedule bails out on the dense
// workspace, reference transformations.cpp:1578-1583); this follows the
// semantics of its CPU two-phase code (lowerer_impl_imperative.cpp:
// 2197-2271, 2554-2561, 2616-2779; mode_format_compressed.cpp:217-271):
//   assemble  row i's pattern is the sorted set of distinct structural
//             columns {crd_B[kB] : jA in row i of A, kB in row crd_A[jA] of
//             B} (structural: numeric zeros are kept);  pos is the prefix sum
//             of the per-row counts; crd holds each row sorted ascending.
//   compute   C(i,k) accumulates A_ij * B_jk over that row's candidates in
//             encounter order (jA outer, kB inner), first product stored,
//             later ones added — the order of taco's workspace loop.
//
// Row kernels, one 64-lane wave per row, built on the row helpers of
// include/taco_hip_rt.h (shared with modules emitted by CodeGen_HIP):
//   short rows (<= 64 candidates): candidates in lanes, bitonic sort across
//     the wave with shuffles.  The numeric phase sorts 64-bit keys
//     (column << 32 | encounter index), so equal columns come out in
//     encounter order and are summed sequentially in that order.
//   long rows: a per-wave LDS hash set of up to TACO_HIP_ROW_CAP distinct
//     columns, then an in-LDS bitonic sort; the numeric phase accumulates
//     into LDS slots, adding candidates one at a time in encounter order.

#include <climits>

#include "common.h"
#include "device.h"
#include "taco_hip_rt.h"

namespace taco_hip {
namespace {

constexpr int kWavesPerBlock = TACO_HIP_ROW_WAVES;
constexpr int kHashCap = TACO_HIP_ROW_CAP;

// Symbolic pass.  FILL == false: counts[i] = distinct columns of row i.
// FILL == true: writes row i's sorted columns to ccrd[cpos[i] ...].
// status bit 1 is set when a long row exceeds kHashCap distinct columns.
template <bool FILL>
__global__ __launch_bounds__(256) void k_spgemm_symbolic(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const int32_t* __restrict__ bpos, const int32_t* __restrict__ bcrd,
    int32_t* __restrict__ counts, const int32_t* __restrict__ cpos,
    int32_t* __restrict__ ccrd, int32_t* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (i >= m) return;
  const int32_t c = taco_hip_spgemm_row_symbolic<FILL>(
      apos, acrd, bpos, bcrd, (int32_t)i, FILL ? cpos[i] : 0, ccrd, status);
  if (!FILL && __lane_id() == 0) counts[i] = c;
}

// Numeric pass on the assembled pattern.  Rows are strided over a grid the
// device holds at once (taco_hip_row_grid, 4 workgroups per CU; the kernel
// fits 5): one workgroup per four rows left about 4
// waves resident per CU, since a row takes only a few thousand cycles
// (0.56 -> 0.51 ms at 500k rows).  The symbolic passes measured faster with
// one workgroup per four rows, where running waves stream through
// neighbouring rows (0.63 against 0.68-0.73 ms).
__global__ __launch_bounds__(256) void k_spgemm_numeric(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const double* __restrict__ avals, const int32_t* __restrict__ bpos,
    const int32_t* __restrict__ bcrd, const double* __restrict__ bvals,
    const int32_t* __restrict__ cpos, const int32_t* __restrict__ ccrd,
    double* __restrict__ cvals) {
  const int64_t stride = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t i = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; i < m;
       i += stride)
    taco_hip_spgemm_row_numeric(apos, acrd, avals, bpos, bcrd, bvals, (int32_t)i,
                                cpos, ccrd, cvals);
}


// ---- counts -> pos (exclusive prefix sum with pos[m] = total) ----
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int64_t block_exclusive_sum(int64_t v,
                                                       int64_t* s_tmp,
                                                       int64_t* total) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  int64_t inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int64_t u = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += u;
  }
  if (lane == kWave - 1) s_tmp[wave] = inc;
  __syncthreads();
  int64_t wbase = 0, all = 0;
  for (int w = 0; w < kScanBlock / kWave; ++w) {
    if (w < wave) wbase += s_tmp[w];
    all += s_tmp[w];
  }
  __syncthreads();
  *total = all;
  return wbase + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(
    int32_t m, const int32_t* __restrict__ counts, int64_t* __restrict__ tile_sums) {
  __shared__ int64_t s_tmp[kScanBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int64_t v = 0;
  for (int q = 0; q < kScanItems; ++q)
    if (base + q < m) v += counts[base + q];
  int64_t total;
  block_exclusive_sum(v, s_tmp, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tile_sums(
    int32_t ntiles, int64_t* __restrict__ tile_sums) {
  __shared__ int64_t s_tmp[kScanBlock / kWave];
  int64_t carry = 0;
  for (int32_t t0 = 0; t0 < ntiles; t0 += kScanBlock) {
    const int32_t t = t0 + threadIdx.x;
    const int64_t v = t < ntiles ? tile_sums[t] : 0;
    int64_t total;
    const int64_t ex = block_exclusive_sum(v, s_tmp, &total);
    if (t < ntiles) tile_sums[t] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(
    int32_t m, const int32_t* __restrict__ counts,
    const int64_t* __restrict__ tile_offsets, int32_t* __restrict__ pos,
    int32_t* __restrict__ status) {
  __shared__ int64_t s_tmp[kScanBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int64_t v = 0;
  for (int q = 0; q < kScanItems; ++q)
    if (base + q < m) v += counts[base + q];
  int64_t total;
  int64_t run = tile_offsets[blockIdx.x] + block_exclusive_sum(v, s_tmp, &total);
  if (blockIdx.x == 0 && threadIdx.x == 0) pos[0] = 0;
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < m) {
      run += counts[base + q];
      if (run > INT_MAX) atomicOr(status, 2);
      pos[base + q + 1] = (int32_t)run;
    }
  }
}

}  // namespace

// pos[0..m] from counts[0..m); status bit 2 set on int32 overflow.
int device_counts_to_pos(int32_t m, const int32_t* counts, int32_t* pos,
                         int32_t* status, hipStream_t st) {
  const int32_t ntiles = (int32_t)(((int64_t)m + kScanTile - 1) / kScanTile);
  if (m == 0) {
    TH_CHECK(hipMemsetAsync(pos, 0, sizeof(int32_t), st));
    return 0;
  }
  int64_t* sums = (int64_t*)workspace(sizeof(int64_t) * (size_t)ntiles, 3);
  if (!sums) return 1;
  k_scan_tiles<<<ntiles, kScanBlock, 0, st>>>(m, counts, sums);
  k_scan_tile_sums<<<1, kScanBlock, 0, st>>>(ntiles, sums);
  k_scan_apply<<<ntiles, kScanBlock, 0, st>>>(m, counts, sums, pos, status);
  TH_LAUNCH_CHECK("counts_to_pos");
  return 0;
}

}  // namespace taco_hip

using namespace taco_hip;

namespace {

int check_spgemm_operands(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B,
                          CsrView* a, CsrView* b) {
  if (int rc = view_csr(A, "A", a)) return rc;
  if (int rc = view_csr(B, "B", b)) return rc;
  TH_REQUIRE(C && C->order == 2 && C->mode_types[0] == taco_mode_dense &&
                 C->mode_types[1] == taco_mode_sparse,
             "spgemm: C must be CSR");
  TH_REQUIRE(a->n == b->m, "spgemm: A has %d columns, B has %d rows", a->n,
             b->m);
  TH_REQUIRE(C->dimensions[0] == a->m && C->dimensions[1] == b->n,
             "spgemm: C is %dx%d, expected %dx%d", C->dimensions[0],
             C->dimensions[1], a->m, b->n);
  TH_REQUIRE(a->vt == VType::F64 && b->vt == VType::F64 && C->csize == 64,
             "spgemm: f64 values only");
  return 0;
}

}  // namespace

extern "C" int taco_hip_spgemm_assemble(taco_tensor_t* C, taco_tensor_t* A,
                                        taco_tensor_t* B) {
  CsrView a, b;
  if (int rc = check_spgemm_operands(C, A, B, &a, &b)) return rc;
  hipStream_t st = current_stream();
  const int32_t m = a.m;
  int32_t* cpos = (int32_t*)taco_hip_unified_alloc(sizeof(int32_t) * ((size_t)m + 1));
  if (!cpos) return 1;
  // Workspace: counts[m] and a status word.
  char* ws = (char*)workspace(sizeof(int32_t) * ((size_t)m + 64), 2);
  if (!ws) {
    taco_hip_unified_free(cpos);
    return 1;
  }
  int32_t* status = (int32_t*)ws;
  int32_t* counts = (int32_t*)(ws + 256);
  int rc = 0;
  do {
    if (hipMemsetAsync(status, 0, sizeof(int32_t), st) != hipSuccess) {
      rc = fail("spgemm: memset failed");
      break;
    }
    const unsigned grid = (unsigned)(((int64_t)m + kWavesPerBlock - 1) / kWavesPerBlock);
    if (m > 0)
      k_spgemm_symbolic<false><<<grid, 256, 0, st>>>(
          m, a.pos, a.crd, b.pos, b.crd, counts, nullptr, nullptr, status);
    if ((rc = device_counts_to_pos(m, counts, cpos, status, st))) break;
    int32_t host[2] = {0, 0};
    if (hipMemcpyAsync(&host[0], cpos + m, sizeof(int32_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&host[1], status, sizeof(int32_t),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail("spgemm: reading nnz(C) failed");
      break;
    }
    if (host[1] & 1) {
      rc = fail("spgemm: a row of C has more than %d distinct columns, the "
                "limit of the LDS row kernel", kHashCap);
      break;
    }
    if (host[1] & 2) {
      rc = fail("spgemm: nnz(C) exceeds the int32 index range");
      break;
    }
    const int32_t nnz = host[0];
    int32_t* ccrd = (int32_t*)taco_hip_unified_alloc(sizeof(int32_t) * (size_t)nnz);
    double* cvals = (double*)taco_hip_unified_alloc(sizeof(double) * (size_t)nnz);
    if (nnz > 0 && (!ccrd || !cvals)) {
      taco_hip_unified_free(ccrd);
      taco_hip_unified_free(cvals);
      rc = 1;
      break;
    }
    if (m > 0 && nnz > 0)
      k_spgemm_symbolic<true><<<grid, 256, 0, st>>>(
          m, a.pos, a.crd, b.pos, b.crd, nullptr, cpos, ccrd, status);
    if ((rc = finish("taco_hip_spgemm_assemble"))) {
      taco_hip_unified_free(ccrd);
      taco_hip_unified_free(cvals);
      break;
    }
    C->indices[1][0] = (uint8_t*)cpos;
    C->indices[1][1] = (uint8_t*)ccrd;
    C->vals = (uint8_t*)cvals;
    C->vals_size = nnz;
    return 0;
  } while (0);
  taco_hip_unified_free(cpos);
  return rc;
}

extern "C" int taco_hip_spgemm_compute(taco_tensor_t* C, taco_tensor_t* A,
                                       taco_tensor_t* B) {
  CsrView a, b, c;
  if (int rc = check_spgemm_operands(C, A, B, &a, &b)) return rc;
  if (int rc = view_csr(C, "C", &c)) return rc;
  if (a.m == 0 || c.nnz == 0) return finish("taco_hip_spgemm_compute");
  hipStream_t st = current_stream();
  const unsigned grid = (unsigned)taco_hip_row_grid(a.m);
  k_spgemm_numeric<<<grid, 256, 0, st>>>(
      a.m, a.pos, a.crd, (const double*)a.vals, b.pos, b.crd,
      (const double*)b.vals, c.pos, c.crd, (double*)c.vals);
  return finish("taco_hip_spgemm_compute");
}

extern "C" int taco_hip_device_exclusive_scan(const int32_t* counts, int32_t m,
                                              int32_t* pos) {
  hipStream_t st = current_stream();
  int32_t* status = (int32_t*)workspace(sizeof(int32_t), 7);
  if (!status) return 1;
  TH_CHECK(hipMemsetAsync(status, 0, sizeof(int32_t), st));
  if (int rc = device_counts_to_pos(m, counts, pos, status, st)) return rc;
  int32_t host = 0;
  TH_CHECK(hipMemcpyAsync(&host, status, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  TH_REQUIRE(host == 0, "exclusive scan: total exceeds the int32 index range");
  return 0;
}
