// bf16 SpMM on the matrix cores: C(i,k) = A(i,j) * B(j,k) with bf16 A
// values and B, f32 accumulation and f32 C (no reference counterpart; taco
// has no bf16 type, reference include/taco/type.h:19-36).
//
// Tile formulation for v_mfma_f32_16x16x16_bf16 (wave64):
//   M = 16 consecutive rows of A (a row group), K = 16 consecutive stored
//   entries of that group (the group's entries are contiguous in CSR),
//   N = 16 columns of B/C.  A_tile[m][k] = vals[e_k] if entry e_k belongs to
//   row m, else 0; B_tile[k][n] = B[crd[e_k]][n].  Summing tiles over the
//   group's entry chunks gives exactly C's 16 rows.
// Operand layout (lane l): A[m = l%16][k = 4*(l/16)+q], B[k = 4*(l/16)+q]
// [n = l%16], D[m = 4*(l/16)+r][n = l%16], q, r = 0..3.
// Column permutation: within a 64-column slab, MFMA column n of n-block nb
// is C column 4n + nb, so one lane's four B values for an entry are 8
// contiguous bytes and its four results for a row are 16 contiguous bytes.
//
// Load balance: a group with more than kTileMax entries is split into parts
// of kTileMax entries; parts of a split group add into C with f32 atomics
// (their rows are zeroed first), unsplit groups store directly.

#include "common.h"
#include "device.h"

namespace taco_hip {

int device_counts_to_pos(int32_t m, const int32_t* counts, int32_t* pos,
                         int32_t* status, hipStream_t st);

namespace {

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

constexpr int kTileMax = 1024;  // entries per MFMA task

__global__ __launch_bounds__(256) void k_mfma_parts(
    int32_t m, int32_t K, const int32_t* __restrict__ pos,
    int32_t* __restrict__ parts, float* __restrict__ C, int32_t* __restrict__ status) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // The prefix sum that follows flags overflow in *status: cleared here, in
  // a kernel rather than a memset, so the call can be captured in a graph.
  if (g == 0) *status = 0;
  const int64_t ngroups = ((int64_t)m + 15) / 16;
  if (g >= ngroups) return;
  const int64_t r0 = g * 16;
  const int64_t r1 = r0 + 16 < m ? r0 + 16 : m;
  const int64_t T = (int64_t)pos[r1] - pos[r0];
  const int32_t np = T <= kTileMax ? 1 : (int32_t)((T + kTileMax - 1) / kTileMax);
  parts[g] = np;
  if (np > 1) {
    float4* dst = (float4*)(C + r0 * K);
    const int64_t n4 = (r1 - r0) * K / 4;
    for (int64_t q = 0; q < n4; ++q) dst[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Task -> row group map: a task's wave reads its group with one load instead
// of a binary search over part_pos (17 dependent loads at two million rows,
// about as long as a typical task's own gathers).  Group g owns the tasks
// part_pos[g] .. part_pos[g+1]-1; every group has at least one.
__global__ __launch_bounds__(256) void k_mfma_task_groups(
    int32_t ngroups, int64_t max_tasks, const int32_t* __restrict__ part_pos,
    int32_t* __restrict__ task_group) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t s = part_pos[g];
  const int64_t e = (int64_t)part_pos[g + 1] < max_tasks ? part_pos[g + 1] : max_tasks;
  for (int64_t t = s; t < e; ++t) task_group[t] = (int32_t)g;
}

__device__ __forceinline__ short bf16_or_zero(const uint16_t* p, bool ok) {
  return ok ? (short)p[0] : (short)0;
}

// 7 waves per SIMD: measured 0.89 ms on the 2M-row power-law matrix (K=64)
// against 0.92 ms at 6 and 1.28 ms at 8; the compiler's own choice (5) took
// 1.01 ms.  Entry indices are 32-bit (CSR positions are int32) and, when B
// spans less than 4 GiB (kOff32), so are B's byte offsets: global loads then
// take a scalar base plus one offset register, which keeps the kernel within
// 7 waves' registers without spilling.
template <bool kOff32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_spmm_mfma_bf16(
    int32_t m, int32_t K, int32_t ngroups, const int32_t* __restrict__ part_pos,
    const int32_t* __restrict__ task_group,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const uint16_t* __restrict__ vals, const uint16_t* __restrict__ B,
    float* __restrict__ C, bool coarse) {
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int32_t t = (int32_t)blockIdx.x * 4 + wave;
  const int32_t ntasks = part_pos[ngroups];
  if (t >= ntasks) return;
  const int32_t g = task_group[t];
  const int32_t part = t - part_pos[g];
  const int32_t nparts = part_pos[g + 1] - part_pos[g];
  const int32_t r0 = g * 16;
  const int rows = m - r0 < 16 ? m - r0 : 16;
  const int mrow = lane & 15;
  const int kq = lane >> 4;
  const int32_t ps = mrow < rows ? pos[r0 + mrow] : 0;
  const int32_t pe = mrow < rows ? pos[r0 + mrow + 1] : 0;
  const int32_t gs = pos[r0], ge = pos[r0 + rows];
  const int32_t es = gs + part * kTileMax;
  const int32_t ee = ge - es < kTileMax ? ge : es + kTileMax;
  const char* Bb = (const char*)B;

  for (int32_t c0 = 0; c0 < K; c0 += 64) {
    const int32_t col = c0 + 4 * (lane & 15);
    const bool col_ok = col < K;
    v4f32 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[nb] = (v4f32){0.f, 0.f, 0.f, 0.f};
    // A's stream runs one 16-entry step ahead: the next step's columns and
    // values load while this step's B rows are in flight, so a step waits
    // for one dependent load (B) instead of two (crd, then B).
    int32_t jn[4];
    v4i16 an;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t idx = es + 4 * kq + q;
      const bool valid = idx < ee;
      jn[q] = valid ? stream_load(crd + idx) : 0;
      an[q] = bf16_or_zero(vals + idx, valid && idx >= ps && idx < pe);
    }
    for (int32_t e = es; e < ee; e += 16) {
      const v4i16 a = an;
      int32_t j[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) j[q] = jn[q];
      uint2 bq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int32_t idx = e + 4 * kq + q;
        const bool ok = idx < ee && col_ok;
        if constexpr (kOff32) {
          const uint32_t off = ((uint32_t)j[q] * (uint32_t)K + (uint32_t)col) * 2u;
          bq[q] = ok ? *(const uint2*)(Bb + off) : make_uint2(0u, 0u);
        } else {
          bq[q] = ok ? *(const uint2*)(B + (int64_t)j[q] * K + col) : make_uint2(0u, 0u);
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int32_t idx = e + 16 + 4 * kq + q;
        const bool valid = idx < ee;
        jn[q] = valid ? stream_load(crd + idx) : 0;
        an[q] = bf16_or_zero(vals + idx, valid && idx >= ps && idx < pe);
      }
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        v4i16 b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t w = nb < 2 ? bq[q].x : bq[q].y;
          b[q] = (short)((nb & 1) ? (w >> 16) : (w & 0xffffu));
        }
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nb], 0, 0, 0);
      }
    }
    if (!col_ok) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int lr = 4 * kq + r;
      if (lr >= rows) continue;
      float* dst = C + ((int64_t)r0 + lr) * K + col;
      if (nparts == 1) {
        // Non-temporal: C is written once; keeping it out of the caches
        // leaves them to the gathered B rows.
        const v4f32 o = {acc[0][r], acc[1][r], acc[2][r], acc[3][r]};
        __builtin_nontemporal_store(o, (v4f32*)dst);
      } else {
        atomic_add(dst + 0, acc[0][r], coarse);
        atomic_add(dst + 1, acc[1][r], coarse);
        atomic_add(dst + 2, acc[2][r], coarse);
        atomic_add(dst + 3, acc[3][r], coarse);
      }
    }
  }
}

}  // namespace

int spmm_mfma_bf16(const CsrView& a, int32_t K, const void* B, float* C,
                   const taco_hip_schedule& s) {
  (void)s;
  if (a.m == 0 || K == 0) return 0;
  TH_REQUIRE(K % 4 == 0, "spmm: the MFMA path needs K divisible by 4 (K=%d)",
             K);
  TH_REQUIRE(((uintptr_t)B % 8) == 0 && ((uintptr_t)C % 16) == 0,
             "spmm: the MFMA path needs 8-byte aligned B and 16-byte aligned C");
  hipStream_t st = current_stream();
  const int32_t ngroups = (int32_t)(((int64_t)a.m + 15) / 16);
  // Every group has ceil(T/kTileMax) parts, at least one, so the tasks
  // number at most ngroups + ceil(nnz/kTileMax).
  const int64_t max_tasks = (int64_t)ngroups + (a.nnz + kTileMax - 1) / kTileMax;
  char* ws = (char*)workspace(
      sizeof(int32_t) * (2 * (size_t)ngroups + 2 + (size_t)max_tasks) + 512, 1);
  if (!ws) return 1;
  int32_t* status = (int32_t*)ws;
  int32_t* parts = (int32_t*)(ws + 256);
  int32_t* part_pos = parts + ngroups + 1;
  int32_t* task_group = part_pos + ngroups + 1;
  k_mfma_parts<<<(ngroups + 255) / 256, 256, 0, st>>>(a.m, K, a.pos, parts, C, status);
  if (int rc = device_counts_to_pos(ngroups, parts, part_pos, status, st))
    return rc;
  k_mfma_task_groups<<<(ngroups + 255) / 256, 256, 0, st>>>(ngroups, max_tasks,
                                                          part_pos, task_group);
  const unsigned grid = (unsigned)((max_tasks + 3) / 4);
  const bool off32 = (int64_t)a.n * K * 2 <= (int64_t)UINT32_MAX;
  (off32 ? k_spmm_mfma_bf16<true> : k_spmm_mfma_bf16<false>)<<<grid, 256, 0, st>>>(
      a.m, K, ngroups, part_pos, task_group, a.pos, a.crd, (const uint16_t*)a.vals,
      (const uint16_t*)B, C, coarse_grained(C));
  TH_LAUNCH_CHECK("k_spmm_mfma_bf16");
  return 0;
}

}  // namespace taco_hip
