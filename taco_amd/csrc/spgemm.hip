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
// Assemble sorts each row once:
//   1. per-row candidate counts (sum of the B row lengths A's row names) and
//      their prefix sum give every row a scratch region that holds all of
//      its candidates;
//   2. one symbolic pass writes each row's sorted distinct columns into its
//      region and its distinct count;
//   3. the prefix sum of the distinct counts is C's pos; a compaction
//      copies each region's sorted columns to crd.
// Rows with more distinct columns than the LDS row set holds are listed by
// pass 2 and sorted in global memory: their candidates are written out in
// encounter order, sorted per row by a segmented radix sort (rocPRIM) and
// deduplicated into their regions.  When the candidate total exceeds the
// int32 range the two-pass protocol of taco (count, then fill, each
// building the row set) is used instead, with the LDS row-set limit.
//
// Row kernels, one 64-lane wave per row, built on the row helpers of
// include/taco_hip_rt.h (shared with modules emitted by CodeGen_HIP):
//   short rows (<= 64 candidates): candidates in lanes, bitonic sort across
//     the wave by register moves (DPP, lane-half swaps).  The numeric phase
//     sorts (column, encounter index) keys, in 32 bits when the columns
//     allow, so equal columns come out in encounter order and are summed
//     sequentially in that order.  The hand-written short-row kernels take
//     three rows a wave at a time, staging their loads together, and list
//     longer rows for a second pass.
//   long rows: a per-wave LDS hash set of up to TACO_HIP_ROW_CAP distinct
//     columns, then an in-LDS bitonic sort; the numeric phase accumulates
//     into LDS slots, adding candidates one at a time in encounter order
//     (rows of more than TACO_HIP_ROW_CAP columns accumulate in C's values
//     in global memory, in the same order).

#include <climits>
#include <vector>

#include <rocprim/device/device_segmented_radix_sort.hpp>

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

// Rows per wave taken together by the short-row kernels: their dependent
// loads (A's row bounds, A's entries, B's row bounds, B's entries) are
// issued stage by stage for all of them, so that many rows' loads are in
// flight at once.
constexpr int kRowsTogether = 3;

// A wave's kRowsTogether short rows, rows base + q * stride (base and
// stride wave-uniform), staged: each stage issues its loads for every row
// before the next stage waits for them.  Every load reads a valid address
// (indices clamped, A and B each holding at least one entry), so no load
// sits under a branch, and lanes past a row's end mask what they loaded.
// Lane t < la[q] holds A's entry t of row q; a row of more than 64
// candidates (or 64 A entries) has total[q] = 65 and no candidates in
// lanes; otherwise lane c < total[q] holds candidate c (encounter order):
// its column col[q] and, with VALS, its product prod[q].
template <int R, bool VALS>
struct ShortRows {
  bool real[R];
  int32_t a0[R], la[R], total[R], col[R];
  double prod[R];

  __device__ void load(int64_t base, int64_t stride, int32_t m, const int32_t* apos,
                       const int32_t* acrd, const double* avals, int32_t nnz_a,
                       const int32_t* bpos, const int32_t* bcrd, const double* bvals,
                       int32_t nnz_b) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int64_t i = base + q * stride;
      real[q] = i < m;
      const int64_t ic = real[q] ? i : m;
      a0[q] = apos[ic];
      la[q] = real[q] ? apos[ic + 1] - a0[q] : 0;
    }
    int32_t j[R];
    double av[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int32_t e = a0[q] + (lane < la[q] ? lane : la[q] - 1);
      e = e < 0 ? 0 : (e < nnz_a ? e : nnz_a - 1);
      j[q] = acrd[e];
      av[q] = VALS ? avals[e] : 0.0;
    }
    int32_t b0[R], b1[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (lane >= la[q] || la[q] > 64) j[q] = 0;
      b0[q] = bpos[j[q]];
      b1[q] = bpos[j[q] + 1];
    }
    int32_t kB[R];
    double a[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool mine = lane < la[q] && la[q] <= 64;
      const int32_t len = (b1[q] - b0[q]) & (mine ? -1 : 0);
      // B row lengths clamped to 65: the sums stay small and a total above
      // 64 still reads as above 64.
      const int32_t lc = len < 65 ? len : 65;
      const int32_t inc = taco_hip_wave_scan_add_i32(lc);
      total[q] = la[q] > 64 ? 65 : __builtin_amdgcn_readlane(inc, 63);
      // Entry t holding candidate `lane`: the entries s < la - 1 whose
      // inclusive sum is <= lane.
      const int32_t tops = total[q] <= 64 ? la[q] - 1 : 0;
      int t = 0;
      for (int s2 = 0; s2 < tops; ++s2) t += __builtin_amdgcn_readlane(inc, s2) <= lane ? 1 : 0;
      const int32_t excl = __shfl(inc - lc, t, 64);
      const int32_t bt = __shfl(b0[q], t, 64);
      a[q] = VALS ? __shfl(av[q], t, 64) : 0.0;
      int32_t k = bt + (lane - excl);
      kB[q] = k < 0 ? 0 : (k < nnz_b ? k : nnz_b - 1);
    }
    int32_t c[R];
    double bv[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      c[q] = bcrd[kB[q]];
      bv[q] = VALS ? bvals[kB[q]] : 0.0;
    }
    // Masks applied arithmetically, so that every lane uses what it loaded
    // and the compiler keeps the loads where they are (a select would let
    // it move them under a branch and wait there).  Columns are
    // non-negative, so c | INT_MAX = INT_MAX; x * 1.0 is exact.
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool cand = total[q] <= 64 && lane < total[q];
      col[q] = c[q] | (cand ? 0 : INT_MAX);
      prod[q] = VALS ? a[q] * bv[q] * (cand ? 1.0 : 0.0) : 0.0;
    }
  }
};

// Numeric pass over the rows of at most 64 candidates, whose products stay
// in lanes: this kernel uses no LDS, so twice as many waves stay resident as
// with the general row kernel (whose long-row accumulators take 32 KB of
// LDS per workgroup), and each wave stages the loads of kRowsTogether rows
// at once.  Longer rows are listed (long_rows, *nlong) for
// k_spgemm_numeric_listed, the general row kernel over the list.
__global__ __launch_bounds__(256) void k_spgemm_numeric_short(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const double* __restrict__ avals, int32_t nnz_a, const int32_t* __restrict__ bpos,
    const int32_t* __restrict__ bcrd, const double* __restrict__ bvals, int32_t nnz_b,
    const int32_t* __restrict__ cpos, double* __restrict__ cvals,
    int32_t* __restrict__ long_rows, int32_t* __restrict__ nlong) {
  const int lane = (int)__lane_id();
  const int64_t stride = (int64_t)gridDim.x * kWavesPerBlock;
  const int64_t first = (int64_t)blockIdx.x * kWavesPerBlock +
                        __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  if (nnz_a == 0 || nnz_b == 0) return;  // C has no entries
  const TacoHipSortMasks sm;
  for (int64_t base = first; base < m; base += kRowsTogether * stride) {
    ShortRows<kRowsTogether, true> r;
    r.load(base, stride, m, apos, acrd, avals, nnz_a, bpos, bcrd, bvals, nnz_b);
#pragma unroll
    for (int q = 0; q < kRowsTogether; ++q) {
      if (!r.real[q]) continue;
      const int64_t i = base + q * stride;
      if (r.total[q] > 64) {
        if (lane == 0) long_rows[atomicAdd(nlong, 1)] = (int32_t)i;
        continue;
      }
      const int32_t c0 = cpos[i];
      taco_hip_spgemm_row_sum_store(r.total[q], r.col[q], r.prod[q], c0, cpos[i + 1] - c0,
                                    cvals, &sm);
    }
  }
}

// Numeric pass over the rows listed by k_spgemm_numeric_short.
__global__ __launch_bounds__(256) void k_spgemm_numeric_listed(
    const int32_t* __restrict__ long_rows, const int32_t* __restrict__ nlong,
    const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const double* __restrict__ avals, const int32_t* __restrict__ bpos,
    const int32_t* __restrict__ bcrd, const double* __restrict__ bvals,
    const int32_t* __restrict__ cpos, const int32_t* __restrict__ ccrd,
    double* __restrict__ cvals) {
  const int32_t count = *nlong;
  const int64_t stride = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t r = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; r < count;
       r += stride)
    taco_hip_spgemm_row_numeric(apos, acrd, avals, bpos, bcrd, bvals, long_rows[r], cpos,
                                ccrd, cvals);
}

__global__ void k_spgemm_zero_count(int32_t* __restrict__ count) {
  if (threadIdx.x == 0) *count = 0;
}

// out[0..4] = status[0..2], *candoff_m, *cpos_m: what the host reads back
// after the row-set pass, gathered for one copy.
__global__ void k_spgemm_gather_result(const int32_t* __restrict__ status,
                                       const int32_t* __restrict__ candoff_m,
                                       const int32_t* __restrict__ cpos_m,
                                       int32_t* __restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = status[0];
    out[1] = status[1];
    out[2] = status[2];
    out[3] = *candoff_m;
    out[4] = *cpos_m;
  }
}

// ---- one-sort assemble ----

// cand[i] = candidates of row i (saturated at INT_MAX; the prefix sum then
// reports the overflow).  kCandLanes lanes per row, each striding the row's
// A entries, so a wave covers 64 / kCandLanes rows and a row's loads are in
// flight together.
constexpr int kCandLanes = 8;
// It also clears the assemble's status words (status[0..15]), which no
// kernel reads before the prefix sum that follows it.
__global__ __launch_bounds__(256) void k_spgemm_row_cands(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const int32_t* __restrict__ bpos, int32_t* __restrict__ cand,
    int32_t* __restrict__ status) {
  if (blockIdx.x == 0 && threadIdx.x < 16) status[threadIdx.x] = 0;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kCandLanes;
  const int sub = (int)(t % kCandLanes);
  int64_t n = 0;
  if (i < m)
    for (int32_t p = apos[i] + sub; p < apos[i + 1]; p += kCandLanes) {
      const int32_t j = acrd[p];
      n += bpos[j + 1] - bpos[j];
    }
#pragma unroll
  for (int off = kCandLanes / 2; off > 0; off >>= 1) n += __shfl_xor(n, off, kCandLanes);
  if (i < m && sub == 0) cand[i] = n > INT_MAX ? INT_MAX : (int32_t)n;
}

// Row i's sorted distinct columns into dst[0..count).  Returns the count,
// or -1 (nothing written) when the row has more distinct columns than the
// LDS row set holds.
__device__ inline int32_t row_sorted_set(const int32_t* apos, const int32_t* acrd,
                                         const int32_t* bpos, const int32_t* bcrd,
                                         int32_t i, int32_t* dst) {
  __shared__ TacoHipRowSet s_sets[kWavesPerBlock];
  const int lane = (int)__lane_id();
  int32_t col = INT_MAX, jA = 0, kB = 0;
  int64_t n = taco_hip_row_lanes(apos, acrd, bpos, bcrd, i, &col, &jA, &kB);
  if (n < 0) {
    n = taco_hip_row_candidates(apos, acrd, bpos, i);
    if (n <= 64 && lane < n)
      taco_hip_row_candidate(apos, acrd, bpos, bcrd, i, lane, &col, &jA, &kB);
  }
  if (n <= 64) {
    const int32_t s = (int32_t)taco_hip_wave_sort_u32((uint32_t)col);
    const int32_t prev = __shfl_up(s, 1, 64);
    const bool head = s != INT_MAX && (lane == 0 || prev != s);
    const uint64_t heads = __ballot(head);
    if (head) dst[__popcll(heads & ((1ull << lane) - 1))] = s;
    return __popcll(heads);
  }
  TacoHipRowSet* w = &s_sets[(threadIdx.x / kWave) % kWavesPerBlock];
  const int32_t cnt = taco_hip_row_hash(w, apos, acrd, bpos, bcrd, i);
  if (cnt < 0) return -1;
  taco_hip_row_hash_sort(w);
  for (int q = lane; q < cnt; q += kWave) dst[q] = w->keys[q];
  return cnt;
}


// Pass 2 over the rows of at most 64 candidates (no LDS, so more waves stay
// resident): each row's sorted distinct columns go to its scratch region
// and its distinct count to counts.  Longer rows are listed (mid_rows,
// status[3]) for k_spgemm_symbolic_listed.  status[0] bit 4: the scratch
// holds fewer entries than the candidate total, or the candidate offsets
// overflowed (nothing done).
__global__ __launch_bounds__(256) void k_spgemm_symbolic_short(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    int32_t nnz_a, const int32_t* __restrict__ bpos, const int32_t* __restrict__ bcrd,
    int32_t nnz_b, const int32_t* __restrict__ candoff, int64_t capacity,
    int32_t* __restrict__ scratch, int32_t* __restrict__ counts,
    int32_t* __restrict__ mid_rows, int32_t* __restrict__ status) {
  const int lane = (int)__lane_id();
  const int64_t first = (int64_t)blockIdx.x * kWavesPerBlock +
                        __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  // The candidate offsets overflowed int32 (status[2] bit 2: the caller
  // falls back to the two-pass form), or the scratch is too small: write
  // nothing.
  if ((status[2] & 2) || (int64_t)candoff[m] > capacity) {
    if (first == 0 && lane == 0) atomicOr(status, 4);
    return;
  }
  const int64_t stride = (int64_t)gridDim.x * kWavesPerBlock;
  if (nnz_a == 0 || nnz_b == 0) {
    // No candidates anywhere (and nothing the staged loads may read).
    for (int64_t i = first; i < m; i += stride)
      if (lane == 0) counts[i] = 0;
    return;
  }
  const TacoHipSortMasks sm;
  for (int64_t base = first; base < m; base += kRowsTogether * stride) {
    ShortRows<kRowsTogether, false> r;
    r.load(base, stride, m, apos, acrd, nullptr, nnz_a, bpos, bcrd, nullptr, nnz_b);
#pragma unroll
    for (int q = 0; q < kRowsTogether; ++q) {
      if (!r.real[q]) continue;
      const int64_t i = base + q * stride;
      if (r.total[q] > 64) {
        if (lane == 0) mid_rows[atomicAdd(status + 3, 1)] = (int32_t)i;
        continue;
      }
      const int32_t v = (int32_t)taco_hip_wave_sort_u32((uint32_t)r.col[q], sm);
      const int32_t prev = __shfl_up(v, 1, 64);
      const bool head = v != INT_MAX && (lane == 0 || prev != v);
      const uint64_t heads = __ballot(head);
      if (head) scratch[candoff[i] + __popcll(heads & ((1ull << lane) - 1))] = v;
      if (lane == 0) counts[i] = __popcll(heads);
    }
  }
}

// Pass 2 over the rows listed by k_spgemm_symbolic_short: the LDS row set
// for up to TACO_HIP_ROW_CAP distinct columns; rows beyond it are listed in
// long_rows (status[1]) for the global-memory sort, their count left 0.
__global__ __launch_bounds__(256) void k_spgemm_symbolic_listed(
    int32_t m, const int32_t* __restrict__ mid_rows, const int32_t* __restrict__ apos,
    const int32_t* __restrict__ acrd, const int32_t* __restrict__ bpos,
    const int32_t* __restrict__ bcrd, const int32_t* __restrict__ candoff,
    int32_t* __restrict__ scratch, int32_t* __restrict__ counts,
    int32_t* __restrict__ long_rows, int32_t* __restrict__ status) {
  if (status[0] & 4) return;
  const int32_t count = status[3];
  const int64_t stride = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t r = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; r < count;
       r += stride) {
    const int32_t i = mid_rows[r];
    const int32_t c = row_sorted_set(apos, acrd, bpos, bcrd, i, scratch + candoff[i]);
    if (__lane_id() == 0) {
      if (c < 0) long_rows[atomicAdd(status + 1, 1)] = i;
      counts[i] = c < 0 ? 0 : c;
    }
  }
}

// Candidates of the listed long rows, lens[r] (candidate counts from the
// first prefix sum).
__global__ __launch_bounds__(256) void k_spgemm_long_lengths(
    int32_t nlong, const int32_t* __restrict__ long_rows, const int32_t* __restrict__ candoff,
    int32_t* __restrict__ lens) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nlong) return;
  const int32_t i = long_rows[r];
  lens[r] = candoff[i + 1] - candoff[i];
}

// Long row r's candidate columns, in encounter order, at keys[loff[r] ...].
__global__ __launch_bounds__(256) void k_spgemm_long_fill(
    int32_t nlong, const int32_t* __restrict__ long_rows, const uint32_t* __restrict__ loff,
    const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const int32_t* __restrict__ bpos, const int32_t* __restrict__ bcrd,
    uint32_t* __restrict__ keys) {
  const int64_t r = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (r >= nlong) return;
  const int lane = (int)__lane_id();
  const int32_t i = long_rows[r];
  uint32_t o = loff[r];
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    const int32_t b0 = bpos[j], b1 = bpos[j + 1];
    for (int32_t kB = b0 + lane; kB < b1; kB += kWave) keys[o + (uint32_t)(kB - b0)] = (uint32_t)bcrd[kB];
    o += (uint32_t)(b1 - b0);
  }
}

// Long row r's sorted candidates deduplicated into its scratch region;
// counts[i] = the distinct count.
__global__ __launch_bounds__(256) void k_spgemm_long_dedup(
    int32_t nlong, const int32_t* __restrict__ long_rows, const uint32_t* __restrict__ loff,
    const uint32_t* __restrict__ sorted, const int32_t* __restrict__ candoff,
    int32_t* __restrict__ scratch, int32_t* __restrict__ counts) {
  const int64_t r = (int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (r >= nlong) return;
  const int lane = (int)__lane_id();
  const int32_t i = long_rows[r];
  const uint32_t q0 = loff[r], q1 = loff[r + 1];
  int32_t* dst = scratch + candoff[i];
  int32_t base = 0;
  for (uint32_t c0 = q0; c0 < q1; c0 += kWave) {
    const uint32_t q = c0 + (uint32_t)lane;
    bool head = false;
    uint32_t v = 0;
    if (q < q1) {
      v = sorted[q];
      head = q == q0 || sorted[q - 1] != v;
    }
    const uint64_t heads = __ballot(head);
    if (head) dst[base + __popcll(heads & ((1ull << lane) - 1))] = (int32_t)v;
    base += __popcll(heads);
  }
  if (lane == 0) counts[i] = base;
}

// crd: each row's sorted columns copied from its scratch region.  A
// workgroup takes kCompactRows consecutive rows, whose output range in crd
// is contiguous: their pos and region starts go to LDS, and the threads
// stride the range, each finding its row by binary search in LDS, so the
// stores are coalesced whatever the row lengths.
constexpr int kCompactRows = 256;
constexpr int kCompactUnroll = 4;
__global__ __launch_bounds__(256) void k_spgemm_compact(
    int32_t m, const int32_t* __restrict__ candoff, const int32_t* __restrict__ cpos,
    const int32_t* __restrict__ scratch, int32_t* __restrict__ ccrd) {
  __shared__ int32_t s_pos[kCompactRows + 1];
  __shared__ int32_t s_src[kCompactRows];
  const int64_t r0 = (int64_t)blockIdx.x * kCompactRows;
  const int32_t nr = (int32_t)(m - r0 < kCompactRows ? m - r0 : kCompactRows);
  for (int q = threadIdx.x; q <= nr; q += blockDim.x) s_pos[q] = cpos[r0 + q];
  for (int q = threadIdx.x; q < nr; q += blockDim.x) s_src[q] = candoff[r0 + q];
  __syncthreads();
  const int32_t e0 = s_pos[0], e1 = s_pos[nr];
  // kCompactUnroll entries per thread at a time, their row searches
  // interleaved and their loads in flight together; entries past the end
  // are clamped to the last entry of the range.
  for (int32_t e = e0 + (int32_t)threadIdx.x; e < e1;
       e += kCompactUnroll * (int32_t)blockDim.x) {
    int32_t ee[kCompactUnroll], lo[kCompactUnroll];
#pragma unroll
    for (int u = 0; u < kCompactUnroll; ++u) {
      const int32_t x = e + u * (int32_t)blockDim.x;
      ee[u] = x < e1 ? x : e1 - 1;
      lo[u] = 0;
    }
    // The row of entry ee: the largest r < nr with s_pos[r] <= ee.
#pragma unroll
    for (int step = kCompactRows / 2; step > 0; step >>= 1) {
#pragma unroll
      for (int u = 0; u < kCompactUnroll; ++u) {
        const int c = lo[u] + step;
        const int32_t pv = s_pos[c < nr ? c : nr - 1];  // read unconditionally
        lo[u] = ((c < nr) & (pv <= ee[u])) ? c : lo[u];
      }
    }
    int32_t v[kCompactUnroll];
#pragma unroll
    for (int u = 0; u < kCompactUnroll; ++u)
      v[u] = scratch[s_src[lo[u]] + (ee[u] - s_pos[lo[u]])];
    // Entries past the end store the last entry's own value again (ee is
    // clamped), so every lane stores and no load waits under a branch.
#pragma unroll
    for (int u = 0; u < kCompactUnroll; ++u) __builtin_nontemporal_store(v[u], ccrd + ee[u]);
  }
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

namespace {

void publish(taco_tensor_t* C, int32_t* cpos, int32_t* ccrd, double* cvals, int32_t nnz) {
  C->indices[1][0] = (uint8_t*)cpos;
  C->indices[1][1] = (uint8_t*)ccrd;
  C->vals = (uint8_t*)cvals;
  C->vals_size = nnz;
}

// taco's two-phase protocol as it stands: a count pass and a fill pass, each
// building the row set (rows limited to kHashCap distinct columns).  Used
// when the candidate total exceeds the int32 range of the one-sort scratch.
int assemble_two_pass(const CsrView& a, const CsrView& b, taco_tensor_t* C,
                      hipStream_t st) {
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
                "limit of the LDS row kernel (two-pass form, used when the "
                "candidate total exceeds the int32 range)", kHashCap);
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
    publish(C, cpos, ccrd, cvals, nnz);
    return 0;
  } while (0);
  taco_hip_unified_free(cpos);
  return rc;
}

// Workspace slots of the one-sort assemble, and of compute's row list.
constexpr int kSlotCandOff = 8, kSlotScratch = 9, kSlotLongMeta = 10, kSlotLongKeys = 11,
              kSlotSortTemp = 12, kSlotComputeList = 13;
// Workgroups per CU of k_spgemm_numeric_short's resident grid.
constexpr int kShortBlocksPerCu = 8;

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    cus = n > 0 ? n : 256;
  }
  return cus;
}

// Resident grid of the short-row kernels (4 waves per workgroup, each
// taking kRowsTogether rows per step), no larger than the rows need.
unsigned short_rows_grid(int32_t m) {
  const int64_t need = ((int64_t)m + kWavesPerBlock * kRowsTogether - 1) /
                       (kWavesPerBlock * kRowsTogether);
  const int64_t resident = (int64_t)kShortBlocksPerCu * device_cus();
  return (unsigned)(need < resident ? (need > 0 ? need : 1) : resident);
}

// Rows listed by the symbolic pass (more than kHashCap distinct columns):
// candidates written out, sorted per row in global memory, deduplicated into
// their scratch regions with their distinct counts.
int assemble_long_rows(const CsrView& a, const CsrView& b, int32_t nlong,
                       const int32_t* long_rows, const int32_t* candoff, int32_t* scratch,
                       int32_t* counts, hipStream_t st) {
  char* meta = (char*)workspace(sizeof(uint32_t) * (2 * (size_t)nlong + 2), kSlotLongMeta);
  if (!meta) return 1;
  int32_t* lens = (int32_t*)meta;
  uint32_t* loff = (uint32_t*)meta + nlong + 1;
  k_spgemm_long_lengths<<<(nlong + 255) / 256, 256, 0, st>>>(nlong, long_rows, candoff, lens);
  TH_LAUNCH_CHECK("spgemm long rows");
  std::vector<int32_t> h_lens((size_t)nlong);
  TH_CHECK(hipMemcpyAsync(h_lens.data(), lens, sizeof(int32_t) * (size_t)nlong,
                          hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  std::vector<uint32_t> h_off((size_t)nlong + 1, 0);
  uint64_t total = 0;
  for (int32_t r = 0; r < nlong; ++r) {
    h_off[(size_t)r] = (uint32_t)total;
    total += (uint64_t)h_lens[(size_t)r];
  }
  TH_REQUIRE(total < 0xFFFFFFFFull, "spgemm: long rows hold too many candidates");
  h_off[(size_t)nlong] = (uint32_t)total;
  TH_CHECK(hipMemcpyAsync(loff, h_off.data(), sizeof(uint32_t) * ((size_t)nlong + 1),
                          hipMemcpyHostToDevice, st));
  uint32_t* keys = (uint32_t*)workspace(sizeof(uint32_t) * 2 * (size_t)(total ? total : 1),
                                        kSlotLongKeys);
  if (!keys) return 1;
  uint32_t* sorted = keys + total;
  const unsigned grid = (unsigned)(((int64_t)nlong + kWavesPerBlock - 1) / kWavesPerBlock);
  k_spgemm_long_fill<<<grid, 256, 0, st>>>(nlong, long_rows, loff, a.pos, a.crd, b.pos, b.crd,
                                           keys);
  TH_LAUNCH_CHECK("spgemm long rows: candidates");
  unsigned bits = 1;
  while (bits < 32 && (1ull << bits) < (uint64_t)(b.n > 1 ? b.n : 2)) ++bits;
  size_t temp_bytes = 0;
  TH_CHECK(rocprim::segmented_radix_sort_keys(nullptr, temp_bytes, keys, sorted,
                                              (unsigned)total, (unsigned)nlong, loff, loff + 1,
                                              0, bits, st));
  void* temp = workspace(temp_bytes ? temp_bytes : 1, kSlotSortTemp);
  if (!temp) return 1;
  TH_CHECK(rocprim::segmented_radix_sort_keys(temp, temp_bytes, keys, sorted, (unsigned)total,
                                              (unsigned)nlong, loff, loff + 1, 0, bits, st));
  k_spgemm_long_dedup<<<grid, 256, 0, st>>>(nlong, long_rows, loff, sorted, candoff, scratch,
                                            counts);
  TH_LAUNCH_CHECK("spgemm long rows: distinct columns");
  return 0;
}

// The one-sort assemble (see the file comment).  *two_pass is set when the
// candidate total exceeds the int32 range; nothing is published then.
int assemble_one_sort(const CsrView& a, const CsrView& b, taco_tensor_t* C, hipStream_t st,
                      bool* two_pass) {
  *two_pass = false;
  const int32_t m = a.m;
  // status[0]: flags (2: nnz(C) beyond int32, 4: scratch too small);
  // status[1]: long rows listed; status[2]: candidate-total flags (2).
  char* ws = (char*)workspace(256 + sizeof(int32_t) * 3 * ((size_t)m + 1), 2);
  int32_t* candoff = (int32_t*)workspace(sizeof(int32_t) * ((size_t)m + 1), kSlotCandOff);
  if (!ws || !candoff) return 1;
  int32_t* status = (int32_t*)ws;
  int32_t* counts = (int32_t*)(ws + 256);
  int32_t* long_rows = counts + m + 1;
  int32_t* mid_rows = long_rows + m + 1;
  int32_t* cpos = (int32_t*)taco_hip_unified_alloc(sizeof(int32_t) * ((size_t)m + 1));
  if (!cpos) return 1;
  const unsigned short_grid = short_rows_grid(m);
  const unsigned listed_grid = (unsigned)taco_hip_row_grid(m);
  int rc = 0;
  do {
    // The status words are cleared by k_spgemm_row_cands (a memset costs a
    // runtime call of its own).
    if (m > 0)
      k_spgemm_row_cands<<<(unsigned)(((int64_t)m * kCandLanes + 255) / 256), 256, 0, st>>>(
          m, a.pos, a.crd, b.pos, counts, status);
    else if (hipMemsetAsync(status, 0, 16 * sizeof(int32_t), st) != hipSuccess) {
      rc = fail("spgemm: memset failed");
      break;
    }
    if ((rc = device_counts_to_pos(m, counts, candoff, status + 2, st))) break;
    size_t cap_bytes = workspace_size(kSlotScratch);
    int32_t* scratch = (int32_t*)workspace(cap_bytes ? cap_bytes : 4096, kSlotScratch);
    if (!scratch) {
      rc = 1;
      break;
    }
    cap_bytes = workspace_size(kSlotScratch);
    int32_t host[5] = {0, 0, 0, 0, 0};  // status[0..2], candoff[m], cpos[m]
    for (int attempt = 0; attempt < 2; ++attempt) {
      if (attempt > 0 &&
          (hipMemsetAsync(status, 0, 2 * sizeof(int32_t), st) != hipSuccess ||
           hipMemsetAsync(status + 3, 0, sizeof(int32_t), st) != hipSuccess)) {
        rc = fail("spgemm: memset failed");
        break;
      }
      if (m > 0) {
        k_spgemm_symbolic_short<<<short_grid, 256, 0, st>>>(
            m, a.pos, a.crd, (int32_t)a.nnz, b.pos, b.crd, (int32_t)b.nnz, candoff,
            (int64_t)(cap_bytes / sizeof(int32_t)), scratch, counts, mid_rows, status);
        k_spgemm_symbolic_listed<<<listed_grid, 256, 0, st>>>(
            m, mid_rows, a.pos, a.crd, b.pos, b.crd, candoff, scratch, counts, long_rows,
            status);
      }
      if ((rc = device_counts_to_pos(m, counts, cpos, status, st))) break;
      k_spgemm_gather_result<<<1, 64, 0, st>>>(status, candoff + m, cpos + m, status + 8);
      if ((rc = read_back_i32(status + 8, 5, host, st))) break;
      if ((host[2] & 2) || !(host[0] & 4)) break;
      // The scratch was too small for the candidate total: grow, run again.
      cap_bytes = sizeof(int32_t) * (size_t)host[3];
      scratch = (int32_t*)workspace(cap_bytes, kSlotScratch);
      if (!scratch) {
        rc = 1;
        break;
      }
    }
    if (rc) break;
    if (host[2] & 2) {
      *two_pass = true;
      break;
    }
    TH_REQUIRE(!(host[0] & 4), "spgemm: scratch sizing failed");
    if (host[1] > 0) {
      if ((rc = assemble_long_rows(a, b, host[1], long_rows, candoff, scratch, counts, st)))
        break;
      if (hipMemsetAsync(status, 0, sizeof(int32_t), st) != hipSuccess) {
        rc = fail("spgemm: memset failed");
        break;
      }
      if ((rc = device_counts_to_pos(m, counts, cpos, status, st))) break;
      k_spgemm_gather_result<<<1, 64, 0, st>>>(status, candoff + m, cpos + m, status + 8);
      if ((rc = read_back_i32(status + 8, 5, host, st))) break;
    }
    if (host[0] & 2) {
      rc = fail("spgemm: nnz(C) exceeds the int32 index range");
      break;
    }
    const int32_t nnz = host[4];
    int32_t* ccrd = (int32_t*)taco_hip_unified_alloc(sizeof(int32_t) * (size_t)nnz);
    double* cvals = (double*)taco_hip_unified_alloc(sizeof(double) * (size_t)nnz);
    if (nnz > 0 && (!ccrd || !cvals)) {
      taco_hip_unified_free(ccrd);
      taco_hip_unified_free(cvals);
      rc = 1;
      break;
    }
    if (m > 0 && nnz > 0)
      k_spgemm_compact<<<(unsigned)(((int64_t)m + kCompactRows - 1) / kCompactRows), 256, 0,
                         st>>>(m, candoff, cpos, scratch, ccrd);
    if ((rc = finish("taco_hip_spgemm_assemble"))) {
      taco_hip_unified_free(ccrd);
      taco_hip_unified_free(cvals);
      break;
    }
    publish(C, cpos, ccrd, cvals, nnz);
    return 0;
  } while (0);
  taco_hip_unified_free(cpos);
  return rc;
}

}  // namespace

extern "C" int taco_hip_spgemm_assemble(taco_tensor_t* C, taco_tensor_t* A,
                                        taco_tensor_t* B) {
  CsrView a, b;
  if (int rc = check_spgemm_operands(C, A, B, &a, &b)) return rc;
  hipStream_t st = current_stream();
  bool two_pass = false;
  if (int rc = assemble_one_sort(a, b, C, st, &two_pass)) return rc;
  return two_pass ? assemble_two_pass(a, b, C, st) : 0;
}

extern "C" int taco_hip_spgemm_compute(taco_tensor_t* C, taco_tensor_t* A,
                                       taco_tensor_t* B) {
  CsrView a, b, c;
  if (int rc = check_spgemm_operands(C, A, B, &a, &b)) return rc;
  if (int rc = view_csr(C, "C", &c)) return rc;
  if (a.m == 0 || c.nnz == 0) return finish("taco_hip_spgemm_compute");
  hipStream_t st = current_stream();
  // Workspace: the count of listed rows, then the list.
  char* ws = (char*)workspace(256 + sizeof(int32_t) * (size_t)a.m, kSlotComputeList);
  if (!ws) return 1;
  int32_t* nlong = (int32_t*)ws;
  int32_t* long_rows = (int32_t*)(ws + 256);
  k_spgemm_zero_count<<<1, 64, 0, st>>>(nlong);
  k_spgemm_numeric_short<<<short_rows_grid(a.m), 256, 0, st>>>(
      a.m, a.pos, a.crd, (const double*)a.vals, (int32_t)a.nnz, b.pos, b.crd,
      (const double*)b.vals, (int32_t)b.nnz, c.pos, (double*)c.vals, long_rows, nlong);
  k_spgemm_numeric_listed<<<(unsigned)taco_hip_row_grid(a.m), 256, 0, st>>>(
      long_rows, nlong, a.pos, a.crd, (const double*)a.vals, b.pos, b.crd,
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
