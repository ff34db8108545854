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
// Row kernels, one 64-lane wave per row:
//   short rows (<= 64 candidates): candidates in lanes, bitonic sort across
//     the wave with shuffles.  The numeric phase sorts 64-bit keys
//     (column << 32 | encounter index), so equal columns come out in
//     encounter order and are summed sequentially in that order.
//   long rows: a per-wave LDS hash set of up to kHashCap distinct columns,
//     then an in-LDS bitonic sort of the distinct columns; the numeric phase
//     accumulates into LDS slots, adding candidates one at a time in
//     encounter order.

#include <climits>

#include "common.h"
#include "device.h"

namespace taco_hip {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kHashCap = 1024;  // distinct columns per long row

template <typename K>
__device__ __forceinline__ K wave_bitonic_sort(K v) {
  const int lane = (int)__lane_id();
#pragma unroll
  for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const K p = __shfl_xor(v, j, kWave);
      const bool up = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      v = (lower == up) ? (p < v ? p : v) : (p > v ? p : v);
    }
  }
  return v;
}

// Candidate count of row i: sum of B row lengths over row i of A.
__device__ __forceinline__ int64_t row_candidates(
    const int32_t* apos, const int32_t* acrd, const int32_t* bpos, int32_t i) {
  const int lane = (int)__lane_id();
  int64_t n = 0;
  for (int32_t p = apos[i] + lane; p < apos[i + 1]; p += kWave) {
    const int32_t j = acrd[p];
    n += bpos[j + 1] - bpos[j];
  }
  return wave_sum(n);
}

// Candidate q (0-based, encounter order) of row i for short rows: lane q
// walks the row of A to find which B row holds its candidate.  Every lane
// scans A's row; rows are short on this path (<= 64 candidates).
__device__ __forceinline__ bool short_candidate(
    const int32_t* apos, const int32_t* acrd, const int32_t* bpos,
    const int32_t* bcrd, int32_t i, int q, int32_t* col, int32_t* jA_out,
    int32_t* kB_out) {
  int acc = 0;
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    const int32_t b0 = bpos[j], len = bpos[j + 1] - b0;
    if (q < acc + len) {
      *jA_out = p;
      *kB_out = b0 + (q - acc);
      *col = bcrd[b0 + (q - acc)];
      return true;
    }
    acc += len;
  }
  return false;
}

struct WaveLds {
  int32_t keys[kHashCap];
  int32_t count;
};

// Inserts every candidate column of row i into the wave's LDS hash set.
// Returns the number of distinct columns, or -1 when it exceeds kHashCap.
__device__ int32_t hash_row(WaveLds* w, const int32_t* apos,
                            const int32_t* acrd, const int32_t* bpos,
                            const int32_t* bcrd, int32_t i) {
  const int lane = (int)__lane_id();
  for (int q = lane; q < kHashCap; q += kWave) w->keys[q] = -1;
  if (lane == 0) w->count = 0;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  bool overflow = false;
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    for (int32_t kB = bpos[j] + lane; kB < bpos[j + 1]; kB += kWave) {
      const int32_t col = bcrd[kB];
      uint32_t h = ((uint32_t)col * 2654435761u) & (kHashCap - 1);
      for (int probe = 0; probe < kHashCap; ++probe) {
        const int32_t prev = atomicCAS(&w->keys[h], -1, col);
        if (prev == -1) {
          atomicAdd(&w->count, 1);
          break;
        }
        if (prev == col) break;
        h = (h + 1) & (kHashCap - 1);
        if (probe == kHashCap - 1) overflow = true;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const bool any_overflow = __ballot(overflow) != 0;
  const int32_t cnt = w->count;
  return any_overflow || cnt > kHashCap ? -1 : cnt;
}

// Sorts the distinct keys of the hash set into w->keys[0..cnt) ascending.
__device__ void compact_sort(WaveLds* w, int32_t cnt) {
  const int lane = (int)__lane_id();
  // Compact: empty slots become INT_MAX and sort to the end.
  for (int q = lane; q < kHashCap; q += kWave)
    if (w->keys[q] == -1) w->keys[q] = INT_MAX;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int k = 2; k <= kHashCap; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int q = lane; q < kHashCap; q += kWave) {
        const int partner = q ^ j;
        if (partner > q) {
          const int32_t a = w->keys[q], b = w->keys[partner];
          const bool up = (q & k) == 0;
          if ((a > b) == up) {
            w->keys[q] = b;
            w->keys[partner] = a;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  (void)cnt;
}

// Symbolic pass.  fill == false: counts[i] = distinct columns of row i.
// fill == true: writes row i's sorted columns to ccrd[cpos[i] ...].
// status[0] is set to 1 when a long row exceeds kHashCap distinct columns.
template <bool FILL>
__global__ __launch_bounds__(256) void k_spgemm_symbolic(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const int32_t* __restrict__ bpos, const int32_t* __restrict__ bcrd,
    int32_t* __restrict__ counts, const int32_t* __restrict__ cpos,
    int32_t* __restrict__ ccrd, int32_t* __restrict__ status) {
  __shared__ WaveLds s_w[kWavesPerBlock];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int64_t i64 = (int64_t)blockIdx.x * kWavesPerBlock + wave;
  if (i64 >= m) return;
  const int32_t i = (int32_t)i64;
  const int64_t n = row_candidates(apos, acrd, bpos, i);
  if (n <= kWave) {
    int32_t col = INT_MAX, jA = 0, kB = 0;
    if (lane < n) short_candidate(apos, acrd, bpos, bcrd, i, lane, &col, &jA, &kB);
    const int32_t s = wave_bitonic_sort(col);
    const int32_t prev = __shfl_up(s, 1, kWave);
    const bool head = s != INT_MAX && (lane == 0 || prev != s);
    const uint64_t heads = __ballot(head);
    if (!FILL) {
      if (lane == 0) counts[i] = __popcll(heads);
    } else if (head) {
      const int rank = __popcll(heads & ((1ull << lane) - 1));
      ccrd[cpos[i] + rank] = s;
    }
    return;
  }
  WaveLds* w = &s_w[wave];
  const int32_t cnt = hash_row(w, apos, acrd, bpos, bcrd, i);
  if (cnt < 0) {
    if (lane == 0) {
      atomicOr(status, 1);
      if (!FILL) counts[i] = 0;
    }
    return;
  }
  if (!FILL) {
    if (lane == 0) counts[i] = cnt;
    return;
  }
  compact_sort(w, cnt);
  const int32_t base = cpos[i];
  for (int q = lane; q < cnt; q += kWave) ccrd[base + q] = w->keys[q];
}

// Numeric pass on the assembled pattern.
__global__ __launch_bounds__(256) void k_spgemm_numeric(
    int32_t m, const int32_t* __restrict__ apos, const int32_t* __restrict__ acrd,
    const double* __restrict__ avals, const int32_t* __restrict__ bpos,
    const int32_t* __restrict__ bcrd, const double* __restrict__ bvals,
    const int32_t* __restrict__ cpos, const int32_t* __restrict__ ccrd,
    double* __restrict__ cvals) {
  __shared__ double s_acc[kWavesPerBlock][kHashCap];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int64_t i64 = (int64_t)blockIdx.x * kWavesPerBlock + wave;
  if (i64 >= m) return;
  const int32_t i = (int32_t)i64;
  const int64_t n = row_candidates(apos, acrd, bpos, i);
  const int32_t c0 = cpos[i], nc = cpos[i + 1] - c0;
  if (n <= kWave) {
    int32_t col = INT_MAX, jA = 0, kB = 0;
    double prod = 0;
    if (lane < n) {
      short_candidate(apos, acrd, bpos, bcrd, i, lane, &col, &jA, &kB);
      prod = avals[jA] * bvals[kB];
    }
    // Key: column, then encounter index; carries the product's source lane.
    uint64_t key = ((uint64_t)(uint32_t)col << 32) | (uint32_t)lane;
    key = wave_bitonic_sort(key);
    const int32_t scol = (int32_t)(key >> 32);
    const int src = (int)(key & 63);
    const double v = __shfl(prod, src, kWave);
    const int32_t prev = __shfl_up(scol, 1, kWave);
    const bool valid = scol != INT_MAX;
    const bool head = valid && (lane == 0 || prev != scol);
    const uint64_t heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    // Run of a head: lanes up to the next head (or the end of valid lanes).
    int run = 0;
    if (head) {
      const uint64_t above = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1);
      const int next = above ? __ffsll((unsigned long long)above) - 1 : nvalid;
      run = next - lane;
    }
    int maxrun = run;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const int o = __shfl_xor(maxrun, off, kWave);
      maxrun = o > maxrun ? o : maxrun;
    }
    // Sequential sum of each run in encounter order: s = v0; s = s + v1; ...
    double s = v;
    for (int q = 1; q < maxrun; ++q) {
      const double u = __shfl(v, (lane + q) & 63, kWave);
      if (head && q < run) s = s + u;
    }
    if (head) {
      const int rank = __popcll(heads & ((1ull << lane) - 1));
      if (rank < nc) cvals[c0 + rank] = s;
    }
    return;
  }
  // Long row: LDS slot per pattern entry (nc <= kHashCap was checked by the
  // symbolic pass), candidates added one at a time in encounter order.
  // A caller-supplied pattern longer than the slot array is left unwritten
  // rather than overrunning LDS.
  if (nc > kHashCap) return;
  double* acc = s_acc[wave];
  for (int q = lane; q < nc; q += kWave) acc[q] = 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int32_t* crow = ccrd + c0;
  for (int32_t p = apos[i]; p < apos[i + 1]; ++p) {
    const int32_t j = acrd[p];
    const double a = avals[p];
    for (int32_t kb0 = bpos[j]; kb0 < bpos[j + 1]; kb0 += kWave) {
      const int32_t kB = kb0 + lane;
      int slot = -1;
      double prod = 0;
      if (kB < bpos[j + 1]) {
        const int32_t col = bcrd[kB];
        prod = a * bvals[kB];
        int lo = 0, hi = nc;  // crow sorted ascending
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (crow[mid] < col) lo = mid + 1; else hi = mid;
        }
        slot = lo < nc && crow[lo] == col ? lo : -1;
      }
      const int lim = bpos[j + 1] - kb0 < kWave ? bpos[j + 1] - kb0 : kWave;
      for (int q = 0; q < lim; ++q) {
        const int sq = __shfl(slot, q, kWave);
        const double vq = __shfl(prod, q, kWave);
        if (lane == 0 && sq >= 0) acc[sq] = acc[sq] + vq;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  for (int q = lane; q < nc; q += kWave) cvals[c0 + q] = acc[q];
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
  const unsigned grid = (unsigned)(((int64_t)a.m + kWavesPerBlock - 1) / kWavesPerBlock);
  k_spgemm_numeric<<<grid, 256, 0, st>>>(
      a.m, a.pos, a.crd, (const double*)a.vals, b.pos, b.crd,
      (const double*)b.vals, c.pos, c.crd, (double*)c.vals);
  return finish("taco_hip_spgemm_compute");
}
