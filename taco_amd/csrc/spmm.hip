// Hand-written HIP SpMM kernels, C(i,k) = A(i,j) * B(j,k): A in CSR, B and C
// dense row-major.  Value forms: f64, f32, and bf16 A/B with f32 C.
//
// Schedules (wave64 forms of the reference's GPU SpMM schedules):
//   THREAD_PER_ROW  taco's default split(i,256) -> GPUBlock/GPUThread
//                   (reference transformations.cpp:1585-1596): one thread
//                   owns a row of C and runs the generated i-j-k loop nest.
//   GROUP_PER_ROW   scheduleSpMMRowsGPU (tests-scheduling-eval.cpp:345-355):
//                   a lane group owns rows, lanes split k.
//   NNZ_SPLIT       scheduleSpMMGPU (tests-scheduling-eval.cpp:249-268):
//                   fuse(i,j) -> pos -> fixed entry chunks per lane group,
//                   lanes split k, atomics at row changes.
//   MERGE_PATH      merge-path over (row ends, entries) per lane group with
//                   lanes on k; a group's unfinished last row is written to a
//                   carry slot and added in order by a fix-up kernel.
//   MFMA            bf16 only; see spmm_mfma.hip.

#include <climits>

#include "common.h"
#include "device.h"

namespace taco_hip {

// bf16 stored as raw 16-bit patterns; widening to f32 is exact.
struct bf16_t {
  uint16_t bits;
};
__device__ __forceinline__ float widen(bf16_t v) {
  return __uint_as_float((uint32_t)v.bits << 16);
}
__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ bf16_t stream_load(const bf16_t* p) {
  return bf16_t{__builtin_nontemporal_load(&p->bits)};
}
__device__ __forceinline__ double widen(double v) { return v; }

namespace {

template <typename TA, typename TC>
__global__ __launch_bounds__(1024) void k_spmm_thread_per_row(
    int32_t m, int32_t K, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ crd, const TA* __restrict__ vals,
    const TA* __restrict__ B, TC* __restrict__ C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  TC* crow = C + i * K;
  for (int32_t k = 0; k < K; ++k) crow[k] = 0;
  for (int32_t p = pos[i]; p < pos[i + 1]; ++p) {
    const TC a = widen(vals[p]);
    const TA* brow = B + (int64_t)crd[p] * K;
    for (int32_t k = 0; k < K; ++k) crow[k] += a * widen(brow[k]);
  }
}

// G lanes per row team, CPL columns per lane (K <= G * CPL).
template <typename TA, typename TC, int G, int CPL>
__global__ __launch_bounds__(256) void k_spmm_group_per_row(
    int32_t m, int32_t K, int32_t rows_per_group,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const TA* __restrict__ vals, const TA* __restrict__ B,
    TC* __restrict__ C) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t team = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  for (int32_t rr = 0; rr < rows_per_group; ++rr) {
    const int64_t i = team * rows_per_group + rr;
    if (i >= m) break;
    TC acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c] = 0;
    const int32_t end = pos[i + 1];
    int32_t p = pos[i];
    for (; p + 4 <= end; p += 4) {
      int32_t j[4];
      TC a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        j[u] = stream_load(crd + p + u);
        a[u] = widen(stream_load(vals + p + u));
      }
      TC b[4][CPL];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const int k = lane + c * G;
          b[u][c] = k < K ? widen(B[(int64_t)j[u] * K + k]) : TC(0);
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] += a[u] * b[u][c];
    }
    for (; p < end; ++p) {
      const int32_t j = stream_load(crd + p);
      const TC a = widen(stream_load(vals + p));
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const int k = lane + c * G;
        if (k < K) acc[c] += a * widen(B[(int64_t)j * K + k]);
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int k = lane + c * G;
      if (k < K) C[i * K + k] = acc[c];
    }
  }
}

template <typename TA, typename TC, int G, int CPL>
__global__ __launch_bounds__(256) void k_spmm_nnz_split(
    int32_t m, int64_t nnz, int32_t K, int32_t nnz_per_team,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const TA* __restrict__ vals, const TA* __restrict__ B,
    TC* __restrict__ C, bool coarse) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t team = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int64_t z0 = team * nnz_per_team;
  if (z0 >= nnz) return;
  const int64_t z1 = z0 + nnz_per_team < nnz ? z0 + nnz_per_team : nnz;
  // Row holding entry z0 (taco_binarySearchBefore).
  int32_t lo = 0, hi = m;
  while (hi - lo > 1) {
    int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)pos[mid] <= z0) lo = mid; else hi = mid;
  }
  int32_t r = lo;
  int64_t rend = pos[r + 1];
  TC acc[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) acc[c] = 0;
  for (int64_t z = z0; z < z1; ++z) {
    if (z >= rend) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const int k = lane + c * G;
        if (k < K) atomic_add(C + (int64_t)r * K + k, acc[c], coarse);
        acc[c] = 0;
      }
      do {
        ++r;
        rend = pos[r + 1];
      } while (z >= rend);
    }
    const int32_t j = stream_load(crd + z);
    const TC a = widen(stream_load(vals + z));
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int k = lane + c * G;
      if (k < K) acc[c] += a * widen(B[(int64_t)j * K + k]);
    }
  }
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int k = lane + c * G;
    if (k < K) atomic_add(C + (int64_t)r * K + k, acc[c], coarse);
  }
}

// Merge-path split (see wave_merge_search) for a team of G lanes.
template <int G>
__device__ __forceinline__ int64_t group_merge_search(const int32_t* pos,
                                                      int64_t m, int64_t nnz,
                                                      int64_t d) {
  int64_t lo = d - nnz > 0 ? d - nnz : 0;
  int64_t hi = d < m ? d : m;
  const int lane = (int)__lane_id() & (G - 1);
  const int base = (int)__lane_id() & ~(G - 1);
  while (lo < hi) {
    const int64_t step = (hi - lo + G - 1) / G;
    const int64_t p = lo + (int64_t)lane * step;
    bool pred = false;
    if (p < hi) pred = (int64_t)pos[p + 1] <= d - p - 1;
    uint64_t mask = __ballot(pred) >> base;
    if (G < 64) mask &= ((1ull << G) - 1);
    const int c = __popcll(mask);
    const int64_t nlo = c == 0 ? lo : lo + (int64_t)(c - 1) * step + 1;
    int64_t nhi = lo + (int64_t)c * step;
    if (nhi > hi) nhi = hi;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// Value of lane `idx` of this lane's team of G lanes (idx uniform across
// the team).  A whole-wave team reads it with v_readlane into a scalar.
template <int G, typename T>
__device__ __forceinline__ T team_bcast(T v, int idx) {
  if constexpr (G == 64) {
    if constexpr (sizeof(T) == 8) {
      const uint64_t b = __builtin_bit_cast(uint64_t, v);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, idx);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), idx);
      return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
      return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), idx));
    }
  } else {
    return __shfl(v, idx, G);
  }
}

template <typename TA, typename TC, int G, int CPL>
__global__ __launch_bounds__(256) void k_spmm_merge(
    int32_t m, int64_t nnz, int32_t K, int32_t items_per_team,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const TA* __restrict__ vals, const TA* __restrict__ B,
    TC* __restrict__ C, int64_t nteams, int32_t* __restrict__ carry_row,
    TC* __restrict__ carry_val) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t team = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  // Whole teams only: the grid is sized so a team is either fully in range
  // or fully out, and ballots inside the search see complete teams.
  if (team >= nteams) return;
  const int64_t total = (int64_t)m + nnz;
  int64_t d = team * items_per_team;
  const int64_t dend = d + items_per_team < total ? d + items_per_team : total;
  int64_t r = group_merge_search<G>(pos, m, nnz, d);
  int64_t z = d - r;
  TC acc[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) acc[c] = 0;
  bool partial = false;  // acc holds entries of row r
  while (d < dend) {
    const int64_t rend = pos[r + 1];  // r < m while items remain
    int64_t zstop = z + (dend - d);
    if (zstop > rend) zstop = rend;
    const int64_t take = zstop - z;
    if (take > 0) partial = true;
    // The team loads up to G consecutive entries at once (one per lane,
    // coalesced) and broadcasts them one by one; 8 B-row gathers in flight.
    for (int64_t zc = z; zc < zstop; zc += G) {
      const int cnt = (int)(zstop - zc < G ? zstop - zc : G);
      const int32_t jl = lane < cnt ? stream_load(crd + zc + lane) : 0;
      const TC al = lane < cnt ? widen(stream_load(vals + zc + lane)) : TC(0);
      int u = 0;
      for (; u + 8 <= cnt; u += 8) {
        int32_t j[8];
        TC a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          j[q] = team_bcast<G>(jl, u + q);
          a[q] = team_bcast<G>(al, u + q);
        }
        TC b[8][CPL];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int k = lane + c * G;
            b[q][c] = k < K ? widen(B[(int64_t)j[q] * K + k]) : TC(0);
          }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int c = 0; c < CPL; ++c) acc[c] += a[q] * b[q][c];
      }
      for (; u < cnt; ++u) {
        const int32_t j = team_bcast<G>(jl, u);
        const TC a = team_bcast<G>(al, u);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const int k = lane + c * G;
          if (k < K) acc[c] += a * widen(B[(int64_t)j * K + k]);
        }
      }
    }
    z = zstop;
    d += take;
    if (d < dend && z == rend) {
      // Row end item: row r completes in this team (its earlier entries, if
      // any, arrive through the carries of preceding teams).
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const int k = lane + c * G;
        // non-temporal: C is written once, B rows keep the caches
        if (k < K) __builtin_nontemporal_store(acc[c], C + r * K + k);
        acc[c] = 0;
      }
      partial = false;
      ++r;
      ++d;
    }
  }
  const bool has_carry = partial && r < m;
  if (lane == 0) carry_row[team] = has_carry ? (int32_t)r : -1;
  if (has_carry) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int k = lane + c * G;
      if (k < K) carry_val[team * K + k] = acc[c];
    }
  }
}

// One 64-lane wave per carry slot; the first slot of each run of equal rows
// adds the whole run, in team order.
template <typename TC>
__global__ __launch_bounds__(256) void k_spmm_merge_fixup(
    int64_t nteams, int32_t K, const int32_t* __restrict__ carry_row,
    const TC* __restrict__ carry_val, TC* __restrict__ C) {
  const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x % kWave;
  if (t >= nteams) return;
  const int32_t r = carry_row[t];
  if (r < 0) return;
  if (t > 0 && carry_row[t - 1] == r) return;
  int64_t e = t + 1;
  while (e < nteams && carry_row[e] == r) ++e;
  for (int32_t k = lane; k < K; k += kWave) {
    TC s = carry_val[t * K + k];
    int64_t q = t + 1;
    // Loads 8 carries ahead, adds them in team order (long rows have
    // hundreds of carries; the order of the sum stays fixed).
    for (; q + 8 <= e; q += 8) {
      TC v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = carry_val[(q + u) * K + k];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; q < e; ++q) s += carry_val[q * K + k];
    C[(int64_t)r * K + k] += s;
  }
}

template <typename TA, typename TC>
int run_spmm(const CsrView& a, int32_t K, const TA* B, TC* C,
             const taco_hip_schedule& s) {
  hipStream_t st = current_stream();
  if (a.m == 0 || K == 0) return 0;
  const int algo = s.algo;
  if (algo == TACO_HIP_THREAD_PER_ROW) {
    int bs = s.block_size > 0 ? s.block_size : 256;
    TH_REQUIRE(bs % kWave == 0 && bs <= 1024, "spmm: bad block size %d", bs);
    unsigned grid = (unsigned)((a.m + bs - 1) / bs);
    k_spmm_thread_per_row<TA, TC><<<grid, bs, 0, st>>>(
        a.m, K, a.pos, a.crd, (const TA*)a.vals, B, C);
    return 0;
  }
  // Lane-team shape: G lanes along k, CPL columns per lane.
  int G = s.group_size > 0 ? s.group_size : (K >= 64 ? 64 : 1);
  if (s.group_size <= 0)
    while (G < K && G < 64) G <<= 1;
  TH_REQUIRE(G == 8 || G == 16 || G == 32 || G == 64,
             "spmm: group_size %d must be 8, 16, 32 or 64", G);
  const int cpl = (K + G - 1) / G;
  TH_REQUIRE(cpl <= 4, "spmm: K=%d too wide for %d lanes (max 4 per lane)",
             K, G);
  const int CPL = cpl <= 1 ? 1 : cpl <= 2 ? 2 : 4;
  const int bs = 256;
  const int teams_per_block = bs / G;

#define TH_SHAPES(X)                                                       \
  X(8, 1) X(8, 2) X(8, 4) X(16, 1) X(16, 2) X(16, 4) X(32, 1) X(32, 2)     \
  X(32, 4) X(64, 1) X(64, 2) X(64, 4)

  if (algo == TACO_HIP_GROUP_PER_ROW) {
    int rpg = s.rows_per_group > 0 ? s.rows_per_group : 1;
    int64_t teams = ((int64_t)a.m + rpg - 1) / rpg;
    unsigned grid = (unsigned)((teams + teams_per_block - 1) / teams_per_block);
#define TH_X(g, c)                                                          \
  if (G == g && CPL == c) {                                                 \
    k_spmm_group_per_row<TA, TC, g, c><<<grid, bs, 0, st>>>(                \
        a.m, K, rpg, a.pos, a.crd, (const TA*)a.vals, B, C);                \
    return 0;                                                               \
  }
    TH_SHAPES(TH_X)
#undef TH_X
    return fail("spmm: no kernel for G=%d CPL=%d", G, CPL);
  }
  if (algo == TACO_HIP_NNZ_SPLIT) {
    int npt = s.items_per_thread > 0 ? s.items_per_thread : 8;
    if (int rc = zero_fill(C, sizeof(TC) * (size_t)a.m * K, st)) return rc;
    if (a.nnz == 0) return 0;
    int64_t teams = (a.nnz + npt - 1) / npt;
    unsigned grid = (unsigned)((teams + teams_per_block - 1) / teams_per_block);
#define TH_X(g, c)                                                          \
  if (G == g && CPL == c) {                                                 \
    k_spmm_nnz_split<TA, TC, g, c><<<grid, bs, 0, st>>>(                    \
        a.m, a.nnz, K, npt, a.pos, a.crd, (const TA*)a.vals, B, C,          \
        coarse_grained(C));                                                 \
    return 0;                                                               \
  }
    TH_SHAPES(TH_X)
#undef TH_X
    return fail("spmm: no kernel for G=%d CPL=%d", G, CPL);
  }
  if (algo == TACO_HIP_MERGE_PATH) {
    int ipt = s.items_per_thread > 0 ? s.items_per_thread : 256;
    const int64_t total = (int64_t)a.m + a.nnz;
    const int64_t nteams = (total + ipt - 1) / ipt;
    const size_t row_bytes = ((size_t)nteams * sizeof(int32_t) + 255) / 256 * 256;
    void* ws = workspace(row_bytes + (size_t)nteams * K * sizeof(TC), 1);
    if (!ws) return 1;
    int32_t* crow = (int32_t*)ws;
    TC* cval = (TC*)((char*)ws + row_bytes);
    unsigned grid = (unsigned)((nteams + teams_per_block - 1) / teams_per_block);
    bool launched = false;
#define TH_X(g, c)                                                          \
  if (!launched && G == g && CPL == c) {                                    \
    k_spmm_merge<TA, TC, g, c><<<grid, bs, 0, st>>>(                        \
        a.m, a.nnz, K, ipt, a.pos, a.crd, (const TA*)a.vals, B, C, nteams,  \
        crow, cval);                                                        \
    launched = true;                                                        \
  }
    TH_SHAPES(TH_X)
#undef TH_X
    TH_REQUIRE(launched, "spmm: no kernel for G=%d CPL=%d", G, CPL);
    TH_LAUNCH_CHECK("k_spmm_merge");
    unsigned g2 = (unsigned)((nteams * kWave + 255) / 256);
    k_spmm_merge_fixup<TC><<<g2, 256, 0, st>>>(nteams, K, crow, cval, C);
    return 0;
  }
#undef TH_SHAPES
  return fail("spmm: schedule algo %d not available here", algo);
}

}  // namespace

int spmm_mfma_bf16(const CsrView& a, int32_t K, const void* B, float* C,
                   const taco_hip_schedule& s);

}  // namespace taco_hip

using namespace taco_hip;

extern "C" int taco_hip_spmm_compute(taco_tensor_t* C, taco_tensor_t* A,
                                     taco_tensor_t* B,
                                     const taco_hip_schedule* sched) {
  CsrView a;
  DenseView bv, cv;
  if (int rc = view_csr(A, "A", &a)) return rc;
  if (int rc = view_dense(B, "B", 2, &bv)) return rc;
  if (int rc = view_dense(C, "C", 2, &cv)) return rc;
  TH_REQUIRE(bv.dims[0] == a.n, "spmm: B has %d rows, A has %d columns",
             bv.dims[0], a.n);
  TH_REQUIRE(cv.dims[0] == a.m && cv.dims[1] == bv.dims[1],
             "spmm: C is %dx%d, expected %dx%d", cv.dims[0], cv.dims[1], a.m,
             bv.dims[1]);
  const int32_t K = bv.dims[1];
  taco_hip_schedule s;
  if (sched)
    s = *sched;
  else
    taco_hip_default_schedule(1, a.m, a.nnz, K, &s);
  int rc;
  if (a.vt == VType::F64 && bv.vt == VType::F64 && cv.vt == VType::F64) {
    TH_REQUIRE(s.algo != TACO_HIP_MFMA, "spmm: MFMA schedule needs bf16 A/B");
    rc = run_spmm<double, double>(a, K, (const double*)bv.vals,
                                  (double*)cv.vals, s);
  } else if (a.vt == VType::F32 && bv.vt == VType::F32 &&
             cv.vt == VType::F32) {
    TH_REQUIRE(s.algo != TACO_HIP_MFMA, "spmm: MFMA schedule needs bf16 A/B");
    rc = run_spmm<float, float>(a, K, (const float*)bv.vals, (float*)cv.vals,
                                s);
  } else if (a.vt == VType::BF16 && bv.vt == VType::BF16 &&
             cv.vt == VType::F32) {
    if (s.algo == TACO_HIP_MFMA)
      rc = spmm_mfma_bf16(a, K, bv.vals, (float*)cv.vals, s);
    else
      rc = run_spmm<bf16_t, float>(a, K, (const bf16_t*)bv.vals,
                                   (float*)cv.vals, s);
  } else {
    return fail("spmm: supported value forms are f64, f32, and bf16 A/B "
                "with f32 C");
  }
  if (rc) return rc;
  return finish("taco_hip_spmm_compute");
}
