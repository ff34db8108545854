// Hand-written HIP SpMV kernels, y(i) = A(i,j) * x(j), A in CSR.
//
// Four schedules, each the wave64 form of a schedule family taco emits or
// that the power-law workload needs:
//   THREAD_PER_ROW  taco's default GPU schedule, split(i,256) -> GPUBlock /
//                   GPUThread (reference src/index_notation/
//                   transformations.cpp:1585-1596): the generated body
//                   `for (jA = pos[i]; jA < pos[i+1]; jA++) y += A*x[crd]`.
//   GROUP_PER_ROW   scheduleSpMVRowsGPU (test/tests-scheduling-eval.cpp:
//                   211-223): a lane group strides one row's entries and the
//                   group reduces (GPUWarpReduction) — here a shuffle tree
//                   within the 64-lane wave instead of lane 0's serial loop
//                   (codegen_cuda.cpp:875-890).
//   NNZ_SPLIT       scheduleSpMVGPU / scheduleSpMVSplitPosGPU (tests-
//                   scheduling-eval.cpp:193-209, 233-247): fuse(i,j) -> pos ->
//                   split into fixed entry chunks per thread, row found by
//                   binary search (taco_binarySearchBefore, codegen_cuda.cpp:
//                   89-109), atomics at row changes, wave-aggregated when the
//                   whole wave is in one row (atomicAddWarp, codegen_cuda.cpp:
//                   142-159, here over 64 lanes).
//   MERGE_PATH      merge-path decomposition of (row ends, entries): equal
//                   work per thread regardless of row lengths; within a
//                   workgroup the partial rows are combined by a segmented
//                   scan; rows crossing workgroups are fixed up in order from
//                   a carry array by a second kernel.  No atomics, so results
//                   are deterministic run to run.

#include <climits>
#include <mutex>

#include "common.h"
#include "device.h"

namespace taco_hip {
namespace {

template <typename T>
__global__ __launch_bounds__(1024) void k_spmv_thread_per_row(
    int32_t m, const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const T* __restrict__ vals, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  T acc = 0;
  const int32_t end = pos[i + 1];
  for (int32_t p = pos[i]; p < end; ++p) acc += vals[p] * x[crd[p]];
  y[i] = acc;
}

template <typename T, int G>
__global__ __launch_bounds__(1024) void k_spmv_group_per_row(
    int32_t m, int32_t rows_per_group, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ crd, const T* __restrict__ vals,
    const T* __restrict__ x, T* __restrict__ y) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t group = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int64_t first = group * rows_per_group;
  for (int32_t rr = 0; rr < rows_per_group; ++rr) {
    const int64_t i = first + rr;
    if (i >= m) break;  // uniform across the group
    T acc = 0;
    const int32_t end = pos[i + 1];
    for (int32_t p = pos[i] + lane; p < end; p += G)
      acc += stream_load(vals + p) * x[stream_load(crd + p)];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (lane == 0) y[i] = acc;
  }
}

// Largest r in [0, m) with pos[r] <= z (the row holding entry z < nnz).
__device__ __forceinline__ int32_t row_of_entry(const int32_t* pos, int32_t m,
                                                int64_t z) {
  int32_t lo = 0, hi = m;  // pos[lo] <= z < pos[hi]
  while (hi - lo > 1) {
    int32_t mid = lo + (hi - lo) / 2;
    if ((int64_t)pos[mid] <= z)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename T, int K>
__global__ __launch_bounds__(1024) void k_spmv_nnz_split(
    int32_t m, int64_t nnz, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ crd, const T* __restrict__ vals,
    const T* __restrict__ x, T* __restrict__ y, bool coarse) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t z0 = t * K;
  int32_t r = 0;
  T acc = 0;
  const bool active = z0 < nnz;
  if (active) {
    const int64_t z1 = z0 + K < nnz ? z0 + K : nnz;
    r = row_of_entry(pos, m, z0);
    int64_t rend = pos[r + 1];
    for (int64_t z = z0; z < z1; ++z) {
      while (z >= rend) {  // skips empty rows too
        if (acc != T(0)) atomic_add(y + r, acc, coarse);
        acc = 0;
        ++r;
        rend = pos[r + 1];
      }
      acc += stream_load(vals + z) * x[stream_load(crd + z)];
    }
  }
  // Wave-aggregated flush: one atomic when every active lane ends in the
  // same row, otherwise one per lane.
  const uint64_t act = __ballot(active);
  const int leader = act ? __ffsll((unsigned long long)act) - 1 : 0;
  const int32_t r_lead = __shfl(r, leader, kWave);
  const bool same = !active || r == r_lead;
  if (__ballot(same) == ~0ull) {
    T s = wave_sum(active ? acc : T(0));
    if ((int)__lane_id() == leader && act) atomic_add(y + r_lead, s, coarse);
  } else if (active) {
    atomic_add(y + r, acc, coarse);
  }
}

// Merge-path SpMV: one workgroup of 256 threads per IPW = 256*K merge items.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_spmv_merge(
    int32_t m, int64_t nnz, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ crd, const T* __restrict__ vals,
    const T* __restrict__ x, T* __restrict__ y, int32_t* __restrict__ carry_row,
    T* __restrict__ carry_val) {
  constexpr int B = 256;
  constexpr int IPW = B * K;
  constexpr int NW = B / kWave;
  __shared__ T s_prod[IPW];
  __shared__ int32_t s_row_end[IPW + 1];
  __shared__ int64_t s_coord[2];
  __shared__ int32_t s_wlast[NW], s_wfirst[NW];
  __shared__ T s_wval[NW];

  const int tid = threadIdx.x;
  const int wave = tid / kWave;
  const int lane = tid % kWave;
  const int64_t total = (int64_t)m + nnz;
  const int64_t d0 = (int64_t)blockIdx.x * IPW;
  const int64_t d1 = d0 + IPW < total ? d0 + IPW : total;

  // Tile coordinates: wave 0 finds the start split, wave 1 the end split.
  if (wave < 2) {
    const int64_t r = wave_merge_search(pos, m, nnz, wave == 0 ? d0 : d1);
    if (lane == 0) s_coord[wave] = r;
  }
  __syncthreads();
  const int64_t r0 = s_coord[0], r1 = s_coord[1];
  const int64_t z0 = d0 - r0;
  const int nrows = (int)(r1 - r0);       // row ends inside the tile
  const int nz = (int)((d1 - r1) - z0);   // entries inside the tile

  // Stage the tile's row ends (plus the end of the row in progress at the
  // tile end) and the products A_z * x[crd_z], coalesced.
  for (int q = tid; q <= nrows; q += B) {
    const int64_t rr = r0 + 1 + q;
    s_row_end[q] = rr <= m ? pos[rr] : INT_MAX;
  }
  {
    int32_t c[K];
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int q = tid + k * B;
      c[k] = 0;
      v[k] = 0;
      if (q < nz) {
        c[k] = stream_load(crd + z0 + q);
        v[k] = stream_load(vals + z0 + q);
      }
    }
    T xv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) xv[k] = (tid + k * B < nz) ? x[c[k]] : T(0);
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (tid + k * B < nz) s_prod[tid + k * B] = v[k] * xv[k];
  }
  __syncthreads();

  // This thread's start on the tile-local merge path.
  const int dl = tid * K;
  const int tile_items = nrows + nz;
  int rl, zl;
  if (dl >= tile_items) {
    rl = nrows;
    zl = nz;
  } else {
    int lo = dl - nz > 0 ? dl - nz : 0;
    int hi = dl < nrows ? dl : nrows;
    while (lo < hi) {
      const int p = (lo + hi) >> 1;
      if ((int64_t)s_row_end[p] <= z0 + dl - p - 1)
        lo = p + 1;
      else
        hi = p;
    }
    rl = lo;
    zl = dl - lo;
  }

  // Consume K merge items.  Completed rows after the first are final within
  // this thread; the first completed row still lacks the partial sums of the
  // threads before it, added after the scan.
  T acc = 0;
  bool have_first = false;
  int first_rl = 0;
  T first_val = 0;
  const int items = tile_items - dl < K ? (tile_items - dl > 0 ? tile_items - dl : 0) : K;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < items) {
      if (z0 + zl < (int64_t)s_row_end[rl]) {
        acc += s_prod[zl];
        ++zl;
      } else {
        if (!have_first) {
          have_first = true;
          first_rl = rl;
          first_val = acc;
        } else {
          y[r0 + rl] = acc;
        }
        acc = 0;
        ++rl;
      }
    }
  }

  // Segmented inclusive scan of the carries (row in progress, partial sum)
  // over the workgroup, keyed by row.
  T inc = wave_segmented_inclusive_sum(acc, rl);
  if (lane == kWave - 1) {
    s_wlast[wave] = rl;
    s_wval[wave] = inc;
  }
  if (lane == 0) s_wfirst[wave] = rl;
  __syncthreads();
  int cin_row = -1;
  T cin = 0;
  if (wave > 0) {
    cin_row = s_wlast[wave - 1];
    cin = s_wval[wave - 1];
    for (int w2 = wave - 2; w2 >= 0; --w2) {
      if (s_wfirst[w2 + 1] == cin_row && s_wlast[w2] == cin_row)
        cin += s_wval[w2];
      else
        break;
    }
    if (rl == cin_row) inc += cin;
  }
  T prev_inc = __shfl_up(inc, 1, kWave);
  int prev_row = __shfl_up(rl, 1, kWave);
  if (lane == 0) {
    prev_inc = cin;
    prev_row = cin_row;
  }
  if (have_first) {
    T v = first_val;
    if (prev_row == first_rl) v += prev_inc;
    y[r0 + first_rl] = v;
  }
  if (tid == B - 1) {
    carry_row[blockIdx.x] = (int32_t)(r0 + rl);
    carry_val[blockIdx.x] = inc;
  }
}

// Adds each row's cross-workgroup partial sums, in workgroup order.
template <typename T>
__global__ __launch_bounds__(256) void k_spmv_merge_fixup(
    int32_t nblocks, int32_t m, const int32_t* __restrict__ carry_row,
    const T* __restrict__ carry_val, T* __restrict__ y) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const int32_t r = carry_row[b];
  if (r >= m) return;
  if (b > 0 && carry_row[b - 1] == r) return;  // not the first of its run
  T s = carry_val[b];
  for (int32_t c = b + 1; c < nblocks && carry_row[c] == r; ++c)
    s += carry_val[c];
  y[r] += s;
}

// ext_scratch / ext_bytes: caller-provided scratch for the merge path's
// carries (the no-allocation entry point); nullptr: the library workspace.
template <typename T>
int run_spmv(const CsrView& a, const T* x, T* y, const taco_hip_schedule& s,
             void* ext_scratch = nullptr, size_t ext_bytes = 0) {
  hipStream_t st = current_stream();
  if (a.m == 0) return 0;
  const int algo = s.algo;
  if (algo == TACO_HIP_THREAD_PER_ROW) {
    int bs = s.block_size > 0 ? s.block_size : 256;
    TH_REQUIRE(bs % kWave == 0 && bs <= 1024, "spmv: bad block size %d", bs);
    unsigned grid = (unsigned)((a.m + bs - 1) / bs);
    k_spmv_thread_per_row<T><<<grid, bs, 0, st>>>(
        a.m, a.pos, a.crd, (const T*)a.vals, x, y);
    return 0;
  }
  if (algo == TACO_HIP_GROUP_PER_ROW) {
    int bs = s.block_size > 0 ? s.block_size : 256;
    int g = s.group_size > 0 ? s.group_size : 8;
    int rpg = s.rows_per_group > 0 ? s.rows_per_group : 1;
    TH_REQUIRE(bs % kWave == 0 && bs <= 1024, "spmv: bad block size %d", bs);
    int64_t groups = ((int64_t)a.m + rpg - 1) / rpg;
    int64_t threads = groups * g;
    unsigned grid = (unsigned)((threads + bs - 1) / bs);
#define TH_GPR(G)                                                          \
  case G:                                                                  \
    k_spmv_group_per_row<T, G><<<grid, bs, 0, st>>>(                       \
        a.m, rpg, a.pos, a.crd, (const T*)a.vals, x, y);                   \
    return 0;
    switch (g) {
      TH_GPR(1) TH_GPR(2) TH_GPR(4) TH_GPR(8) TH_GPR(16) TH_GPR(32) TH_GPR(64)
      default:
        return fail("spmv: group_size %d not a power of two <= 64", g);
    }
#undef TH_GPR
  }
  if (algo == TACO_HIP_NNZ_SPLIT) {
    int bs = s.block_size > 0 ? s.block_size : 256;
    int k = s.items_per_thread > 0 ? s.items_per_thread : 8;
    TH_REQUIRE(bs % kWave == 0 && bs <= 1024, "spmv: bad block size %d", bs);
    if (int rc = zero_fill(y, sizeof(T) * (size_t)a.m, st)) return rc;
    if (a.nnz == 0) return 0;
    int64_t threads = (a.nnz + k - 1) / k;
    unsigned grid = (unsigned)((threads + bs - 1) / bs);
    const bool coarse = coarse_grained(y);
#define TH_NS(K)                                                           \
  case K:                                                                  \
    k_spmv_nnz_split<T, K><<<grid, bs, 0, st>>>(a.m, a.nnz, a.pos, a.crd,  \
                                                (const T*)a.vals, x, y,    \
                                                coarse);                   \
    return 0;
    switch (k) {
      TH_NS(1) TH_NS(2) TH_NS(4) TH_NS(8) TH_NS(16) TH_NS(32)
      default:
        return fail("spmv: items_per_thread %d unsupported for NNZ_SPLIT", k);
    }
#undef TH_NS
  }
  if (algo == TACO_HIP_MERGE_PATH) {
    int k = s.items_per_thread > 0 ? s.items_per_thread : 8;
    TH_REQUIRE(s.block_size == 0 || s.block_size == 256,
               "spmv: MERGE_PATH uses 256-thread workgroups");
    const int64_t total = (int64_t)a.m + a.nnz;
    const int64_t ipw = 256 * (int64_t)k;
    const int64_t nblocks = (total + ipw - 1) / ipw;
    TH_REQUIRE(nblocks < INT32_MAX, "spmv: problem too large");
    const size_t need = (size_t)nblocks * (sizeof(int32_t) + sizeof(T)) + 64;
    void* ws = ext_scratch;
    if (ws) {
      TH_REQUIRE(need <= ext_bytes,
                 "spmv (no-allocation form): the merge path needs %zu bytes of scratch "
                 "for %lld rows + entries at %d items per thread; the limit is %zu bytes",
                 need, (long long)total, k, ext_bytes);
    } else {
      ws = workspace(need, 0);
      if (!ws) return 1;
    }
    int32_t* crow = (int32_t*)ws;
    T* cval = (T*)((char*)ws + ((size_t)nblocks * sizeof(int32_t) + 15) / 16 * 16);
#define TH_MP(K)                                                           \
  case K:                                                                  \
    k_spmv_merge<T, K><<<(unsigned)nblocks, 256, 0, st>>>(                 \
        a.m, a.nnz, a.pos, a.crd, (const T*)a.vals, x, y, crow, cval);     \
    break;
    switch (k) {
      TH_MP(2) TH_MP(4) TH_MP(6) TH_MP(8) TH_MP(12) TH_MP(16)
      default:
        return fail("spmv: items_per_thread %d unsupported for MERGE_PATH", k);
    }
#undef TH_MP
    TH_LAUNCH_CHECK("k_spmv_merge");
    unsigned g2 = (unsigned)((nblocks + 255) / 256);
    k_spmv_merge_fixup<T><<<g2, 256, 0, st>>>((int32_t)nblocks, a.m, crow,
                                               cval, y);
    return 0;
  }
  return fail("spmv: schedule algo %d not available for SpMV", algo);
}

__global__ __launch_bounds__(256) void k_zero_fill(uint32_t* __restrict__ p, int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
    p[i] = 0u;
}

}  // namespace

int zero_fill(void* p, size_t bytes, hipStream_t st) {
  TH_REQUIRE(bytes % 4 == 0 && ((uintptr_t)p % 4) == 0, "zero fill: %zu bytes at %p", bytes, p);
  const int64_t words = (int64_t)(bytes / 4);
  if (words == 0) return 0;
  const int64_t blocks = (words + 255) / 256;
  k_zero_fill<<<(unsigned)(blocks < 8192 ? blocks : 8192), 256, 0, st>>>((uint32_t*)p, words);
  TH_LAUNCH_CHECK("k_zero_fill");
  return 0;
}

}  // namespace taco_hip

using namespace taco_hip;

namespace {

// Scratch of the no-allocation entry point: the merge path's per-workgroup
// carries (row, partial sum), written by every workgroup before the fix-up
// kernel reads them, so nothing needs clearing between calls.  Each device
// holds its own copy until the process ends; host code uses the address
// hipGetSymbolAddress returns for the current device, never the symbol.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};

// The current device's scratch address, resolved once per device.
int noalloc_scratch(void** out) {
  int dev = 0;
  TH_CHECK(hipGetDevice(&dev));
  TH_REQUIRE(dev >= 0 && dev < kScratchDevices, "spmv (no-allocation form): device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TH_REQUIRE(p != nullptr, "spmv (no-allocation form): no scratch address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int spmv_entry(taco_tensor_t* y, taco_tensor_t* A, taco_tensor_t* x,
               const taco_hip_schedule* sched, bool no_alloc) {
  CsrView a;
  DenseView xv, yv;
  if (int rc = view_csr(A, "A", &a)) return rc;
  if (int rc = view_dense(x, "x", 1, &xv)) return rc;
  if (int rc = view_dense(y, "y", 1, &yv)) return rc;
  TH_REQUIRE(xv.dims[0] == a.n, "spmv: x has %d entries, A has %d columns",
             xv.dims[0], a.n);
  TH_REQUIRE(yv.dims[0] == a.m, "spmv: y has %d entries, A has %d rows",
             yv.dims[0], a.m);
  TH_REQUIRE(a.vt == xv.vt && a.vt == yv.vt,
             "spmv: A, x and y must share one value type");
  taco_hip_schedule s;
  if (sched)
    s = *sched;
  else
    taco_hip_default_schedule(0, a.m, a.nnz, 1, &s);
  void* ext = nullptr;
  if (no_alloc)
    if (int rc = noalloc_scratch(&ext)) return rc;
  const size_t ext_bytes = no_alloc ? kNoAllocScratchBytes : 0;
  int rc;
  if (a.vt == VType::F64)
    rc = run_spmv<double>(a, (const double*)xv.vals, (double*)yv.vals, s, ext, ext_bytes);
  else if (a.vt == VType::F32)
    rc = run_spmv<float>(a, (const float*)xv.vals, (float*)yv.vals, s, ext, ext_bytes);
  else
    return fail("spmv: value type must be f64 or f32");
  if (rc) return rc;
  return finish(no_alloc ? "taco_hip_spmv_compute_noalloc" : "taco_hip_spmv_compute");
}

}  // namespace

extern "C" int taco_hip_spmv_compute(taco_tensor_t* y, taco_tensor_t* A,
                                     taco_tensor_t* x,
                                     const taco_hip_schedule* sched) {
  return spmv_entry(y, A, x, sched, false);
}

extern "C" int taco_hip_spmv_compute_noalloc(taco_tensor_t* y, taco_tensor_t* A,
                                             taco_tensor_t* x,
                                             const taco_hip_schedule* sched) {
  return spmv_entry(y, A, x, sched, true);
}
