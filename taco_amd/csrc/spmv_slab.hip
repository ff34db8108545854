// Slab-tiled SpMV: an MI355X storage plan for CSR SpMV whose x does not fit
// one XCD's L2.
//
// Random gathers x[crd] over a vector larger than the 4 MB L2 of an XCD are
// served from the memory-side MALL at a fraction of the L2 hit rate, and they
// dominate SpMV on matrices with scattered columns.  The plan splits the
// columns into slabs of `slab_width` (2^17 columns = 1 MB of f64 by default)
// and the rows into P parts sized so that one part's row partial sums fit in
// LDS and all parts are resident at once.  Inside each part the entries are
// reordered by slab (stable, so rows stay ordered within a slab).  Every
// workgroup then sweeps slab 0, 1, 2, ... over its own entries; since all
// workgroups progress at the same rate, each XCD's L2 holds the slab being
// swept and the gathers hit L2.  A is still read exactly once.
//
// Partition: a weighted merge path over (row ends, entries) in which a row
// end weighs `row_weight` entries; part p starts at (r_p, z_p) with
// pos[r_p] <= z_p <= pos[r_p + 1], owns local rows r_p .. r_{p+1}, completes
// rows r_p .. r_{p+1}-1 (writing y) and carries the partial sum of row
// r_{p+1}; carries are added in part order by a fix-up kernel, as in the
// merge-path kernel.
//
// Packed entry: (local row << 17) | (column - slab * slab_width), so an
// entry stays 4 + 8 bytes.  Each (part, slab) segment starts at a multiple
// of kAlign entries (the gaps hold zeroed filler that is never summed).
// Inside a segment the entries are ordered by column (a stable segmented
// radix sort on the column bits, so a column's entries keep their row
// order): the lanes of a wave then gather nearby columns of x, so a wave's
// gathers touch fewer cache lines and consecutive waves of a workgroup reuse
// L1 lines.  On the 2M-row power-law matrix that took the main kernel from
// 235 to 204 us (512 parts, 4 consecutive entries per thread).
//
// Default shape: one part per CU, whose row partials may take up to 156 KB
// of the CU's 160 KB of LDS, as one workgroup of 512 threads.  Twice the
// entries per (part, slab) segment as with two parts per CU means twice
// the entries per line of x under a wave's load: 171 us on the power-law
// matrix, against 195 us for 512 parts of 256 threads (sweeps over
// 256..1024 parts, 1..16 entries per thread, 256..1024 threads, 0.5..2 MB
// slabs).
//
// Three forms of the main kernel accumulate the row partials in LDS:
//   k_slab_spmv_il (default, 8 entries per thread): the workgroup takes
//     8·T consecutive entries at a time and thread t takes entries t, t + T,
//     ..., so one load instruction of a wave covers 64 consecutive entries
//     of the column-ordered segment and lanes whose columns share a line of
//     x are served by one L2 request; each product is added to its row with
//     an LDS atomic (171 us on the power-law matrix);
//   k_slab_spmv: a thread takes U consecutive entries (vector loads), sums
//     the runs of one row among them and adds each run with an LDS atomic —
//     with column-ordered segments nearly every entry is its own run (205 us
//     with 4 entries per thread);
//   k_slab_spmv_wave: a wave takes blocks of 64 consecutive entries; a
//     segmented wave sum by row (DPP row shifts and broadcasts) leaves one
//     total per row run, added by one lane.  It pays off when segments keep
//     rows contiguous (235 us on the power-law matrix with row-ordered
//     segments, against 237 us for k_slab_spmv on the same layout).
// The per-entry LDS atomics are not what bounds the kernel: the x gathers
// are (the L2 request rate and the misses among them).
//
// Row partial sums accumulate with LDS float atomics, so the order of the
// additions inside a row is not fixed: results agree with the sequential sum
// to rounding (the value tolerance) but may differ in the last bits from run
// to run.  The MERGE_PATH schedule is the deterministic one.

#include <climits>
#include <vector>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "common.h"
#include "device.h"

namespace taco_hip {

namespace {

constexpr int kColBits = 17;
constexpr uint32_t kColMask = (1u << kColBits) - 1;
// Workgroup size of the main kernel: 512 threads when each CU holds one
// part, 256 when it holds several.
constexpr int kPlanThreadsOnePerCu = 512;
constexpr int kPlanThreads = 256;
constexpr int kDefaultUnroll = 4;     // k_slab_spmv, k_slab_spmv_wave
constexpr int kDefaultUnrollIl = 8;   // k_slab_spmv_il
// Default parts: one per CU, more (a multiple of the CU count) when the
// rows would otherwise crowd the LDS (parts of at most half its capacity
// on average, so that entry-balanced parts stay within it).
// 2^17 columns = 1 MB of f64 x per slab (a quarter of an XCD's L2).
constexpr int kDefaultSlabWidth = 1 << 17;
// Segment start alignment (entries); the largest per-thread chunk.
constexpr int kAlign = 8;
// Row partials live in dynamic LDS sized by the plan's largest part; this
// caps one part at 156 KB of f64 partials (a CU has 160 KB of LDS).
constexpr int kMaxLocalRows = 19968;

// Largest r in [0, m] with w*r + pos[r] <= d.
__device__ int64_t weighted_row(const int32_t* pos, int64_t m, int64_t w,
                                int64_t d) {
  int64_t lo = 0, hi = m;  // f(lo) <= d holds for lo = 0 when d >= 0
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (w * mid + (int64_t)pos[mid] <= d) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_slab_partition(
    int32_t m, int64_t nnz, int64_t w, int64_t per_part, int32_t parts,
    const int32_t* __restrict__ pos, int32_t* __restrict__ part_row,
    int32_t* __restrict__ part_entry) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > parts) return;
  const int64_t total = w * (int64_t)m + nnz;
  int64_t d = p * per_part;
  if (d > total || p == parts) d = total;
  // Part 0 starts at row 0 even when leading rows are empty (with w = 0 the
  // search would otherwise skip them).
  int64_t r = p == 0 ? 0 : weighted_row(pos, m, w, d);
  int64_t z = d - w * r;
  if (r < m) {
    if (z > pos[r + 1]) z = pos[r + 1];
  } else {
    r = m;
    z = nnz;
  }
  if (z < pos[r]) z = pos[r];
  part_row[p] = (int32_t)r;
  part_entry[p] = (int32_t)z;
}

// Per part: entries per slab (into cnt) and the part's padded size.
__global__ __launch_bounds__(256) void k_slab_counts(
    int32_t parts, int32_t slabs, int32_t slab_width,
    const int32_t* __restrict__ crd, const int32_t* __restrict__ part_entry,
    int32_t* __restrict__ cnt, int32_t* __restrict__ part_size) {
  __shared__ int32_t s_count[1024];
  const int32_t p = blockIdx.x;
  for (int s = threadIdx.x; s < slabs; s += blockDim.x) s_count[s] = 0;
  __syncthreads();
  const int32_t z0 = part_entry[p], z1 = part_entry[p + 1];
  for (int32_t z = z0 + threadIdx.x; z < z1; z += blockDim.x)
    atomicAdd(&s_count[crd[z] / slab_width], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int s = 0; s < slabs; ++s) {
      cnt[(int64_t)p * slabs + s] = s_count[s];
      acc += (s_count[s] + kAlign - 1) / kAlign * kAlign;
    }
    part_size[p] = acc > INT32_MAX ? INT32_MAX : (int32_t)acc;
  }
}

// Per part: aligned segment starts off[p][0..slabs] from the part's base
// (exclusive scan of the padded part sizes) and segment ends seg_end[p][s].
__global__ __launch_bounds__(256) void k_slab_layout(
    int32_t parts, int32_t slabs, const int32_t* __restrict__ part_base,
    const int32_t* __restrict__ cnt, int32_t* __restrict__ off,
    int32_t* __restrict__ seg_end) {
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= parts) return;
  int32_t at = part_base[p];
  int32_t* o = off + (int64_t)p * (slabs + 1);
  for (int s = 0; s < slabs; ++s) {
    const int32_t c = cnt[(int64_t)p * slabs + s];
    o[s] = at;
    seg_end[(int64_t)p * slabs + s] = at + c;
    at += (c + kAlign - 1) / kAlign * kAlign;
  }
  o[slabs] = at;
}

__global__ __launch_bounds__(256) void k_slab_scatter(
    int32_t slabs, int32_t slab_width, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ crd, const double* __restrict__ vals,
    const int32_t* __restrict__ part_row, const int32_t* __restrict__ part_entry,
    const int32_t* __restrict__ off, uint32_t* __restrict__ ent,
    double* __restrict__ val2) {
  __shared__ int32_t s_cursor[1024];
  __shared__ int32_t s_wave[4][1024];
  const int32_t p = blockIdx.x;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const int32_t* o = off + (int64_t)p * (slabs + 1);
  for (int s = tid; s < slabs; s += blockDim.x) s_cursor[s] = o[s];
  const int32_t r0 = part_row[p], r1 = part_row[p + 1];
  const int32_t z0 = part_entry[p], z1 = part_entry[p + 1];
  __syncthreads();
  for (int32_t base = z0; base < z1; base += 256) {
    const int32_t z = base + tid;
    const bool valid = z < z1;
    int32_t s = -1, col = 0, lrow = 0;
    double v = 0;
    if (valid) {
      col = crd[z];
      v = vals[z];
      s = col / slab_width;
      // local row: the row of entry z within [r0, r1]
      int32_t lo = r0, hi = r1;  // pos[lo] <= z, answer in [r0, r1]
      while (lo < hi) {
        const int32_t mid = lo + (hi - lo + 1) / 2;
        if (pos[mid] <= z) lo = mid; else hi = mid - 1;
      }
      lrow = lo - r0;
    }
    // Rank within this chunk for each lane's slab.
    for (int q = tid; q < 4 * slabs; q += blockDim.x) s_wave[q / slabs][q % slabs] = 0;
    __syncthreads();
    int rank_in_wave = 0;
    {
      // lanes with the same slab in this wave, below this lane
      uint64_t below = (1ull << lane) - 1;
      for (int ss = 0; ss < slabs; ++ss) {
        const uint64_t mask = __ballot(s == ss);
        if (s == ss) rank_in_wave = __popcll(mask & below);
        if (lane == 0) s_wave[wave][ss] = __popcll(mask);
      }
    }
    __syncthreads();
    if (valid) {
      int32_t before = 0;
      for (int w2 = 0; w2 < wave; ++w2) before += s_wave[w2][s];
      const int32_t dst = s_cursor[s] + before + rank_in_wave;
      ent[dst] = ((uint32_t)lrow << kColBits) | (uint32_t)(col - s * slab_width);
      val2[dst] = v;
    }
    __syncthreads();
    for (int ss = tid; ss < slabs; ss += blockDim.x)
      s_cursor[ss] += s_wave[0][ss] + s_wave[1][ss] + s_wave[2][ss] + s_wave[3][ss];
    __syncthreads();
  }
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup p runs on
// XCD p % 8).  Every workgroup of one XCD starts its sweep at the same slab
// — so its L2 serves one slab at a time — and different XCDs start at
// different slabs, which spreads their concurrent misses over the memory
// side instead of all XCDs fetching the same slab at once.
constexpr int kXcds = 8;

int device_cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      n <= 0)
    return 256;
  return n;
}
__device__ __forceinline__ int32_t first_slab(int32_t part, int32_t slabs) {
  return (int32_t)((int64_t)(part % kXcds) * slabs / kXcds);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Loads U consecutive packed entries and values starting at an index that is
// a multiple of U (16-byte vector loads for U >= 2).
template <int U>
__device__ __forceinline__ void load_chunk(const uint32_t* __restrict__ ent,
                                           const double* __restrict__ val2,
                                           int64_t e, uint32_t (&pk)[U], double (&v)[U]) {
  if constexpr (U == 1) {
    pk[0] = stream_load(ent + e);
    v[0] = stream_load(val2 + e);
  } else if constexpr (U == 2) {
    const f64x2 a = __builtin_nontemporal_load((const f64x2*)(val2 + e));
    pk[0] = stream_load(ent + e);
    pk[1] = stream_load(ent + e + 1);
    v[0] = a.x;
    v[1] = a.y;
  } else {
#pragma unroll
    for (int q = 0; q < U / 4; ++q) {
      const u32x4 a = __builtin_nontemporal_load((const u32x4*)(ent + e) + q);
      pk[4 * q + 0] = a.x;
      pk[4 * q + 1] = a.y;
      pk[4 * q + 2] = a.z;
      pk[4 * q + 3] = a.w;
    }
#pragma unroll
    for (int q = 0; q < U / 2; ++q) {
      const f64x2 a = __builtin_nontemporal_load((const f64x2*)(val2 + e) + q);
      v[2 * q + 0] = a.x;
      v[2 * q + 1] = a.y;
    }
  }
}

// One workgroup per part.  Each thread takes U consecutive entries of the
// current slab segment; entries of one row are adjacent in a segment, so
// the thread sums each run of equal rows in a register and adds the run to
// the row's LDS partial with one atomic.
template <int U, int T>
__global__ __launch_bounds__(T) void k_slab_spmv(
    int32_t m, int32_t slabs, int32_t slab_width,
    const int32_t* __restrict__ part_row, const int32_t* __restrict__ off,
    const int32_t* __restrict__ seg_end,
    const uint32_t* __restrict__ ent, const double* __restrict__ val2,
    const double* __restrict__ x, double* __restrict__ y,
    int32_t* __restrict__ carry_row, double* __restrict__ carry_val) {
  extern __shared__ double s_part[];
  const int32_t p = blockIdx.x;
  const int tid = threadIdx.x;
  const int32_t r0 = part_row[p], r1 = part_row[p + 1];
  const int32_t nrows = r1 - r0 + 1;  // local rows, <= kMaxLocalRows by plan
  for (int32_t q = tid; q < nrows; q += T) s_part[q] = 0;
  __syncthreads();
  const int32_t* o = off + (int64_t)p * (slabs + 1);
  const int32_t* se = seg_end + (int64_t)p * slabs;
  const int32_t s_first = first_slab(p, slabs);
  for (int32_t s0 = 0; s0 < slabs; ++s0) {
    const int32_t s = s0 + s_first < slabs ? s0 + s_first : s0 + s_first - slabs;
    const int32_t e0 = o[s], e1 = se[s];
    const int32_t nchunks = (e1 - e0 + U - 1) / U;
    const double* xs = x + (int64_t)s * slab_width;
    for (int32_t c0 = 0; c0 < nchunks; c0 += T) {
      const int32_t c = c0 + tid;
      // Threads past the last chunk re-read it (a valid, aligned address)
      // and sum nothing, so every thread issues the same loads.
      const int32_t e = e0 + (c < nchunks ? c : nchunks - 1) * U;
      uint32_t pk[U];
      double v[U];
      load_chunk<U>(ent, val2, e, pk, v);
      double xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = xs[pk[u] & kColMask];
      const int32_t live = c < nchunks ? e1 - e : 0;  // entries of this thread
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (u < live) {
          acc = fma(v[u], xv[u], acc);
          const uint32_t row = pk[u] >> kColBits;
          const bool last = u == U - 1 || u + 1 >= live || (pk[u + 1] >> kColBits) != row;
          if (last) {
            atomicAdd(&s_part[row], acc);
            acc = 0.0;
          }
        }
      }
    }
  }
  __syncthreads();
  // Rows r0 .. r1-1 complete here; row r1 is carried.
  for (int32_t q = tid; q < nrows - 1; q += T) y[r0 + q] = s_part[q];
  if (tid == 0) {
    carry_row[p] = r1;
    carry_val[p] = s_part[nrows - 1];
  }
}

// One workgroup per part.  The workgroup takes U·T consecutive entries of
// the current slab segment at a time and thread t takes entries t, t + T,
// ..., so one load instruction of a wave covers 64 consecutive entries: in a
// column-ordered segment, lanes whose columns share a cache line of x are
// served by one L2 request.  Each entry's product goes to its row's LDS
// partial with one atomic.
template <int U, int T>
__global__ __launch_bounds__(T) void k_slab_spmv_il(
    int32_t m, int32_t slabs, int32_t slab_width,
    const int32_t* __restrict__ part_row, const int32_t* __restrict__ off,
    const int32_t* __restrict__ seg_end,
    const uint32_t* __restrict__ ent, const double* __restrict__ val2,
    const double* __restrict__ x, double* __restrict__ y,
    int32_t* __restrict__ carry_row, double* __restrict__ carry_val) {
  extern __shared__ double s_part[];
  const int32_t p = blockIdx.x;
  const int tid = threadIdx.x;
  const int32_t r0 = part_row[p], r1 = part_row[p + 1];
  const int32_t nrows = r1 - r0 + 1;  // local rows, <= kMaxLocalRows by plan
  for (int32_t q = tid; q < nrows; q += T) s_part[q] = 0;
  __syncthreads();
  const int32_t* o = off + (int64_t)p * (slabs + 1);
  const int32_t* se = seg_end + (int64_t)p * slabs;
  const int32_t s_first = first_slab(p, slabs);
  for (int32_t s0 = 0; s0 < slabs; ++s0) {
    const int32_t s = s0 + s_first < slabs ? s0 + s_first : s0 + s_first - slabs;
    const int32_t e0 = o[s], e1 = se[s];
    const double* xs = x + (int64_t)s * slab_width;
    for (int32_t b = e0; b < e1; b += U * T) {
      // Threads past the segment end re-read its last entry (a valid
      // address) and add nothing, so every thread issues the same loads.
      uint32_t pk[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t e = b + u * T + tid;
        const int32_t ec = e < e1 ? e : e1 - 1;
        pk[u] = stream_load(ent + ec);
        v[u] = stream_load(val2 + ec);
      }
      double xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = xs[pk[u] & kColMask];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (b + u * T + tid < e1) atomicAdd(&s_part[pk[u] >> kColBits], v[u] * xv[u]);
    }
  }
  __syncthreads();
  // Rows r0 .. r1-1 complete here; row r1 is carried.
  for (int32_t q = tid; q < nrows - 1; q += T) y[r0 + q] = s_part[q];
  if (tid == 0) {
    carry_row[p] = r1;
    carry_val[p] = s_part[nrows - 1];
  }
}

// f64 DPP move: lanes whose source is invalid or outside row_mask read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// Segmented inclusive sum over the wave: lane l gets the sum of lanes
// ss .. l, where ss <= l is the first lane of l's segment.  Row shifts by
// 1, 2, 4, 8 within each 16-lane row, then the row broadcasts of lanes 15
// (into rows 1 and 3), 47 (into row 3) and 31 (into rows 2 and 3).
__device__ __forceinline__ double wave_segmented_sum(double v, int lane, int ss) {
  double t;
  t = dpp_f64<0x111, 0xf>(v);  // row_shr:1
  if (lane - 1 >= ss) v += t;
  t = dpp_f64<0x112, 0xf>(v);  // row_shr:2
  if (lane - 2 >= ss) v += t;
  t = dpp_f64<0x114, 0xf>(v);  // row_shr:4
  if (lane - 4 >= ss) v += t;
  t = dpp_f64<0x118, 0xf>(v);  // row_shr:8
  if (lane - 8 >= ss) v += t;
  t = dpp_f64<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3 read lanes 15, 47
  if ((lane & 16) && (lane & ~15) - 1 >= ss) v += t;
  t = dpp_f64<0x143, 0xc>(v);  // row_bcast:31 -> rows 2, 3 read lane 31
  if ((lane & 32) && 31 >= ss) v += t;
  return v;
}

// Wave-reduced variant: each wave takes U blocks of 64 consecutive entries
// per step (lane l holds entry l of a block).  Entries of one row are
// adjacent, so a segmented wave sum by row leaves each row's block total in
// its last lane, which alone adds it to the LDS partial: a row spanning
// the whole block costs one LDS atomic instead of 64 to one address.
template <int U, int T>
__global__ __launch_bounds__(T) void k_slab_spmv_wave(
    int32_t m, int32_t slabs, int32_t slab_width,
    const int32_t* __restrict__ part_row, const int32_t* __restrict__ off,
    const int32_t* __restrict__ seg_end,
    const uint32_t* __restrict__ ent, const double* __restrict__ val2,
    const double* __restrict__ x, double* __restrict__ y,
    int32_t* __restrict__ carry_row, double* __restrict__ carry_val) {
  extern __shared__ double s_part[];
  constexpr int kWaves = T / kWave;
  const int32_t p = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid % kWave, wave = tid / kWave;
  const int32_t r0 = part_row[p], r1 = part_row[p + 1];
  const int32_t nrows = r1 - r0 + 1;  // local rows, <= kMaxLocalRows by plan
  for (int32_t q = tid; q < nrows; q += T) s_part[q] = 0;
  __syncthreads();
  const int32_t* o = off + (int64_t)p * (slabs + 1);
  const int32_t* se = seg_end + (int64_t)p * slabs;
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const int32_t s_first = first_slab(p, slabs);
  for (int32_t s0 = 0; s0 < slabs; ++s0) {
    const int32_t s = s0 + s_first < slabs ? s0 + s_first : s0 + s_first - slabs;
    const int32_t e0 = o[s], e1 = se[s];
    const double* xs = x + (int64_t)s * slab_width;
    for (int32_t base = e0; base < e1; base += T * U) {
      uint32_t pk[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t e = base + (u * kWaves + wave) * kWave + lane;
        const int32_t ec = e < e1 ? e : e1 - 1;  // valid address; summed as 0
        pk[u] = stream_load(ent + ec);
        v[u] = stream_load(val2 + ec);
      }
      double xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) xv[u] = xs[pk[u] & kColMask];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t e = base + (u * kWaves + wave) * kWave + lane;
        const bool live = e < e1;
        const uint32_t row = pk[u] >> kColBits;
        const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(
            -1, (int)row, 0x138, 0xf, 0xf, false);  // wave_shr:1, lane 0 reads -1
        const bool head = !live || lane == 0 || prev != row;
        const uint64_t heads = __ballot(head);
        const int ss = 63 - __clzll(heads & upto);
        const double sum = wave_segmented_sum(live ? v[u] * xv[u] : 0.0, lane, ss);
        const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1);
        if (live && tail) atomicAdd(&s_part[row], sum);
      }
    }
  }
  __syncthreads();
  // Rows r0 .. r1-1 complete here; row r1 is carried.
  for (int32_t q = tid; q < nrows - 1; q += T) y[r0 + q] = s_part[q];
  if (tid == 0) {
    carry_row[p] = r1;
    carry_val[p] = s_part[nrows - 1];
  }
}

__global__ __launch_bounds__(256) void k_slab_fixup(
    int32_t parts, int32_t m, const int32_t* __restrict__ carry_row,
    const double* __restrict__ carry_val, double* __restrict__ y) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= parts) return;
  const int32_t r = carry_row[b];
  if (r >= m) return;
  if (b > 0 && carry_row[b - 1] == r) return;
  double s = carry_val[b];
  for (int32_t c = b + 1; c < parts && carry_row[c] == r; ++c) s += carry_val[c];
  y[r] += s;
}

__global__ void k_max_rows(int32_t parts, const int32_t* __restrict__ part_row,
                           int32_t* __restrict__ out) {
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= parts) return;
  atomicMax(out, part_row[p + 1] - part_row[p] + 1);
}

}  // namespace
}  // namespace taco_hip

using namespace taco_hip;

struct taco_hip_spmv_plan {
  int32_t m = 0, n = 0;
  int64_t nnz = 0;
  int32_t parts = 0, slabs = 0, slab_width = 0;
  int64_t row_weight = 0;
  int32_t* part_row = nullptr;    // parts + 1
  int32_t* part_entry = nullptr;  // parts + 1
  int32_t* off = nullptr;         // parts * (slabs + 1), aligned segment starts
  int32_t* seg_end = nullptr;     // parts * slabs, segment ends
  int64_t padded = 0;             // entries in the padded layout
  uint32_t* ent = nullptr;        // padded
  double* val2 = nullptr;         // padded
  int32_t* carry_row = nullptr;   // parts
  double* carry_val = nullptr;    // parts
  int32_t max_local_rows = 0;
  int32_t threads = kPlanThreads;  // workgroup size of the main kernel
  // Main kernel: k_slab_spmv (0), k_slab_spmv_wave (1), k_slab_spmv_il (2).
  int32_t form = 2;
  // Optional kernel timing: event pairs around k_slab_spmv of the next
  // `timing_cap` compute calls.
  std::vector<hipEvent_t> timing_ev;
  int32_t timing_cap = 0, timing_count = 0;
};

namespace {
void timing_release(taco_hip_spmv_plan* p) {
  for (hipEvent_t e : p->timing_ev)
    if (e) (void)hipEventDestroy(e);
  p->timing_ev.clear();
  p->timing_cap = p->timing_count = 0;
}

void plan_free(taco_hip_spmv_plan* p) {
  if (!p) return;
  timing_release(p);
  for (void* q : {(void*)p->part_row, (void*)p->part_entry, (void*)p->off,
                  (void*)p->seg_end, (void*)p->ent, (void*)p->val2,
                  (void*)p->carry_row, (void*)p->carry_val})
    if (q) (void)hipFree(q);
  delete p;
}
}  // namespace

namespace {
// Segment bounds of the (part, slab) segments, flattened p * slabs + s.
__global__ __launch_bounds__(256) void k_segment_bounds(int32_t parts, int32_t slabs,
                                                        const int32_t* __restrict__ off,
                                                        const int32_t* __restrict__ seg_end,
                                                        uint32_t* __restrict__ beg,
                                                        uint32_t* __restrict__ end) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)parts * slabs) return;
  const int64_t p = q / slabs, s = q % slabs;
  beg[q] = (uint32_t)off[p * (slabs + 1) + s];
  end[q] = (uint32_t)seg_end[q];
}
}  // namespace

namespace {
// Orders each (part, slab) segment by column (stable, so a column's entries
// keep their row order).
int sort_segments_by_column(taco_hip_spmv_plan* p, hipStream_t st) {
  if (p->padded == 0) return 0;
  const int64_t nseg = (int64_t)p->parts * p->slabs;
  uint32_t *beg = nullptr, *end = nullptr, *ent2 = nullptr;
  double* val2b = nullptr;
  void* temp = nullptr;
  int rc = 0;
  auto release = [&]() {
    for (void* q : {(void*)beg, (void*)end, (void*)ent2, (void*)val2b, temp})
      if (q) (void)hipFree(q);
  };
  if (hipMalloc((void**)&beg, sizeof(uint32_t) * nseg) ||
      hipMalloc((void**)&end, sizeof(uint32_t) * nseg) ||
      hipMalloc((void**)&ent2, sizeof(uint32_t) * (size_t)p->padded) ||
      hipMalloc((void**)&val2b, sizeof(double) * (size_t)p->padded)) {
    release();
    return fail("spmv plan: device allocation failed");
  }
  k_segment_bounds<<<(unsigned)((nseg + 255) / 256), 256, 0, st>>>(p->parts, p->slabs, p->off,
                                                                   p->seg_end, beg, end);
  // Padding between a segment's end and the next segment's start keeps its
  // filler (never summed): copy everything first, then sort the segments.
  (void)hipMemcpyAsync(ent2, p->ent, sizeof(uint32_t) * (size_t)p->padded,
                       hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(val2b, p->val2, sizeof(double) * (size_t)p->padded,
                       hipMemcpyDeviceToDevice, st);
  size_t bytes = 0;
  if (rocprim::segmented_radix_sort_pairs(nullptr, bytes, p->ent, ent2, p->val2, val2b,
                                          (unsigned)p->padded, (unsigned)nseg, beg, end, 0,
                                          kColBits, st) != hipSuccess ||
      hipMalloc(&temp, bytes ? bytes : 16) != hipSuccess ||
      rocprim::segmented_radix_sort_pairs(temp, bytes, p->ent, ent2, p->val2, val2b,
                                          (unsigned)p->padded, (unsigned)nseg, beg, end, 0,
                                          kColBits, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    rc = fail("spmv plan: column sort failed");
  } else {
    std::swap(p->ent, ent2);
    std::swap(p->val2, val2b);
  }
  release();
  return rc;
}

}  // namespace

extern "C" int taco_hip_spmv_plan_create(taco_tensor_t* A, int32_t parts,
                                         int32_t slab_width,
                                         taco_hip_spmv_plan** out) {
  TH_REQUIRE(out, "spmv plan: null output");
  *out = nullptr;
  CsrView a;
  if (int rc = view_csr(A, "A", &a)) return rc;
  TH_REQUIRE(a.vt == VType::F64, "spmv plan: f64 values only");
  if (slab_width <= 0) slab_width = kDefaultSlabWidth;
  TH_REQUIRE(slab_width <= (1 << kColBits), "spmv plan: slab width above 2^17");
  const int cus = device_cu_count();
  if (parts <= 0) {
    const int64_t per_cu_rows = (int64_t)cus * (kMaxLocalRows / 2);
    parts = (int32_t)((int64_t)cus * (((int64_t)a.m + per_cu_rows - 1) / per_cu_rows));
    if (parts < cus) parts = cus;
  }
  const int32_t slabs = a.n == 0 ? 1 : (int32_t)(((int64_t)a.n + slab_width - 1) / slab_width);
  TH_REQUIRE(slabs <= 1024, "spmv plan: more than 1024 slabs");
  // Row weight w: first w = 0, parts of equal entry counts (so every
  // workgroup sweeps the slabs at the same pace), accepted when no part
  // exceeds the LDS row capacity; otherwise the smallest power of two for
  // which (w*m + nnz) / parts bounds the rows of every part.
  const int64_t cap = kMaxLocalRows - 2;
  int64_t w_bound = 1;
  while (w_bound < (1 << 20)) {
    const int64_t per = (w_bound * (int64_t)a.m + a.nnz + parts - 1) / parts;
    if (per / w_bound + 1 <= cap) break;
    w_bound *= 2;
  }
  TH_REQUIRE((w_bound * (int64_t)a.m + a.nnz + parts - 1) / parts / w_bound + 1 <= cap,
             "spmv plan: %d parts cannot hold %d rows; use more parts", parts, a.m);
  TH_REQUIRE((w_bound * (int64_t)a.m + a.nnz + parts - 1) / parts < INT32_MAX,
             "spmv plan: problem too large");

  auto* p = new taco_hip_spmv_plan();
  p->m = a.m;
  p->n = a.n;
  p->nnz = a.nnz;
  p->parts = parts;
  p->slabs = slabs;
  p->slab_width = slab_width;
  p->threads = parts <= cus ? kPlanThreadsOnePerCu : kPlanThreads;
  hipStream_t st = current_stream();
  auto alloc = [&](void** q, size_t bytes) {
    return hipMalloc(q, bytes < 16 ? 16 : bytes);
  };
  if (alloc((void**)&p->part_row, sizeof(int32_t) * (parts + 1)) ||
      alloc((void**)&p->part_entry, sizeof(int32_t) * (parts + 1)) ||
      alloc((void**)&p->off, sizeof(int32_t) * (size_t)parts * (slabs + 1)) ||
      alloc((void**)&p->seg_end, sizeof(int32_t) * (size_t)parts * slabs) ||
      alloc((void**)&p->carry_row, sizeof(int32_t) * parts) ||
      alloc((void**)&p->carry_val, sizeof(double) * parts)) {
    plan_free(p);
    return fail("spmv plan: device allocation failed");
  }
  int32_t* maxrows = (int32_t*)workspace(sizeof(int32_t), 3);
  if (!maxrows) {
    plan_free(p);
    return 1;
  }
  int32_t host_max = 0;
  int64_t w = a.nnz > 0 ? 0 : w_bound;
  for (;;) {
    const int64_t per_part = (w * (int64_t)a.m + a.nnz + parts - 1) / parts;
    k_slab_partition<<<(parts + 1 + 255) / 256, 256, 0, st>>>(
        a.m, a.nnz, w, per_part, parts, a.pos, p->part_row, p->part_entry);
    (void)hipMemsetAsync(maxrows, 0, sizeof(int32_t), st);
    k_max_rows<<<(parts + 255) / 256, 256, 0, st>>>(parts, p->part_row, maxrows);
    if (hipMemcpyAsync(&host_max, maxrows, sizeof(int32_t), hipMemcpyDeviceToHost, st) ||
        hipStreamSynchronize(st)) {
      plan_free(p);
      return fail("spmv plan: partition failed");
    }
    if (host_max <= kMaxLocalRows || w >= w_bound) break;
    w = w == 0 ? 1 : 2 * w;
  }
  if (host_max > kMaxLocalRows) {
    plan_free(p);
    return fail("spmv plan: a part holds %d rows (max %d)", host_max, kMaxLocalRows);
  }
  p->row_weight = w;
  p->max_local_rows = host_max;
  // Layout: per-slab counts, padded part sizes, their scan, aligned starts.
  int32_t* cnt = nullptr;
  int32_t* part_size = nullptr;
  int32_t* part_base = nullptr;
  auto release_tmp = [&]() {
    for (void* q : {(void*)cnt, (void*)part_size, (void*)part_base})
      if (q) (void)hipFree(q);
  };
  if (alloc((void**)&cnt, sizeof(int32_t) * (size_t)parts * slabs) ||
      alloc((void**)&part_size, sizeof(int32_t) * parts) ||
      alloc((void**)&part_base, sizeof(int32_t) * (parts + 1))) {
    release_tmp();
    plan_free(p);
    return fail("spmv plan: device allocation failed");
  }
  k_slab_counts<<<parts, 256, 0, st>>>(parts, slabs, slab_width, a.crd, p->part_entry, cnt,
                                        part_size);
  if (int rc = taco_hip_device_exclusive_scan(part_size, parts, part_base)) {
    release_tmp();
    plan_free(p);
    return rc;
  }
  int32_t total = 0;
  if (hipMemcpyAsync(&total, part_base + parts, sizeof(int32_t), hipMemcpyDeviceToHost, st) ||
      hipStreamSynchronize(st)) {
    release_tmp();
    plan_free(p);
    return fail("spmv plan: layout scan failed");
  }
  p->padded = total;
  if (alloc((void**)&p->ent, sizeof(uint32_t) * (size_t)total) ||
      alloc((void**)&p->val2, sizeof(double) * (size_t)total)) {
    release_tmp();
    plan_free(p);
    return fail("spmv plan: device allocation failed");
  }
  // Filler entries (row 0, column 0, value 0) are loaded but never summed.
  (void)hipMemsetAsync(p->ent, 0, sizeof(uint32_t) * (size_t)total, st);
  (void)hipMemsetAsync(p->val2, 0, sizeof(double) * (size_t)total, st);
  k_slab_layout<<<(parts + 255) / 256, 256, 0, st>>>(parts, slabs, part_base, cnt, p->off,
                                                     p->seg_end);
  if (a.nnz > 0)
    k_slab_scatter<<<parts, 256, 0, st>>>(slabs, slab_width, a.pos, a.crd,
                                         (const double*)a.vals, p->part_row,
                                         p->part_entry, p->off, p->ent, p->val2);
  const hipError_t build_err = hipStreamSynchronize(st);
  release_tmp();
  if (build_err != hipSuccess) {
    plan_free(p);
    return fail("spmv plan: build kernels failed");
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    plan_free(p);
    return fail("spmv plan: build kernels failed");
  }
  if (int rc = sort_segments_by_column(p, st)) {
    plan_free(p);
    return rc;
  }
  *out = p;
  return 0;
}

extern "C" int taco_hip_spmv_plan_compute(taco_hip_spmv_plan* p,
                                          taco_tensor_t* y, taco_tensor_t* x,
                                          int32_t unroll) {
  TH_REQUIRE(p, "spmv plan: null plan");
  DenseView xv, yv;
  if (int rc = view_dense(x, "x", 1, &xv)) return rc;
  if (int rc = view_dense(y, "y", 1, &yv)) return rc;
  TH_REQUIRE(xv.vt == VType::F64 && yv.vt == VType::F64, "spmv plan: f64 x and y");
  TH_REQUIRE(xv.dims[0] == p->n && yv.dims[0] == p->m,
             "spmv plan: x/y sizes do not match the planned matrix");
  if (p->m == 0) return 0;
  hipStream_t st = current_stream();
  const double* xp = (const double*)xv.vals;
  double* yp = (double*)yv.vals;
  const size_t lds = sizeof(double) * (size_t)(p->max_local_rows > 0 ? p->max_local_rows : 1);
  const int u = unroll > 0 ? unroll : p->form == 2 ? kDefaultUnrollIl : kDefaultUnroll;
  const int th = p->threads;
  if (u != 1 && u != 2 && u != 4 && u != 8 && !(u == 16 && p->form == 2))
    return fail("spmv plan: unroll %d unsupported", unroll);
  TH_REQUIRE(th != 1024 || (p->form == 2 && u >= 4),
             "spmv plan: 1024 threads for the interleaved form with 4, 8 or 16 entries per thread");
  const bool timed = p->timing_count < p->timing_cap;
  if (timed) TH_CHECK(hipEventRecord(p->timing_ev[2 * (size_t)p->timing_count], st));
#define TH_SLAB(U, T)                                                           \
  if (u == U && th == T)                                                        \
    k_slab_spmv<U, T><<<p->parts, T, lds, st>>>(                                \
        p->m, p->slabs, p->slab_width, p->part_row, p->off, p->seg_end, p->ent, \
        p->val2, xp, yp, p->carry_row, p->carry_val);
#define TH_SLAB_W(U, T)                                                         \
  if (u == U && th == T)                                                        \
    k_slab_spmv_wave<U, T><<<p->parts, T, lds, st>>>(                           \
        p->m, p->slabs, p->slab_width, p->part_row, p->off, p->seg_end, p->ent, \
        p->val2, xp, yp, p->carry_row, p->carry_val);
#define TH_SLAB_I(U, T)                                                         \
  if (u == U && th == T)                                                        \
    k_slab_spmv_il<U, T><<<p->parts, T, lds, st>>>(                             \
        p->m, p->slabs, p->slab_width, p->part_row, p->off, p->seg_end, p->ent, \
        p->val2, xp, yp, p->carry_row, p->carry_val);
  if (p->form == 1) {
    TH_SLAB_W(1, 256) TH_SLAB_W(2, 256) TH_SLAB_W(4, 256) TH_SLAB_W(8, 256)
    TH_SLAB_W(1, 512) TH_SLAB_W(2, 512) TH_SLAB_W(4, 512) TH_SLAB_W(8, 512)
  } else if (p->form == 2) {
    TH_SLAB_I(1, 256) TH_SLAB_I(2, 256) TH_SLAB_I(4, 256) TH_SLAB_I(8, 256) TH_SLAB_I(16, 256)
    TH_SLAB_I(1, 512) TH_SLAB_I(2, 512) TH_SLAB_I(4, 512) TH_SLAB_I(8, 512) TH_SLAB_I(16, 512)
    TH_SLAB_I(4, 1024) TH_SLAB_I(8, 1024) TH_SLAB_I(16, 1024)
  } else {
    TH_SLAB(1, 256) TH_SLAB(2, 256) TH_SLAB(4, 256) TH_SLAB(8, 256)
    TH_SLAB(1, 512) TH_SLAB(2, 512) TH_SLAB(4, 512) TH_SLAB(8, 512)
  }
#undef TH_SLAB
#undef TH_SLAB_W
#undef TH_SLAB_I
  TH_LAUNCH_CHECK("k_slab_spmv");
  if (timed) TH_CHECK(hipEventRecord(p->timing_ev[2 * (size_t)p->timing_count++ + 1], st));
  k_slab_fixup<<<(p->parts + 255) / 256, 256, 0, st>>>(p->parts, p->m, p->carry_row,
                                                       p->carry_val, yp);
  return finish("taco_hip_spmv_plan_compute");
}


extern "C" void taco_hip_spmv_plan_destroy(taco_hip_spmv_plan* p) { plan_free(p); }

extern "C" int taco_hip_spmv_plan_configure(taco_hip_spmv_plan* p, int32_t threads,
                                           int32_t form) {
  TH_REQUIRE(p, "spmv plan: null plan");
  if (threads <= 0) threads = p->parts <= device_cu_count() ? kPlanThreadsOnePerCu : kPlanThreads;
  TH_REQUIRE(threads == 256 || threads == 512 || threads == 1024,
             "spmv plan: threads must be 256, 512 or 1024");
  p->threads = threads;
  TH_REQUIRE(form >= 0 && form <= 2, "spmv plan: kernel form %d unknown", form);
  p->form = form;
  return 0;
}

extern "C" int taco_hip_spmv_plan_time_kernel(taco_hip_spmv_plan* p, int32_t launches) {
  TH_REQUIRE(p && launches >= 0, "spmv plan: bad timing arguments");
  timing_release(p);
  p->timing_ev.assign(2 * (size_t)launches, nullptr);
  for (hipEvent_t& e : p->timing_ev)
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      timing_release(p);
      return fail("spmv plan: hipEventCreate failed");
    }
  p->timing_cap = launches;
  return 0;
}

extern "C" int taco_hip_spmv_plan_kernel_time(taco_hip_spmv_plan* p, double* total_ms,
                                              int64_t* launches) {
  TH_REQUIRE(p && total_ms && launches, "spmv plan: null argument");
  double sum = 0.0;
  if (p->timing_count > 0)
    TH_CHECK(hipEventSynchronize(p->timing_ev[2 * (size_t)p->timing_count - 1]));
  for (int32_t i = 0; i < p->timing_count; ++i) {
    float ms = 0.f;
    TH_CHECK(hipEventElapsedTime(&ms, p->timing_ev[2 * (size_t)i], p->timing_ev[2 * (size_t)i + 1]));
    sum += ms;
  }
  *total_ms = sum;
  *launches = p->timing_count;
  return 0;
}

extern "C" int taco_hip_spmv_plan_shape(const taco_hip_spmv_plan* p, int32_t* threads,
                                        int32_t* form, int32_t* unroll) {
  TH_REQUIRE(p && threads && form && unroll, "spmv plan: null argument");
  *threads = p->threads;
  *form = p->form;
  *unroll = p->form == 2 ? kDefaultUnrollIl : kDefaultUnroll;
  return 0;
}

extern "C" int taco_hip_spmv_plan_info(const taco_hip_spmv_plan* p, int64_t* info) {
  TH_REQUIRE(p && info, "spmv plan: null argument");
  info[0] = p->parts;
  info[1] = p->slabs;
  info[2] = p->slab_width;
  info[3] = p->row_weight;
  info[4] = p->max_local_rows;
  return 0;
}
