// Batched SpMV / SpMM over many uneven CSR instances in one persistent
// launch (no reference counterpart; taco compiles one kernel per statement
// and launches it per tensor).
//
// Storage: the instances are concatenated.  Instance i owns
//   pos[pos_off_i .. pos_off_i + m_i]           (local entry offsets, int32)
//   crd/vals[ent_off_i .. ent_off_i + nnz_i)      (local column ids, f32)
//   x[col_off_i .. + n_i), y[row_off_i .. + m_i)  (SpMV)
//   B[col_off_i*K ..], C[row_off_i*K ..]          (SpMM, row-major, width K)
// with 64-bit offsets, so the batch may exceed 2^31 entries while indices
// inside an instance stay int32.
//
// Work: the host splits every instance into tasks of `rows_per_task` rows;
// a persistent grid (a few workgroups per CU) pulls tasks from a queue
// counter.  Each instance carries its own schedule:
//   SpMV: G lanes per row, G in {1, 2, 4, ..., 64} (group-per-row with a
//         shuffle-tree reduction; G = 1 is thread-per-row), or G = 0, the
//         stream schedule: the workgroup's products go through LDS in
//         coalesced chunks and each thread sums one row from LDS (balanced
//         by entries instead of rows; suits short rows)
//   SpMM: L lanes per row along k (K / L columns per lane), L in {8,16,32,64};
//         when K = 4L each lane owns 4 adjacent columns (16-byte accesses)
// XCD affinity: a whole-batch launch gives each of the 8 XCDs its own task
// queue.  The host assigns instances to XCDs by work (largest first, to the
// least loaded), so an XCD's L2 holds only the x / B rows of its own
// instances.  Workgroup w runs on XCD w % 8 (round-robin dispatch); it drains
// its XCD's queue, then takes tasks from the other queues in turn, so the
// launch ends balanced.  Which workgroup runs a task never changes its
// result.
// A launch may also cover a contiguous range of instances (their tasks are
// contiguous) through one queue; the autotuner times one instance at a time
// that way with HIP events.  The queue counters are reset by a small kernel,
// so the pair (reset, batch) can be captured into a hipGraph and replayed.

#include <algorithm>
#include <vector>

#include "common.h"
#include "device.h"

namespace taco_hip {
namespace {

struct Inst {
  int64_t pos_off, ent_off, row_off, col_off;
  int32_t m, n, sched, pad;
};

struct Task {
  int32_t inst, r0, r1, pad;
};

constexpr int kBatchThreads = 256;
// Products staged in LDS per chunk by the stream schedule (16 KB).
constexpr int kStreamChunk = 4096;

constexpr int kXcds = 8;
// Queue counters sit 128 bytes apart.
constexpr int kQStride = 32;

__global__ void k_reset_queue(uint32_t* q, int32_t nq) {
  if ((int32_t)threadIdx.x < nq) q[threadIdx.x * kQStride] = 0;
}

// The queues of one launch: nq == 1 serves task slots [lo, hi) directly;
// nq > 1 serves slots [bound[k], bound[k+1]) of queue k, mapped to tasks by
// order[].
struct Queues {
  int32_t nq, lo, hi;
  const int32_t* bound;
  const int32_t* order;
  uint32_t* counter;
};

// Calls body(task index) for every task the workgroup takes.  All threads of
// the workgroup call it together; every counter ends past its queue's end,
// so every workgroup leaves the loop.
template <typename Body>
__device__ __forceinline__ void for_each_task(const Queues& qs, int32_t* s_slot, Body body) {
  const int32_t home = qs.nq > 1 ? (int32_t)(blockIdx.x % (uint32_t)qs.nq) : 0;
  for (int32_t k = 0; k < qs.nq; ++k) {
    const int32_t qi = home + k < qs.nq ? home + k : home + k - qs.nq;
    const int32_t lo = qs.nq > 1 ? qs.bound[qi] : qs.lo;
    const int32_t hi = qs.nq > 1 ? qs.bound[qi + 1] : qs.hi;
    for (;;) {
      if (threadIdx.x == 0) {
        const uint32_t c = atomicAdd(qs.counter + qi * kQStride, 1u);
        *s_slot = c < (uint32_t)(hi - lo) ? lo + (int32_t)c : -1;
      }
      __syncthreads();
      const int32_t slot = *s_slot;
      __syncthreads();
      if (slot < 0) break;
      body(qs.order ? qs.order[slot] : slot);
    }
  }
}

__global__ __launch_bounds__(kBatchThreads) void k_batch_spmv(
    Queues qs, const Task* __restrict__ tasks, const Inst* __restrict__ insts,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ crd,
    const float* __restrict__ vals, const float* __restrict__ x,
    float* __restrict__ y) {
  __shared__ int32_t s_slot;
  __shared__ float s_prod[kStreamChunk];
  const int tid = threadIdx.x;
  for_each_task(qs, &s_slot, [&](int32_t t) {
    const Task tk = tasks[t];
    const Inst in = insts[tk.inst];
    const int G = in.sched;  // lanes per row (power of two <= 64), 0 = stream
    const int32_t* ip = pos + in.pos_off;
    const int32_t* ic = crd + in.ent_off;
    const float* iv = vals + in.ent_off;
    const float* ix = x + in.col_off;
    float* iy = y + in.row_off;
    if (G == 0) {
      // Stream: the workgroup computes the products of up to kStreamChunk
      // consecutive entries into LDS (coalesced loads, one entry per
      // thread), then each thread adds up its own row's products.  Passes
      // of kBatchThreads rows; rows longer than a chunk span chunks.
      for (int32_t rb = tk.r0; rb < tk.r1; rb += kBatchThreads) {
        const int32_t re_ = rb + kBatchThreads < tk.r1 ? rb + kBatchThreads : tk.r1;
        const int32_t r = rb + tid;
        const int32_t q0 = r < re_ ? ip[r] : 0, q1 = r < re_ ? ip[r + 1] : 0;
        const int32_t e_begin = ip[rb], e_end = ip[re_];
        float acc = 0.f;
        for (int32_t c0 = e_begin; c0 < e_end; c0 += kStreamChunk) {
          const int32_t c1 = c0 + kStreamChunk < e_end ? c0 + kStreamChunk : e_end;
          for (int32_t q = c0 + tid; q < c1; q += kBatchThreads)
            s_prod[q - c0] = stream_load(iv + q) * ix[stream_load(ic + q)];
          __syncthreads();
          const int32_t a = q0 > c0 ? q0 : c0, b = q1 < c1 ? q1 : c1;
          for (int32_t q = a; q < b; ++q) acc += s_prod[q - c0];
          __syncthreads();
        }
        if (r < re_) iy[r] = acc;
      }
      return;
    }
    const int lane = tid & (G - 1);
    const int rows_per_pass = kBatchThreads / G;
    for (int32_t r = tk.r0 + tid / G; r < tk.r1; r += rows_per_pass) {
      float acc = 0.f;
      const int32_t end = ip[r + 1];
      for (int32_t q = ip[r] + lane; q < end; q += G)
        acc += stream_load(iv + q) * ix[stream_load(ic + q)];
      for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
      if (lane == 0) iy[r] = acc;
    }
  });
}

__global__ __launch_bounds__(kBatchThreads) void k_batch_spmm(
    Queues qs, int32_t K, const Task* __restrict__ tasks,
    const Inst* __restrict__ insts, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ crd, const float* __restrict__ vals,
    const float* __restrict__ B, float* __restrict__ C) {
  __shared__ int32_t s_slot;
  const int tid = threadIdx.x;
  for_each_task(qs, &s_slot, [&](int32_t t) {
    const Task tk = tasks[t];
    const Inst in = insts[tk.inst];
    const int L = in.sched;  // lanes per row along k
    const int lane = tid & (L - 1);
    const int rows_per_pass = kBatchThreads / L;
    const int32_t* ip = pos + in.pos_off;
    const int32_t* ic = crd + in.ent_off;
    const float* iv = vals + in.ent_off;
    const float* iB = B + in.col_off * (int64_t)K;
    float* iC = C + in.row_off * (int64_t)K;
    if (K == 4 * L) {
      // Each lane owns 4 adjacent columns: 16-byte loads of B rows and
      // stores of C rows, four entries' B rows in flight.
      typedef float f32x4 __attribute__((ext_vector_type(4)));
      for (int32_t r = tk.r0 + tid / L; r < tk.r1; r += rows_per_pass) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const int32_t end = ip[r + 1];
        int32_t q = ip[r];
        for (; q + 4 <= end; q += 4) {
          float a[4];
          int32_t j[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            a[u] = stream_load(iv + q + u);
            j[u] = stream_load(ic + q + u);
          }
          f32x4 b[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) b[u] = *(const f32x4*)(iB + (int64_t)j[u] * K + 4 * lane);
#pragma unroll
          for (int u = 0; u < 4; ++u) acc += a[u] * b[u];
        }
        for (; q < end; ++q) {
          const float a = stream_load(iv + q);
          acc += a * *(const f32x4*)(iB + (int64_t)stream_load(ic + q) * K + 4 * lane);
        }
        // non-temporal: C is written once, B rows keep the caches
        __builtin_nontemporal_store(acc, (f32x4*)(iC + (int64_t)r * K + 4 * lane));
      }
      return;
    }
    for (int32_t r = tk.r0 + tid / L; r < tk.r1; r += rows_per_pass) {
      float acc[4] = {0.f, 0.f, 0.f, 0.f};  // K <= 4 * L
      const int32_t end = ip[r + 1];
      for (int32_t q = ip[r]; q < end; ++q) {
        const float a = stream_load(iv + q);
        const float* brow = iB + (int64_t)stream_load(ic + q) * K;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = lane + c * L;
          if (k < K) acc[c] += a * brow[k];
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = lane + c * L;
        if (k < K) iC[(int64_t)r * K + k] = acc[c];
      }
    }
  });
}

// 32-bit columns of a generated f32 value stream: U[lo, hi) from the same
// splitmix64 counter formula as taco_hip_gen_uniform_f64, rounded to f32.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gen_uniform_f32(int64_t n, uint64_t key,
                                                         double lo, double span,
                                                         float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double u = (double)(mix64(key + (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
    out[i] = (float)(lo + span * u);
  }
}

}  // namespace
}  // namespace taco_hip

using namespace taco_hip;

struct taco_hip_batch {
  int32_t ninst = 0, ntasks = 0, K = 0;
  int op = 0;  // 0 SpMV, 1 SpMM
  Inst* insts = nullptr;
  Task* tasks = nullptr;
  uint32_t* queue = nullptr;   // kXcds counters, kQStride apart
  int32_t* order = nullptr;    // task indices grouped by XCD queue
  int32_t* qbound = nullptr;   // kXcds + 1 slot offsets into order
  const int32_t* pos = nullptr;
  const int32_t* crd = nullptr;
  const float* vals = nullptr;
  const float* in = nullptr;  // x or B
  float* out = nullptr;       // y or C
  int grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<Inst> host_insts;
  std::vector<int32_t> task_begin;  // first task of each instance, ninst + 1
};

namespace {
void batch_free(taco_hip_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void* q : {(void*)b->insts, (void*)b->tasks, (void*)b->queue, (void*)b->order,
                  (void*)b->qbound})
    if (q) (void)hipFree(q);
  delete b;
}

// whole: every task through the per-XCD queues; otherwise tasks [t0, t0+nt)
// through one queue.
int launch_batch(taco_hip_batch* b, hipStream_t st, int32_t t0, int32_t nt, bool whole) {
  if (nt <= 0) return 0;
  Queues qs{1, t0, t0 + nt, nullptr, nullptr, b->queue};
  if (whole) qs = Queues{kXcds, 0, 0, b->qbound, b->order, b->queue};
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, st, b->queue, qs.nq);
  const unsigned grid = (unsigned)(b->grid < nt ? b->grid : nt);
  if (b->op == 0)
    hipLaunchKernelGGL(k_batch_spmv, dim3(grid), dim3(kBatchThreads), 0, st, qs, b->tasks,
                       b->insts, b->pos, b->crd, b->vals, b->in, b->out);
  else
    hipLaunchKernelGGL(k_batch_spmm, dim3(grid), dim3(kBatchThreads), 0, st, qs, b->K,
                       b->tasks, b->insts, b->pos, b->crd, b->vals, b->in, b->out);
  TH_LAUNCH_CHECK("batched kernel");
  return 0;
}

// Instances to XCD queues, largest first to the least loaded (work ~ rows +
// entries); returns the task order grouped by queue and the queue bounds.
void xcd_order(const std::vector<Inst>& insts, const int64_t* nnz,
               const std::vector<int32_t>& task_begin, std::vector<int32_t>* order,
               std::vector<int32_t>* bound) {
  const int32_t ninst = (int32_t)insts.size();
  std::vector<int32_t> by_work((size_t)ninst);
  for (int32_t i = 0; i < ninst; ++i) by_work[(size_t)i] = i;
  auto work = [&](int32_t i) { return (int64_t)insts[(size_t)i].m + nnz[i]; };
  std::stable_sort(by_work.begin(), by_work.end(),
                   [&](int32_t a, int32_t c) { return work(a) > work(c); });
  std::vector<int64_t> load(kXcds, 0);
  std::vector<int32_t> home((size_t)ninst, 0);
  for (int32_t i : by_work) {
    int x = 0;
    for (int k = 1; k < kXcds; ++k)
      if (load[(size_t)k] < load[(size_t)x]) x = k;
    home[(size_t)i] = x;
    load[(size_t)x] += work(i);
  }
  order->clear();
  bound->assign(kXcds + 1, 0);
  for (int x = 0; x < kXcds; ++x) {
    for (int32_t i = 0; i < ninst; ++i)
      if (home[(size_t)i] == x)
        for (int32_t t = task_begin[(size_t)i]; t < task_begin[(size_t)i + 1]; ++t)
          order->push_back(t);
    (*bound)[(size_t)x + 1] = (int32_t)order->size();
  }
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
// SpMV: 0 (stream) or lanes per row 1..64; SpMM: lanes 8..64 with K <= 4*lanes.
bool valid_sched(int op, int s, int K) {
  if (op == 0) return s == 0 || (pow2(s) && s <= 64);
  return pow2(s) && s >= 8 && s <= 64 && K <= 4 * s;
}
}  // namespace

extern "C" {

int taco_hip_batch_create(int32_t op, int32_t ninst, const int32_t* m,
                          const int32_t* n, const int64_t* nnz,
                          const int32_t* sched, int32_t rows_per_task,
                          int32_t K, const int32_t* pos, const int32_t* crd,
                          const float* vals, const float* in, float* out,
                          taco_hip_batch** result) {
  TH_REQUIRE(result, "batch: null output");
  *result = nullptr;
  TH_REQUIRE(op == 0 || op == 1, "batch: op must be 0 (SpMV) or 1 (SpMM)");
  TH_REQUIRE(ninst >= 0 && m && n && nnz && sched, "batch: bad instance arrays");
  TH_REQUIRE(rows_per_task > 0, "batch: rows_per_task must be positive");
  TH_REQUIRE(op == 0 || (K > 0 && K <= 256), "batch: SpMM width must be in 1..256");
  auto* b = new taco_hip_batch();
  b->op = op;
  b->ninst = ninst;
  b->K = K;
  b->pos = pos;
  b->crd = crd;
  b->vals = vals;
  b->in = in;
  b->out = out;
  std::vector<Task> tasks;
  b->host_insts.resize((size_t)ninst);
  int64_t po = 0, eo = 0, ro = 0, co = 0;
  for (int32_t i = 0; i < ninst; ++i) {
    const int s = sched[i];
    if (!valid_sched(op, s, K)) {
      batch_free(b);
      return fail("batch: instance %d has schedule %d, not a valid lane count", i, s);
    }
    if (m[i] < 0 || n[i] < 0 || nnz[i] < 0) {
      batch_free(b);
      return fail("batch: instance %d has negative sizes", i);
    }
    Inst& in_ = b->host_insts[(size_t)i];
    in_.pos_off = po;
    in_.ent_off = eo;
    in_.row_off = ro;
    in_.col_off = co;
    in_.m = m[i];
    in_.n = n[i];
    in_.sched = s;
    in_.pad = 0;
    b->task_begin.push_back((int32_t)tasks.size());
    for (int32_t r = 0; r < m[i]; r += rows_per_task)
      tasks.push_back(Task{i, r, r + rows_per_task < m[i] ? r + rows_per_task : m[i], 0});
    po += (int64_t)m[i] + 1;
    eo += nnz[i];
    ro += m[i];
    co += n[i];
  }
  b->task_begin.push_back((int32_t)tasks.size());
  if (tasks.size() > (size_t)INT32_MAX) {
    batch_free(b);
    return fail("batch: too many tasks");
  }
  b->ntasks = (int32_t)tasks.size();
  std::vector<int32_t> order, bound;
  xcd_order(b->host_insts, nnz, b->task_begin, &order, &bound);
  hipStream_t st = current_stream();
  if (hipMalloc((void**)&b->insts, sizeof(Inst) * (ninst > 0 ? ninst : 1)) ||
      hipMalloc((void**)&b->tasks, sizeof(Task) * (tasks.empty() ? 1 : tasks.size())) ||
      hipMalloc((void**)&b->queue, sizeof(uint32_t) * kQStride * kXcds) ||
      hipMalloc((void**)&b->order, sizeof(int32_t) * (order.empty() ? 1 : order.size())) ||
      hipMalloc((void**)&b->qbound, sizeof(int32_t) * (kXcds + 1))) {
    batch_free(b);
    return fail("batch: device allocation failed");
  }
  if ((ninst > 0 && hipMemcpyAsync(b->insts, b->host_insts.data(), sizeof(Inst) * ninst,
                                   hipMemcpyHostToDevice, st)) ||
      (!tasks.empty() && hipMemcpyAsync(b->tasks, tasks.data(), sizeof(Task) * tasks.size(),
                                        hipMemcpyHostToDevice, st)) ||
      (!order.empty() && hipMemcpyAsync(b->order, order.data(), sizeof(int32_t) * order.size(),
                                        hipMemcpyHostToDevice, st)) ||
      hipMemcpyAsync(b->qbound, bound.data(), sizeof(int32_t) * (kXcds + 1),
                     hipMemcpyHostToDevice, st) ||
      hipStreamSynchronize(st)) {
    batch_free(b);
    return fail("batch: upload failed");
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->grid = cus * 8;
  *result = b;
  return 0;
}

// Replaces the per-instance schedules (same validity rules as create).
int taco_hip_batch_set_schedules(taco_hip_batch* b, const int32_t* sched) {
  TH_REQUIRE(b && sched, "batch: null argument");
  for (int32_t i = 0; i < b->ninst; ++i) {
    const int s = sched[i];
    TH_REQUIRE(valid_sched(b->op, s, b->K),
               "batch: instance %d has schedule %d, not a valid lane count", i, s);
    b->host_insts[(size_t)i].sched = s;
  }
  if (b->ninst > 0)
    TH_CHECK(hipMemcpyAsync(b->insts, b->host_insts.data(), sizeof(Inst) * b->ninst,
                            hipMemcpyHostToDevice, current_stream()));
  TH_CHECK(hipStreamSynchronize(current_stream()));
  return 0;
}

int taco_hip_batch_run(taco_hip_batch* b) {
  TH_REQUIRE(b, "batch: null");
  if (int rc = launch_batch(b, current_stream(), 0, b->ntasks, true)) return rc;
  return finish("taco_hip_batch_run");
}

// Runs only instances [i0, i1) (their tasks are contiguous).
int taco_hip_batch_run_range(taco_hip_batch* b, int32_t i0, int32_t i1) {
  TH_REQUIRE(b && i0 >= 0 && i0 <= i1 && i1 <= b->ninst, "batch: bad instance range");
  const int32_t t0 = b->task_begin[(size_t)i0], t1 = b->task_begin[(size_t)i1];
  if (int rc = launch_batch(b, current_stream(), t0, t1 - t0, false)) return rc;
  return finish("taco_hip_batch_run_range");
}

// Per-instance timing for the autotuner: instance i's tasks are launched
// `reps` times back to back between two HIP events, every instance in turn
// on the current stream, with one synchronisation at the end.  times_ns[i]
// is the mean time of one launch over instance i alone.
int taco_hip_batch_time_instances(taco_hip_batch* b, int32_t reps, double* times_ns) {
  TH_REQUIRE(b && times_ns && reps > 0, "batch: bad timing arguments");
  if (b->ninst == 0) return 0;
  hipStream_t st = current_stream();
  std::vector<hipEvent_t> ev((size_t)b->ninst + 1, nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      rc = fail("batch: hipEventCreate failed");
      break;
    }
  for (int32_t i = 0; rc == 0 && i < b->ninst; ++i) {
    if (hipEventRecord(ev[(size_t)i], st) != hipSuccess) {
      rc = fail("batch: hipEventRecord failed");
      break;
    }
    const int32_t t0 = b->task_begin[(size_t)i], t1 = b->task_begin[(size_t)i + 1];
    for (int32_t r = 0; rc == 0 && r < reps; ++r) rc = launch_batch(b, st, t0, t1 - t0, false);
  }
  if (rc == 0 && hipEventRecord(ev[(size_t)b->ninst], st) != hipSuccess)
    rc = fail("batch: hipEventRecord failed");
  if (rc == 0 && hipEventSynchronize(ev[(size_t)b->ninst]) != hipSuccess)
    rc = fail("batch: hipEventSynchronize failed");
  for (int32_t i = 0; rc == 0 && i < b->ninst; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[(size_t)i], ev[(size_t)i + 1]) != hipSuccess) {
      rc = fail("batch: hipEventElapsedTime failed");
      break;
    }
    times_ns[i] = (double)ms * 1e6 / reps;
  }
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

// Captures (queue reset, batch kernel) into a hipGraph on a private stream.
int taco_hip_batch_capture(taco_hip_batch* b) {
  TH_REQUIRE(b, "batch: null");
  if (b->exec) return 0;
  hipStream_t cap;
  TH_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  int rc = 0;
  if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail("batch: begin capture failed");
  } else {
    rc = launch_batch(b, cap, 0, b->ntasks, true);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(cap, &g) != hipSuccess && !rc) rc = fail("batch: end capture failed");
    if (!rc) {
      b->graph = g;
      if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess)
        rc = fail("batch: graph instantiate failed");
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cap);
  return rc;
}

int taco_hip_batch_launch_graph(taco_hip_batch* b) {
  TH_REQUIRE(b && b->exec, "batch: no captured graph");
  TH_CHECK(hipGraphLaunch(b->exec, current_stream()));
  return finish("taco_hip_batch_launch_graph");
}

void taco_hip_batch_destroy(taco_hip_batch* b) { batch_free(b); }

int taco_hip_batch_info(const taco_hip_batch* b, int64_t* info) {
  TH_REQUIRE(b && info, "batch: null argument");
  info[0] = b->ninst;
  info[1] = b->ntasks;
  info[2] = b->grid;
  return 0;
}

// f32 values U[lo, hi) on the device with the counter formula of
// taco_hip_gen_uniform_f64 (value i = lo + (hi-lo) * unit(i)), rounded.
int taco_hip_gen_uniform_f32_device(int64_t n, uint64_t seed, uint64_t stream,
                                    double lo, double hi, float* out) {
  if (n <= 0) return 0;
  const uint64_t key = mix64(seed ^ mix64(stream));
  int64_t grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_gen_uniform_f32, dim3((unsigned)grid), dim3(256), 0,
                     current_stream(), n, key, lo, hi - lo, out);
  return finish("taco_hip_gen_uniform_f32_device");
}

}  // extern "C"
