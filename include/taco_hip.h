/* taco_hip: MI355X (gfx950, wave64) backend for taco's sparse hot path.
 *
 * This is the C ABI of libtaco_hip.so.  Every entry point takes plain
 * pointers, sizes and taco_tensor_t descriptors; none takes a framework type.
 * All functions returning int return 0 on success and a non-zero code on
 * failure; taco_hip_last_error() then describes the failure (the reference
 * calls exit() or ignores CUDA errors, src/codegen/codegen_cuda.cpp:58-67,
 * src/cuda.cpp:45-51; this backend reports them instead).
 *
 * Execution model: kernels are enqueued on the library's current stream
 * (taco_hip_set_stream; default: the null stream).  Entry points are
 * stream-ordered; callers that need the reference's blocking behaviour
 * (codegen_cuda.cpp:609 synchronises after every launch) call
 * taco_hip_synchronize() or enable taco_hip_set_blocking(1).
 */
#ifndef TACO_HIP_H
#define TACO_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "taco/taco_tensor_t.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================
 * Runtime: replaces taco's CUDA runtime (reference include/taco/cuda.h:15-28,
 * src/cuda.cpp:11-72).
 * ==================================================================== */

/* should_use_CUDA_codegen (cuda.cpp:13): backend selector. */
int taco_hip_should_use_codegen(void);
/* set_CUDA_codegen_enabled (cuda.cpp:21-23). */
void taco_hip_set_codegen_enabled(int enabled);
/* should_use_CUDA_unified_memory (cuda.cpp:17-19). */
int taco_hip_should_use_unified_memory(void);
/* set_CUDA_unified_memory_enabled (cuda.cpp:25-28).  When disabled (the
 * default), taco_hip_unified_alloc returns device-resident memory. */
void taco_hip_set_unified_memory_enabled(int enabled);
/* get_default_CUDA_compiler_flags (cuda.cpp:30-40): hipcc flags for the
 * current device's architecture.  Returned string is owned by the library. */
const char* taco_hip_get_default_compiler_flags(void);
/* cuda_unified_alloc (cuda.cpp:54-64): NULL for size 0; managed memory when
 * unified memory is enabled, device memory otherwise.  On a machine without
 * a GPU (code generation only) it returns host memory. */
void* taco_hip_unified_alloc(size_t size);
/* cuda_unified_free (cuda.cpp:66-72).  Frees memory from
 * taco_hip_unified_alloc or any allocation this library hands out (outputs
 * of assemble functions).  Like hipFree it synchronizes the device.  Device
 * blocks of 1 MiB and more (sizes rounded up to 2 MiB) are kept, up to 32
 * blocks and 32 GiB per process, and handed out again by the next
 * taco_hip_unified_alloc of the same rounded size; an allocation that finds
 * the device out of memory returns the kept blocks first.  Managed blocks
 * are never kept. */
void taco_hip_unified_free(void* ptr);
/* Reallocation for generated host code (taco's realloc emulation,
 * codegen_cuda.cpp:1086-1136): a new taco_hip_unified_alloc block holding the
 * first min(old, new) bytes of ptr, which is released.  NULL on failure. */
void* taco_hip_unified_realloc(void* ptr, size_t old_bytes, size_t new_bytes);
/* Zero-fill of such a block (Allocate with clear, codegen_cuda.cpp:1100-1112),
 * complete when the call returns. */
int taco_hip_unified_clear(void* ptr, size_t bytes);

/* Description of the last failure on the calling thread. */
const char* taco_hip_last_error(void);
/* Number of visible HIP devices (0 on a machine without a GPU). */
int taco_hip_device_count(void);
/* gfx architecture name of the current device, e.g. "gfx950". */
const char* taco_hip_device_arch(void);
/* Stream used by all launches of this library on the calling thread. */
int taco_hip_set_stream(void* stream);
void* taco_hip_get_stream(void);
/* 1: synchronise the stream at the end of every entry point. */
void taco_hip_set_blocking(int blocking);
int taco_hip_synchronize(void);
/* kind: 0 host->device, 1 device->host, 2 device->device. Synchronous. */
int taco_hip_memcpy(void* dst, const void* src, size_t bytes, int kind);
int taco_hip_memset(void* dst, int value, size_t bytes);

/* ======================================================================
 * Seeded synthetic inputs (host, OpenMP).  Counter-based (splitmix64), so a
 * given (seed, row) always yields the same row regardless of thread count.
 * ==================================================================== */

/* Uniform values in [lo, hi) for element indices [0, n) of stream `stream`. */
void taco_hip_gen_uniform_f64(int64_t n, uint64_t seed, uint64_t stream,
                              double lo, double hi, double* out);
/* Per-row entry counts: Zipf(alpha) on [kmin, kmax] drawn by inverse CDF,
 * rescaled so the mean is `mean` (round to nearest, clip to [kmin, kmax]).
 * mean <= 0 disables the rescale.  Returns the total. */
int64_t taco_hip_gen_row_counts_zipf(int32_t m, double alpha, int32_t kmin,
                                     int32_t kmax, double mean, uint64_t seed,
                                     int32_t* counts);
/* Per-row counts uniform in [kmin, kmax]. Returns the total. */
int64_t taco_hip_gen_row_counts_uniform(int32_t m, int32_t kmin, int32_t kmax,
                                        uint64_t seed, int32_t* counts);
/* pos[0] = 0, pos[i+1] = pos[i] + counts[i].  Returns 0, or 1 when the total
 * exceeds INT32_MAX (taco's int32 index arithmetic). */
int taco_hip_counts_to_pos(int32_t m, const int32_t* counts, int32_t* pos);
/* Column indices of a CSR pattern: row i gets pos[i+1]-pos[i] distinct
 * columns in [0, n).  structure 0: uniform without replacement; 1: banded,
 * uniform without replacement from [i-band, i+band] clipped to [0, n);
 * 2: block diagonal, dense `band` x `band` blocks (row count ignored beyond
 * the block width).  Each row is sorted ascending.  Returns 0, or 1 if a row
 * asks for more columns than its structure holds. */
int taco_hip_gen_csr_columns(int32_t m, int32_t n, const int32_t* pos,
                             int32_t structure, int32_t band, uint64_t seed,
                             int32_t* crd);

/* ======================================================================
 * Hand-written HIP kernels for the hot path.  Signatures follow the
 * generated-module ABI (int f(taco_tensor_t* results..., operands...),
 * reference src/codegen/codegen.cpp:514-535): the result first, then the
 * operands.  Value type follows the descriptors' csize.
 * ==================================================================== */

typedef enum taco_hip_algo {
  /* One thread per row: taco's default GPU schedule split(i,256) ->
   * GPUBlock/GPUThread (reference transformations.cpp:1585-1596). */
  TACO_HIP_THREAD_PER_ROW = 0,
  /* group_size lanes per row, lanes stride the row's entries, wave-level
   * reduction: scheduleSpMVRowsGPU (tests-scheduling-eval.cpp:211-223);
   * for SpMM the lanes run along k instead (scheduleSpMMRowsGPU, :345-355). */
  TACO_HIP_GROUP_PER_ROW = 1,
  /* Equal entry chunks per thread/wave, row boundaries resolved with
   * segmented wave reductions and atomics: scheduleSpMVGPU
   * (tests-scheduling-eval.cpp:193-209), scheduleSpMMGPU (:249-268). */
  TACO_HIP_NNZ_SPLIT = 2,
  /* Merge-path over (row ends, entries): every thread gets the same number
   * of merge items; cross-workgroup rows are fixed up from a carry array in
   * order, without atomics.  No reference counterpart. */
  TACO_HIP_MERGE_PATH = 3,
  /* SpMM only: bf16 operands on MFMA, f32 accumulate. */
  TACO_HIP_MFMA = 4
} taco_hip_algo;

typedef struct taco_hip_schedule {
  int32_t algo;             /* taco_hip_algo */
  int32_t block_size;       /* threads per workgroup, multiple of 64 */
  int32_t items_per_thread; /* entries (NNZ_SPLIT) or merge items (MERGE_PATH) */
  int32_t group_size;       /* lanes per row (GROUP_PER_ROW), power of 2 <= 64 */
  int32_t rows_per_group;   /* rows per lane group (GROUP_PER_ROW) */
  int32_t reserved[3];
} taco_hip_schedule;

/* Default schedule per operation kind (0 SpMV, 1 SpMM) for a matrix. */
void taco_hip_default_schedule(int32_t op, int32_t m, int64_t nnz, int32_t k,
                               taco_hip_schedule* out);

/* y(i) = A(i,j) * x(j).  A: CSR {dense, compressed}; x, y: dense vectors.
 * f64 and f32.  sched == NULL uses taco_hip_default_schedule. */
int taco_hip_spmv_compute(taco_tensor_t* y, taco_tensor_t* A,
                          taco_tensor_t* x, const taco_hip_schedule* sched);

/* No counterpart in the reference: taco_hip_spmv_compute for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, schedules and
 * results; it calls no hipMalloc/hipFree.  The MERGE_PATH schedule's
 * carries go to a static 1 MiB device array (one per GPU, held until the
 * process ends) instead of the library workspace, which bounds that
 * schedule to (rows + entries) / (256 * items_per_thread) workgroups of
 * 12 bytes (f64) or 8 bytes (f32) each: 87,376 workgroups in f64, i.e.
 * 178.9M rows + entries at 8 items per thread; larger problems are rejected
 * with an error before anything is launched.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (several streams or host
 * threads) race.  Nothing in it needs clearing between calls (every
 * workgroup writes its carry before the fix-up reads it). */
int taco_hip_spmv_compute_noalloc(taco_tensor_t* y, taco_tensor_t* A,
                                  taco_tensor_t* x, const taco_hip_schedule* sched);

/* C(i,k) = A(i,j) * B(j,k).  A CSR; B, C dense row-major.  f64, f32, and
 * bf16 A/B with f32 C (A->csize == B->csize == 16, C->csize == 32). */
int taco_hip_spmm_compute(taco_tensor_t* C, taco_tensor_t* A,
                          taco_tensor_t* B, const taco_hip_schedule* sched);

/* Two-phase SpGEMM C(i,k) = A(i,j) * B(j,k), all CSR, f64 values.
 * assemble: computes the structural pattern (sorted columns, bit-exact with
 * taco's CPU assemble) and allocates C's pos, crd and vals (freed with
 * taco_hip_unified_free), storing them into C->indices[1] and C->vals.
 * compute: fills C->vals on that pattern. */
int taco_hip_spgemm_assemble(taco_tensor_t* C, taco_tensor_t* A,
                             taco_tensor_t* B);
int taco_hip_spgemm_compute(taco_tensor_t* C, taco_tensor_t* A,
                            taco_tensor_t* B);

/* Slab-tiled SpMV plan (f64).  A storage conversion of a CSR matrix for
 * SpMV whose x exceeds one XCD's L2: columns split into slabs of slab_width
 * (<= 2^17; 0 = 2^17, 1 MB of x), rows into `parts` workgroup parts (0 = one
 * per CU, or a multiple of the CU count when rows would crowd the LDS)
 * whose entries are reordered slab by slab, and by column inside each
 * (part, slab) segment, so that resident workgroups sweep the slabs in step
 * and gather x from L2, a wave's lanes from nearby columns.  compute's
 * `unroll` is the entries per thread (0 = the kernel form's default: 8 for
 * the interleaved form, 4 for the others; 1, 2, 4, 8, and 16 for the
 * interleaved form at 256 threads).  No reference counterpart: the
 * reference's nearest analogue is packing a tensor into another format
 * before compute (src/tensor.cpp pack()).  Built on the device from A's
 * arrays; A's values are copied, so a plan reflects A at creation time.
 * compute: y = A x with the planned A; row sums are accumulated with LDS
 * atomics, so the summation order (and last bits) may vary run to run. */
typedef struct taco_hip_spmv_plan taco_hip_spmv_plan;
int taco_hip_spmv_plan_create(taco_tensor_t* A, int32_t parts,
                              int32_t slab_width, taco_hip_spmv_plan** out);
int taco_hip_spmv_plan_compute(taco_hip_spmv_plan* plan, taco_tensor_t* y,
                               taco_tensor_t* x, int32_t unroll);
void taco_hip_spmv_plan_destroy(taco_hip_spmv_plan* plan);
/* info[0..4] = parts, slabs, slab width, row weight, max rows in a part. */
int taco_hip_spmv_plan_info(const taco_hip_spmv_plan* plan, int64_t* info);
/* Launch form of the main kernel: workgroup size 256 (default) or 512, and
 * the kernel form: 2 (default) thread t of the workgroup takes entries t,
 * t + T, ... of a segment, so a wave's load covers 64 consecutive
 * column-ordered entries (8 entries per thread by default); 0 each thread
 * takes consecutive entries and adds its runs of one row itself (4 by
 * default); 1 runs of one row are summed across the wave before the LDS
 * atomics (4 by default). */
int taco_hip_spmv_plan_configure(taco_hip_spmv_plan* plan, int32_t threads,
                                 int32_t form);
/* The main kernel's launch shape: workgroup size (by default 512 when each
 * CU holds one part, else 256), kernel form, and the entries per thread
 * used when compute's `unroll` is 0. */
int taco_hip_spmv_plan_shape(const taco_hip_spmv_plan* plan, int32_t* threads,
                             int32_t* form, int32_t* unroll);
/* Kernel timing: HIP events bracket the main kernel (k_slab_spmv) of the
 * next `launches` compute calls on their stream (0 stops and releases).
 * plan_kernel_time waits for the last recorded call and returns the summed
 * kernel time and the number of recorded calls. */
int taco_hip_spmv_plan_time_kernel(taco_hip_spmv_plan* plan, int32_t launches);
int taco_hip_spmv_plan_kernel_time(taco_hip_spmv_plan* plan, double* total_ms,
                                   int64_t* launches);

/* ======================================================================
 * Batched uneven SpMV / SpMM (f32): many CSR instances in one persistent
 * launch.  No reference counterpart (taco launches one kernel per tensor).
 * Instances are concatenated: instance i's pos (m_i + 1 int32 local
 * offsets), crd (local columns) and vals follow instance i-1's; x/B and y/C
 * segments likewise (B, C row-major with width K).  sched[i] is the lanes
 * per row: SpMV any power of two 1..64, SpMM 8..64 with K <= 4 * lanes.
 * The device buffers are borrowed and must outlive the batch.
 * ==================================================================== */
typedef struct taco_hip_batch taco_hip_batch;
/* op 0 = SpMV (in = x, out = y), 1 = SpMM (in = B, out = C, width K). */
int taco_hip_batch_create(int32_t op, int32_t ninst, const int32_t* m,
                          const int32_t* n, const int64_t* nnz,
                          const int32_t* sched, int32_t rows_per_task,
                          int32_t K, const int32_t* pos, const int32_t* crd,
                          const float* vals, const float* in, float* out,
                          taco_hip_batch** result);
int taco_hip_batch_set_schedules(taco_hip_batch* b, const int32_t* sched);
/* One launch (queue reset + persistent kernel) on the current stream. */
int taco_hip_batch_run(taco_hip_batch* b);
/* One launch over instances [i0, i1) only (their tasks are contiguous). */
int taco_hip_batch_run_range(taco_hip_batch* b, int32_t i0, int32_t i1);
/* Autotuner timing: each instance alone, `reps` launches back to back
 * between two HIP events on the current stream; times_ns[i] = mean time of
 * one launch over instance i.  Synchronous (one wait at the end). */
int taco_hip_batch_time_instances(taco_hip_batch* b, int32_t reps,
                                  double* times_ns);
/* Capture (queue reset, persistent kernel) into a hipGraph; then replay it
 * on the current stream with taco_hip_batch_launch_graph. */
int taco_hip_batch_capture(taco_hip_batch* b);
int taco_hip_batch_launch_graph(taco_hip_batch* b);
void taco_hip_batch_destroy(taco_hip_batch* b);
/* info[0..2] = instances, tasks, persistent workgroups. */
int taco_hip_batch_info(const taco_hip_batch* b, int64_t* info);
/* Device f32 values U[lo, hi): element i is taco_hip_gen_uniform_f64's
 * element i of the same (seed, stream), rounded to f32. */
int taco_hip_gen_uniform_f32_device(int64_t n, uint64_t seed, uint64_t stream,
                                    double lo, double hi, float* out);

/* Device exclusive scan: pos[0] = 0, pos[i+1] = pos[i] + counts[i], on the
 * current stream.  Returns non-zero (and records an error) if the total
 * exceeds the int32 index range; synchronises the stream to check. */
int taco_hip_device_exclusive_scan(const int32_t* counts, int32_t m,
                                   int32_t* pos);
/* A device buffer of at least `bytes` from the library's reusable workspace
 * (slots 0..7; slots 4..7 are reserved for generated modules).  Valid until
 * the next request of the same slot on this device; never freed by the
 * caller.  NULL on failure. */
void* taco_hip_workspace_get(size_t bytes, int slot);

/* ======================================================================
 * Learned schedule selector.  No reference counterpart (taco schedules are
 * written by hand).  Features of a CSR matrix and an MLP
 * 32 -> 128 -> 128 -> 25 (24 schedule logits: 0..12 SpMV, 13..23 SpMM, and
 * a predicted log time) with forward, backward and Adam kernels in FP32.
 * ==================================================================== */
/* Exact integer statistics of A (30 values): m, n, nnz, then sum of row
 * lengths, sum of squared lengths, max length, empty rows, 16 log2-length
 * bins, entries in rows >= 64 and >= 1024, then over min(m, 4096) evenly
 * spaced rows: rows, entries, entries within 64 of the diagonal column,
 * entries in the 8x8 diagonal block, and the summed distance from the
 * diagonal in 1/4096 units.  Synchronous. */
int taco_hip_csr_stats(taco_tensor_t* A, int64_t* stats);
/* The 32 selector features of an operation (op 0 SpMV, 1 SpMM of width K)
 * from the statistics above (host arithmetic only). */
int taco_hip_features_from_stats(const int64_t* stats, int32_t op, int32_t K,
                                 float* features);
int taco_hip_features(taco_tensor_t* A, int32_t op, int32_t K, float* features);

typedef struct taco_hip_mlp taco_hip_mlp;
/* Parameters: one flat f32 array W1[128][32], b1, W2[128][128], b2,
 * W3[25][128], b3 (23,961 values). */
int64_t taco_hip_mlp_param_count(void);
int taco_hip_mlp_create(taco_hip_mlp** out);
void taco_hip_mlp_destroy(taco_hip_mlp* net);
/* Host parameter copies; set also clears the optimizer state. */
int taco_hip_mlp_set_params(taco_hip_mlp* net, const float* host);
int taco_hip_mlp_get_params(taco_hip_mlp* net, float* host);
/* Z[B][25] = network(X[B][32]) (device pointers, current stream). */
int taco_hip_mlp_forward(taco_hip_mlp* net, const float* X, int32_t B, float* Z);
/* Loss (mean over the batch) and its gradient for device inputs X[B][32],
 * target classes cls[B] and log times logt[B]; gradients copied to
 * grads_host (23,961 floats) when non-null.  Synchronous. */
int taco_hip_mlp_gradients(taco_hip_mlp* net, const float* X, const int32_t* cls,
                           const float* logt, int32_t B, float* grads_host,
                           float* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8; an L2 term weight_decay * p
 * is added to the gradient); loss optional (synchronous when requested). */
int taco_hip_mlp_train_step(taco_hip_mlp* net, const float* X, const int32_t* cls,
                            const float* logt, int32_t B, float lr,
                            float weight_decay, float* loss);

/* ======================================================================
 * Code generation (CodeGen_HIP) and the JIT module loader.
 * ==================================================================== */

/* Statement: 0 SpMV y(i)=A(i,j)*x(j); 1 SpMM C(i,k)=A(i,j)*B(j,k) with dense
 * row-major B, C; 2 SpGEMM C(i,k)=A(i,j)*B(j,k) with CSR C (two-phase). */
typedef enum taco_hip_gen_op {
  TACO_HIP_OP_SPMV = 0,
  TACO_HIP_OP_SPMM = 1,
  TACO_HIP_OP_SPGEMM = 2
} taco_hip_gen_op;

/* Schedule families, each the lowering of one of the reference's GPU
 * schedules (test/tests-scheduling-eval.cpp):
 *   DEFAULT    parallelizeOuterLoop's split(i,block) -> GPUBlock/GPUThread
 *              (src/index_notation/transformations.cpp:1585-1596)
 *   WARP_ROWS  scheduleSpMVRowsGPU (:211-223): rows over lane groups,
 *              entries over the lanes of a group, GPUWarpReduction
 *   NNZ_SPLIT  scheduleSpMVGPU (:193-209) / scheduleSpMMGPU (:249-268):
 *              fuse(i,j) -> pos -> split into block/warp/thread entry
 *              chunks, block-start search, atomics
 *   TWO_PHASE  SpGEMM assemble (pattern) + compute (values), one wave per
 *              row (taco's CPU two-phase SpGEMM, tests-scheduling-eval.cpp:
 *              68-94, made GPU-parallel) */
typedef enum taco_hip_gen_family {
  TACO_HIP_GEN_DEFAULT = 0,
  TACO_HIP_GEN_WARP_ROWS = 1,
  TACO_HIP_GEN_NNZ_SPLIT = 2,
  TACO_HIP_GEN_TWO_PHASE = 3,
  /* Merge-path load balancing over row ends and entries; no reference
   * schedule (the facade maps fuse(i,j,f) split without pos() onto it).
   * SpMV: nnz_per_thread merge items per thread, products staged in LDS.
   * SpMM: teams of warp_size lanes along k with `bound` columns per lane,
   * nnz_per_thread merge items per team, entries loaded by the team and
   * broadcast.  Rows shared between threads/teams are added atomically. */
  TACO_HIP_GEN_MERGE = 4,
  /* SpMV, one workgroup per row: pos(j) split over the workgroup's threads
   * with a Temporary and no warp level, which taco turns into a
   * GPUBlockReduction consumer (transformations.cpp:1049-1056; printed at
   * codegen_cuda.cpp:875-890).  Reduced by taco_hip_block_sum. */
  TACO_HIP_GEN_BLOCK_ROWS = 5
} taco_hip_gen_family;

typedef struct taco_hip_gen_spec {
  int32_t op;             /* taco_hip_gen_op */
  int32_t family;         /* taco_hip_gen_family */
  int32_t csize;          /* value bits: 64 or 32 */
  int32_t block_size;     /* BLOCK_SIZE (threads per workgroup) */
  int32_t nnz_per_thread; /* NNZ_PER_THREAD (SpMV) / NNZ_PER_WARP (SpMM) */
  int32_t warp_size;      /* lanes per group (the reference's WARP_SIZE) */
  int32_t rows_per_warp;  /* ROWS_PER_WARP (WARP_ROWS) */
  int32_t bound;          /* bound(dense_val, B) of scheduleSpMMGPU: K <= B*warp */
} taco_hip_gen_spec;

/* Lower `spec` and print it.  what = 0: HIP module source (CodeGen_HIP);
 * what = 1: the IR.  Returns a malloc'ed string (free with
 * taco_hip_free_string) or NULL with the error recorded. */
char* taco_hip_generate(const taco_hip_gen_spec* spec, int32_t what);
void taco_hip_free_string(char* s);

/* JIT module (reference include/taco/codegen/module.h:19-58,
 * src/codegen/module.cpp:111-215): source text -> hipcc -> shared object ->
 * dlopen; functions are called through their _shim_ wrappers.  Compiled
 * objects are cached by a hash of (source, compiler, flags) in
 * $TACO_HIP_CACHE, default <library dir>/_jit_cache.  The compiler is
 * $TACO_HIPCC (default /opt/rocm/bin/hipcc), flags $TACO_HIPCCFLAGS
 * (default taco_hip_get_default_compiler_flags()). */
typedef struct taco_hip_module taco_hip_module;
taco_hip_module* taco_hip_module_create(void);
int taco_hip_module_set_source(taco_hip_module* m, const char* source);
const char* taco_hip_module_get_source(const taco_hip_module* m);
/* Compiles (or finds in the cache) and loads; returns 0 on success. */
int taco_hip_module_compile(taco_hip_module* m);
/* Path of the loaded shared object ("" before compile). */
const char* taco_hip_module_path(const taco_hip_module* m);
void* taco_hip_module_get_func(taco_hip_module* m, const char* name);
/* callFuncPacked: calls _shim_<name>(args). */
int taco_hip_module_call_packed(taco_hip_module* m, const char* name,
                                void** args);
void taco_hip_module_destroy(taco_hip_module* m);

/* Used by generated modules: error reporting and end-of-call handling. */
int taco_hip_module_report(int code, const char* what);
int taco_hip_module_finish(int rc);
int taco_hip_host_read_bytes(void* dst, const void* src, size_t bytes);

/* ======================================================================
 * Tensor files (reference include/taco/storage/file_io_mtx.h, file_io_tns.h;
 * src/storage/file_io_mtx.cpp:37-160, file_io_tns.cpp:30-87): MatrixMarket
 * (.mtx, .ttx) and FROSTT (.tns) read into a coordinate list, which the
 * caller packs (taco's read() inserts every coordinate and packs).
 * ==================================================================== */

typedef struct taco_hip_coo taco_hip_coo;
/* Reads `path` by its extension.  MatrixMarket: "real" field, "general" or
 * "symmetric" (the mirrored entry is added for i != j), "coordinate" or
 * "array" (every value, first mode fastest); other headers are errors, as
 * in taco.  FROSTT: the order is the first line's token count minus one,
 * each dimension the largest coordinate in its mode. */
int taco_hip_read_tensor_file(const char* path, taco_hip_coo** out);
/* order, dims[0..min(order, max_dims)), entry count, and whether the file
 * was a dense array. */
int taco_hip_coo_info(const taco_hip_coo* coo, int32_t* order, int32_t* dims,
                      int32_t max_dims, int64_t* nnz, int32_t* dense);
/* Copies the 0-based coordinates (nnz x order, row-major) and values. */
int taco_hip_coo_data(const taco_hip_coo* coo, int32_t* coords, double* vals);
void taco_hip_coo_destroy(taco_hip_coo* coo);
/* writeMTX / writeTNS (file_io_mtx.cpp:233-290, file_io_tns.cpp:89-130):
 * by extension; dense = 1 writes a MatrixMarket array (vals holds every
 * position, first mode fastest; coords unused).  Values with 17 significant
 * digits. */
int taco_hip_write_tensor_file(const char* path, int32_t order, const int32_t* dims,
                               int64_t nnz, const int32_t* coords, const double* vals,
                               int32_t dense);

/* ======================================================================
 * Measurement.  No reference counterpart: a bound for the gather-limited
 * SpMV kernels.  Times random 8-byte loads, and nothing else, from an f64
 * array of 2^log2_entries entries (10..25).  mode 0: each lane loads its own
 * element, `unroll` loads in flight per lane (1, 2, 4, 8, 16); mode 1: each
 * wave loads one wave-uniform element per load through the scalar data
 * cache, `unroll` in flight per wave (1..16); mode 2: both (1..8).  Grid:
 * wg_per_cu workgroups of 256 threads per CU (1..32); about target_loads
 * loads per launch.  out[0] = mean ms per launch over `reps` launches after
 * one warm-up, out[1] = lane loads per launch, out[2] = wave-uniform loads
 * per launch.  Allocates and frees its own buffers; synchronous. */
int taco_hip_gather_ceiling(int32_t mode, int32_t unroll, int32_t log2_entries,
                            int32_t wg_per_cu, int64_t target_loads, int32_t reps,
                            double* out);

#ifdef __cplusplus
}
#endif

#endif
