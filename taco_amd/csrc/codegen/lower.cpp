// Iteration-graph builders for the hot-path statements.
//
// Each builder emits the IR taco's lowerer produces for the statement under
// the named schedule (reference src/lower/lowerer_impl_imperative.cpp:
// 647-1581 for the loop forms, 450-607 for assignments, 3055-3364 for result
// initialisation), using this backend's IR vocabulary.  Tensor properties
// are unpacked at the top of each function as taco's codegen does
// (src/codegen/codegen.cpp:232-272): A1_dimension, A2_pos, A2_crd, A_vals...
#include "lower.h"

#include <stdexcept>

namespace taco_hip {

using namespace ir;

namespace {

struct Csr {
  Expr t, dim1, pos, crd, vals, nnz_hint;
  std::vector<Stmt> unpack;
};

Csr unpack_csr(const std::string& n, DType vt, bool want_vals = true) {
  Csr c;
  c.t = tensor_var(n);
  c.dim1 = var(n + "1_dimension", DType::Int32);
  c.pos = var(n + "2_pos", DType::Int32, true);
  c.crd = var(n + "2_crd", DType::Int32, true);
  c.vals = var(n + "_vals", vt, true);
  c.unpack.push_back(decl(c.dim1, get_property(c.t, TensorProperty::Dimension, 0)));
  c.unpack.push_back(decl(c.pos, get_property(c.t, TensorProperty::Indices, 1, 0, "",
                                              DType::Int32, true)));
  c.unpack.push_back(decl(c.crd, get_property(c.t, TensorProperty::Indices, 1, 1, "",
                                              DType::Int32, true)));
  if (want_vals)
    c.unpack.push_back(decl(c.vals, get_property(c.t, TensorProperty::Values, 0, 0, "",
                                                 vt, true)));
  return c;
}

struct Dense {
  Expr t, dim1, dim2, vals;
  std::vector<Stmt> unpack;
};

Dense unpack_dense(const std::string& n, DType vt, int order) {
  Dense d;
  d.t = tensor_var(n);
  d.dim1 = var(n + "1_dimension", DType::Int32);
  d.dim2 = var(n + "2_dimension", DType::Int32);
  d.vals = var(n + "_vals", vt, true);
  d.unpack.push_back(decl(d.dim1, get_property(d.t, TensorProperty::Dimension, 0)));
  if (order == 2)
    d.unpack.push_back(decl(d.dim2, get_property(d.t, TensorProperty::Dimension, 1)));
  d.unpack.push_back(decl(d.vals, get_property(d.t, TensorProperty::Values, 0, 0, "",
                                               vt, true)));
  return d;
}

std::vector<Stmt> cat(std::vector<Stmt> a, const std::vector<Stmt>& b) {
  a.insert(a.end(), b.begin(), b.end());
  return a;
}

Expr i32(int64_t v) { return lit(v, DType::Int32); }
Expr i64(Expr e) { return cast(DType::Int64, e); }
Expr zero(DType vt) { return flit(0.0, vt); }
Expr ceil_div(Expr a, int64_t b) { return div(add(a, i32(b - 1)), i32(b)); }

// Dense result allocation in assemble (taco's assemble allocates result
// values, e.g. `y_vals = malloc(...)`, and packs them back).
Stmt assemble_dense(const Dense& out, Expr count) {
  return block({allocate(out.vals, count),
                assign(get_property(out.t, TensorProperty::Values, 0, 0, "",
                                    out.vals->type, true), out.vals),
                ret(i32(0))});
}

// evaluate = assemble then compute.
Stmt evaluate_body(const std::vector<Expr>& args) {
  Expr rc = var("taco_rc", DType::Int32);
  return block({decl(rc, call("assemble", args, DType::Int32)),
                if_then(neq(rc, i32(0)), ret(rc)),
                ret(call("compute", args, DType::Int32))});
}

// ---- SpMV ----

// parallelizeOuterLoop: split(i, i0, i1, B), i0 -> GPUBlock, i1 -> GPUThread
// (NoRaces); body is the position loop with a scalar-promoted temporary.
Stmt spmv_default(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                  const Dense& x, const Dense& y) {
  const int B = s.block_size > 0 ? s.block_size : 256;
  Expr i0 = var("i0", DType::Int32), i1 = var("i1", DType::Int32);
  Expr i = var("i", DType::Int32), jA = var("jA", DType::Int32);
  Expr j = var("j", DType::Int32), t = var("tjy_val", vt);
  Stmt inner = block({
      decl(t, zero(vt)),
      for_loop(jA, load(A.pos, i), load(A.pos, add(i, i32(1))), i32(1),
               block({decl(j, load(A.crd, jA)),
                      assign(t, add(t, mul(load(A.vals, jA), load(x.vals, j))))})),
      store(y.vals, i, t)});
  Stmt kernel = for_loop(
      i0, i32(0), ceil_div(A.dim1, B), i32(1),
      for_loop(i1, i32(0), i32(B), i32(1),
               block({decl(i, add(mul(i0, i32(B)), i1)),
                      if_then(lt(i, A.dim1), inner)}),
               LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::NoRaces),
      LoopKind::Serial, ParallelUnit::GPUBlock, OutputRaceStrategy::NoRaces);
  return kernel;
}

// scheduleSpMVRowsGPU: split(i, block, block_row, R*B), split(block_row,
// warp_row, warp, B/W), pos(j), split(jpos, thread_nz, thread, W),
// reorder(block, warp, warp_row, thread, thread_nz); thread Temporary.
Stmt spmv_warp_rows(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                    const Dense& x, const Dense& y) {
  const int B = s.block_size > 0 ? s.block_size : 256;
  const int W = s.warp_size > 0 ? s.warp_size : 32;
  const int R = s.rows_per_warp > 0 ? s.rows_per_warp : 1;
  if (B % W) throw std::runtime_error("WARP_ROWS: block size not a multiple of warp size");
  const int warps = B / W;
  const int rows_per_tb = R * B;
  Expr block_v = var("block", DType::Int32), warp = var("warp", DType::Int32);
  Expr warp_row = var("warp_row", DType::Int32), thread = var("thread", DType::Int32);
  Expr thread_nz = var("thread_nz", DType::Int32), i = var("i", DType::Int32);
  Expr jposA = var("jposA", DType::Int32), j = var("j", DType::Int32);
  Expr t = var("tthread_val", vt);
  Expr row_len = sub(load(A.pos, add(i, i32(1))), load(A.pos, i));
  Stmt producer = for_loop(
      thread, i32(0), i32(W), i32(1),
      for_loop(thread_nz, i32(0), ceil_div(row_len, W), i32(1),
               block({decl(jposA, add(load(A.pos, i), add(mul(thread_nz, i32(W)), thread))),
                      if_then(lt(jposA, load(A.pos, add(i, i32(1)))),
                              block({decl(j, load(A.crd, jposA)),
                                     assign(t, add(t, mul(load(A.vals, jposA),
                                                          load(x.vals, j))))}))})),
      LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::Temporary);
  Stmt consumer = for_loop(thread, i32(0), i32(W), i32(1), store(y.vals, i, t),
                           LoopKind::Serial, ParallelUnit::GPUWarpReduction);
  Stmt row = for_loop(
      warp_row, i32(0), i32(rows_per_tb / warps), i32(1),
      block({decl(i, add(mul(block_v, i32(rows_per_tb)),
                         add(mul(warp_row, i32(warps)), warp))),
             if_then(lt(i, A.dim1), block({decl(t, zero(vt)), producer, consumer}))}));
  return for_loop(block_v, i32(0), ceil_div(A.dim1, rows_per_tb), i32(1),
                  for_loop(warp, i32(0), i32(warps), i32(1), row, LoopKind::Serial,
                           ParallelUnit::GPUWarp),
                  LoopKind::Serial, ParallelUnit::GPUBlock);
}

// One workgroup per row: parallelize(i, GPUBlock), pos(j, jpos, A(i,j)),
// split(jpos, jpos0, thread, B), parallelize(thread, GPUThread, Temporary).
// With no GPUWarp parent the Temporary's consumer is a GPUBlockReduction
// (transformations.cpp:1049-1056).
Stmt spmv_block_rows(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                     const Dense& x, const Dense& y) {
  const int B = s.block_size > 0 ? s.block_size : 256;
  if (B > 1024 || (B <= 64 ? (B & (B - 1)) != 0 : B % 64 != 0))
    throw std::runtime_error(
        "BLOCK_ROWS: block size must be a power of two <= 64 or a multiple of 64 <= 1024");
  Expr i = var("i", DType::Int32), thread = var("thread", DType::Int32);
  Expr jpos0 = var("jpos0", DType::Int32), jposA = var("jposA", DType::Int32);
  Expr j = var("j", DType::Int32), t = var("tthread_val", vt);
  Expr row_len = sub(load(A.pos, add(i, i32(1))), load(A.pos, i));
  Stmt producer = for_loop(
      thread, i32(0), i32(B), i32(1),
      for_loop(jpos0, i32(0), ceil_div(row_len, B), i32(1),
               block({decl(jposA, add(load(A.pos, i), add(mul(jpos0, i32(B)), thread))),
                      if_then(lt(jposA, load(A.pos, add(i, i32(1)))),
                              block({decl(j, load(A.crd, jposA)),
                                     assign(t, add(t, mul(load(A.vals, jposA),
                                                          load(x.vals, j))))}))})),
      LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::Temporary);
  Stmt consumer = for_loop(thread, i32(0), i32(B), i32(1), store(y.vals, i, t),
                           LoopKind::Serial, ParallelUnit::GPUBlockReduction);
  return for_loop(i, i32(0), A.dim1, i32(1), block({decl(t, zero(vt)), producer, consumer}),
                  LoopKind::Serial, ParallelUnit::GPUBlock);
}

// Entry count of A read on the host (taco reads A2_pos[A1_dimension]).
Stmt nnz_decl(const Csr& A, Expr nnz) {
  return decl(nnz, load(A.pos, A.dim1));
}

// Block-start search: i_blockStarts[b] = row of entry b * per_block
// (lowerer_impl_imperative.cpp:1062-1106, taco_binarySearchBeforeBlock).
std::vector<Stmt> block_starts(const Csr& A, Expr nnz, Expr num_blocks,
                               Expr starts, int per_block) {
  return {decl(num_blocks, ceil_div(nnz, per_block)),
          decl(starts, lit(0, DType::Int32)),
          allocate_scratch(starts, add(num_blocks, i32(1)), 4),
          launch("taco_hip_binarySearchBeforeBlockLaunch",
                 {A.pos, starts, i32(0), A.dim1, i32(per_block), i32(256), num_blocks})};
}

// scheduleSpMVGPU: fuse(i,j,f), pos(f, fpos, A), split(fpos, block, fpos1,
// K*B), split(fpos1, warp, fpos2, K*W), split(fpos2, thread, thread_nz, K);
// block/warp IgnoreRaces, thread Atomics.  Row changes inside a thread's
// chunk flush with per-lane atomics; the chunk's last row flushes
// wave-aggregated (atomics relative to GPUWarp).
Stmt spmv_nnz_split(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                    const Dense& x, const Dense& y, std::vector<Stmt>& pre) {
  const int B = s.block_size > 0 ? s.block_size : 256;
  const int W = s.warp_size > 0 ? s.warp_size : 32;
  const int K = s.nnz_per_thread > 0 ? s.nnz_per_thread : 8;
  if (B % W) throw std::runtime_error("NNZ_SPLIT: block size not a multiple of warp size");
  const int per_warp = K * W, per_tb = K * B;
  Expr nnz = var("A_nnz", DType::Int32), nb = var("num_blocks", DType::Int32);
  Expr starts = var("i_blockStarts", DType::Int32, true);
  pre.push_back(nnz_decl(A, nnz));
  pre.push_back(launch("taco_hip_fill", {y.vals, i64(A.dim1), zero(vt)}));
  for (auto& st : block_starts(A, nnz, nb, starts, per_tb)) pre.push_back(st);
  Expr block_v = var("block", DType::Int32), warp = var("warp", DType::Int32);
  Expr thread = var("thread", DType::Int32), tnz = var("thread_nz", DType::Int32);
  Expr f0 = var("fpos0", DType::Int32), fA = var("fposA", DType::Int32);
  Expr ip = var("i_pos", DType::Int32), f = var("f", DType::Int32);
  Expr t = var("tval", vt);
  Stmt advance = block({
      store_add(y.vals, ip, t, true, ParallelUnit::GPUThread),
      assign(t, zero(vt)),
      while_loop(gte(fA, load(A.pos, add(ip, i32(1)))), assign(ip, add(ip, i32(1))))});
  Stmt chunk = for_loop(
      tnz, i32(0), i32(K), i32(1),
      block({decl(fA, add(f0, tnz)),
             if_then(lt(fA, nnz),
                     block({decl(f, load(A.crd, fA)),
                            if_then(gte(fA, load(A.pos, add(ip, i32(1)))), advance),
                            assign(t, add(t, mul(load(A.vals, fA), load(x.vals, f))))}))}),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, K);
  // A chunk wholly inside [0, nnz) drops the per-entry guard and reads its K
  // column indices, values and x entries into registers before the row
  // walk: straight-line loads that the compiler merges into vector loads,
  // with every gather in flight at once.  Same products, same order.
  Expr cc = var("chunk_crd", DType::Int32, true), cv = var("chunk_vals", vt, true);
  Expr cx = var("chunk_x", vt, true);
  Stmt prefetch = for_loop(
      tnz, i32(0), i32(K), i32(1),
      block({store(cc, tnz, load(A.crd, add(f0, tnz))),
             store(cv, tnz, load(A.vals, add(f0, tnz)))}),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, K);
  Stmt gather = for_loop(tnz, i32(0), i32(K), i32(1),
                         store(cx, tnz, load(x.vals, load(cc, tnz))), LoopKind::Serial,
                         ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, K);
  Stmt walk = for_loop(
      tnz, i32(0), i32(K), i32(1),
      block({decl(fA, add(f0, tnz)),
             if_then(gte(fA, load(A.pos, add(ip, i32(1)))), advance),
             assign(t, add(t, mul(load(cv, tnz), load(cx, tnz))))}),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, K);
  Stmt full_chunk = block({allocate(cc, i32(K)), allocate(cv, i32(K)), allocate(cx, i32(K)),
                           prefetch, gather, walk});
  Stmt thread_body = block({
      decl(f0, add(add(mul(block_v, i32(per_tb)), mul(warp, i32(per_warp))),
                   mul(thread, i32(K)))),
      if_then(lt(f0, nnz),
              block({decl(ip, call("taco_hip_binarySearchBefore",
                                   {A.pos, load(starts, block_v),
                                    load(starts, add(block_v, i32(1))), f0},
                                   DType::Int32)),
                     decl(t, zero(vt)),
                     if_then(lte(f0, sub(nnz, i32(K))), full_chunk, chunk),
                     store_add(y.vals, ip, t, true, ParallelUnit::GPUWarp)}))});
  return for_loop(
      block_v, i32(0), nb, i32(1),
      for_loop(warp, i32(0), i32(B / W), i32(1),
               for_loop(thread, i32(0), i32(W), i32(1), thread_body, LoopKind::Serial,
                        ParallelUnit::GPUThread, OutputRaceStrategy::Atomics),
               LoopKind::Serial, ParallelUnit::GPUWarp, OutputRaceStrategy::IgnoreRaces),
      LoopKind::Serial, ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces);
}

// Merge-path load balancing (no reference schedule; the HIP backend's
// reading of fuse(i,j,f) split over rows and entries together): the merge
// items are the m row ends and the nnz entries, K per thread, B*K per
// workgroup.  The workgroup first computes the products of its entry range
// into LDS with coalesced loads (one entry per thread per step), then each
// thread finds its start (row i, entry fA) with taco_hip_mergePathSearch,
// adds its entries' products from LDS and, at each row end, adds the sum
// into y[i] atomically (a thread's first and last rows may be shared with
// its neighbours); its last, unfinished row is flushed wave-aggregated.
Stmt spmv_merge(const taco_hip_gen_spec& s, DType vt, const Csr& A, const Dense& x,
                const Dense& y, std::vector<Stmt>& pre) {
  const int B = s.block_size > 0 ? s.block_size : 256;
  const int K = s.nnz_per_thread > 0 ? s.nnz_per_thread : 8;
  Expr nnz = var("A_nnz", DType::Int32), total = var("merge_items", DType::Int64);
  Expr nb = var("num_blocks", DType::Int32);
  pre.push_back(nnz_decl(A, nnz));
  pre.push_back(decl(total, add(i64(A.dim1), i64(nnz))));
  pre.push_back(decl(nb, cast(DType::Int32, div(add(total, lit(B * K - 1, DType::Int64)),
                                                 lit(B * K, DType::Int64)))));
  pre.push_back(launch("taco_hip_fill", {y.vals, i64(A.dim1), zero(vt)}));
  Expr block_v = var("block", DType::Int32), thread = var("thread", DType::Int32);
  Expr D0 = var("block_d0", DType::Int64), D1 = var("block_d1", DType::Int64);
  Expr z0 = var("block_z0", DType::Int32), z1 = var("block_z1", DType::Int32);
  Expr e = var("e", DType::Int32), prod = var("prod", vt, true);
  Expr d0 = var("d0", DType::Int64), q = var("q", DType::Int32);
  Expr i = var("i", DType::Int32), fA = var("fA", DType::Int32), t = var("tval", vt);
  auto search = [&](Expr d) {
    return call("taco_hip_mergePathSearch", {A.pos, A.dim1, d}, DType::Int32);
  };
  Stmt stage = for_loop(
      thread, i32(0), i32(B), i32(1),
      for_loop(e, add(z0, thread), z1, i32(B),
               store(prod, sub(e, z0), mul(load(A.vals, e), load(x.vals, load(A.crd, e))))),
      LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::IgnoreRaces);
  Stmt step = if_then(
      lt(add(d0, i64(q)), total),
      if_then(lt(fA, load(A.pos, add(i, i32(1)))),
              block({assign(t, add(t, load(prod, sub(fA, z0)))), assign(fA, add(fA, i32(1)))}),
              block({store_add(y.vals, i, t, true, ParallelUnit::GPUThread),
                     assign(t, zero(vt)), assign(i, add(i, i32(1)))})));
  Stmt items = for_loop(q, i32(0), i32(K), i32(1), step, LoopKind::Serial,
                        ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, K);
  Stmt consume = for_loop(
      thread, i32(0), i32(B), i32(1),
      block({decl(d0, add(D0, mul(i64(thread), lit(K, DType::Int64)))),
             if_then(lt(d0, total),
                     block({decl(i, search(d0)), decl(fA, cast(DType::Int32, sub(d0, i64(i)))),
                            decl(t, zero(vt)), items,
                            if_then(lt(i, A.dim1),
                                    store_add(y.vals, i, t, true, ParallelUnit::GPUWarp))}))}),
      LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::Atomics);
  Stmt body = block({
      allocate_shared(prod, i32(B * K)),
      decl(D0, mul(i64(block_v), lit(B * K, DType::Int64))),
      decl(D1, min(add(D0, lit(B * K, DType::Int64)), total)),
      decl(z0, cast(DType::Int32, sub(D0, i64(search(D0))))),
      decl(z1, cast(DType::Int32, sub(D1, i64(search(D1))))),
      stage, sync_threads(), consume});
  return for_loop(block_v, i32(0), nb, i32(1), body, LoopKind::Serial,
                  ParallelUnit::GPUBlock, OutputRaceStrategy::IgnoreRaces);
}

std::vector<Stmt> lower_spmv(const taco_hip_gen_spec& s, DType vt) {
  Csr A = unpack_csr("A", vt);
  Dense x = unpack_dense("x", vt, 1), y = unpack_dense("y", vt, 1);
  std::vector<Stmt> pre = cat(cat(y.unpack, A.unpack), x.unpack);
  Stmt k;
  switch (s.family) {
    case TACO_HIP_GEN_DEFAULT: k = spmv_default(s, vt, A, x, y); break;
    case TACO_HIP_GEN_WARP_ROWS: k = spmv_warp_rows(s, vt, A, x, y); break;
    case TACO_HIP_GEN_NNZ_SPLIT: k = spmv_nnz_split(s, vt, A, x, y, pre); break;
    case TACO_HIP_GEN_MERGE: k = spmv_merge(s, vt, A, x, y, pre); break;
    case TACO_HIP_GEN_BLOCK_ROWS: k = spmv_block_rows(s, vt, A, x, y); break;
    default: throw std::runtime_error("SpMV: unsupported schedule family");
  }
  pre.push_back(k);
  pre.push_back(ret(i32(0)));
  std::vector<Expr> outs = {y.t}, ins = {A.t, x.t};
  Stmt compute = function("compute", block(pre), outs, ins);
  Dense ya = unpack_dense("y", vt, 1);
  Stmt assemble = function("assemble",
                           block(cat(ya.unpack, {assemble_dense(ya, ya.dim1)})),
                           outs, ins);
  Stmt evaluate = function("evaluate", evaluate_body({y.t, A.t, x.t}), outs, ins);
  return {assemble, compute, evaluate};
}

// ---- SpMM ----

// Default: split(i, i0, i1, B) over rows, i-j-k loop nest per thread, result
// zero-filled first (initResultArrays).
Stmt spmm_default(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                  const Dense& Bm, const Dense& C, std::vector<Stmt>& pre) {
  const int BS = s.block_size > 0 ? s.block_size : 256;
  pre.push_back(launch("taco_hip_fill", {C.vals, mul(i64(C.dim1), i64(C.dim2)), zero(vt)}));
  Expr i0 = var("i0", DType::Int32), i1 = var("i1", DType::Int32);
  Expr i = var("i", DType::Int32), jA = var("jA", DType::Int32);
  Expr j = var("j", DType::Int32), k = var("k", DType::Int32);
  Expr kC = var("kC", DType::Int64), kB = var("kB", DType::Int64);
  Stmt knest = for_loop(
      k, i32(0), Bm.dim2, i32(1),
      block({decl(kC, add(mul(i64(i), i64(C.dim2)), i64(k))),
             decl(kB, add(mul(i64(j), i64(Bm.dim2)), i64(k))),
             store(C.vals, kC, add(load(C.vals, kC),
                                   mul(load(A.vals, jA), load(Bm.vals, kB))))}));
  Stmt inner = for_loop(jA, load(A.pos, i), load(A.pos, add(i, i32(1))), i32(1),
                        block({decl(j, load(A.crd, jA)), knest}));
  return for_loop(
      i0, i32(0), ceil_div(A.dim1, BS), i32(1),
      for_loop(i1, i32(0), i32(BS), i32(1),
               block({decl(i, add(mul(i0, i32(BS)), i1)), if_then(lt(i, A.dim1), inner)}),
               LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::NoRaces),
      LoopKind::Serial, ParallelUnit::GPUBlock, OutputRaceStrategy::NoRaces);
}

// scheduleSpMMRowsGPU (tests-scheduling-eval.cpp:344-355): split(i, block,
// i1, R*(B/32)), split(i1, warp, warp_row, R), split(k, thread, thread_col,
// 32), reorder(block, warp, warp_row, thread, thread_col, j); block/warp
// IgnoreRaces, thread Atomics.  Each warp owns R consecutive rows; its
// threads own 32-column chunks of C's row (thread_col serial), each
// summing the row's entries in order into C[i,k] with an atomic add (one
// writer per element, so the order is the sequential one).  taco's thread
// loop has extent ceil(K/32), known only at run time; here a lane group has
// kRowsLanes lanes and lane t takes chunks t, t + kRowsLanes, ...  The
// result is zero-filled first (initResultArrays).
constexpr int kRowsLanes = 4;
Stmt spmm_warp_rows(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                    const Dense& Bm, const Dense& C, std::vector<Stmt>& pre) {
  const int B = s.block_size > 0 ? s.block_size : 256;
  const int W = s.warp_size > 0 ? s.warp_size : 32;  // columns per chunk
  const int R = s.rows_per_warp > 0 ? s.rows_per_warp : 4;
  if (B % 32) throw std::runtime_error("WARP_ROWS SpMM: block size not a multiple of 32");
  const int warps = B / 32, rows_per_tb = R * warps;
  pre.push_back(launch("taco_hip_fill", {C.vals, mul(i64(C.dim1), i64(C.dim2)), zero(vt)}));
  Expr block_v = var("block", DType::Int32), warp = var("warp", DType::Int32);
  Expr warp_row = var("warp_row", DType::Int32), lane = var("thread_lane", DType::Int32);
  Expr thread = var("thread", DType::Int32), tc = var("thread_col", DType::Int32);
  Expr i = var("i", DType::Int32), k = var("k", DType::Int32);
  Expr jA = var("jA", DType::Int32), j = var("j", DType::Int32);
  Expr kC = var("kC", DType::Int64);
  Stmt entries = for_loop(
      jA, load(A.pos, i), load(A.pos, add(i, i32(1))), i32(1),
      block({decl(j, load(A.crd, jA)),
             store_add(C.vals, kC,
                       mul(load(A.vals, jA), load(Bm.vals, add(mul(i64(j), i64(Bm.dim2)),
                                                                 i64(k)))),
                       true, ParallelUnit::GPUThread)}));
  Stmt cols = for_loop(
      tc, i32(0), i32(W), i32(1),
      block({decl(k, add(mul(thread, i32(W)), tc)),
             if_then(lt(k, Bm.dim2),
                     block({decl(kC, add(mul(i64(i), i64(C.dim2)), i64(k))), entries}))}));
  Stmt chunks = for_loop(thread, lane, ceil_div(Bm.dim2, W), i32(kRowsLanes), cols);
  Stmt row = for_loop(
      warp_row, i32(0), i32(R), i32(1),
      block({decl(i, add(mul(block_v, i32(rows_per_tb)), add(mul(warp, i32(R)), warp_row))),
             if_then(lt(i, A.dim1),
                     for_loop(lane, i32(0), i32(kRowsLanes), i32(1), chunks, LoopKind::Serial,
                              ParallelUnit::GPUThread, OutputRaceStrategy::Atomics))}));
  return for_loop(block_v, i32(0), ceil_div(A.dim1, rows_per_tb), i32(1),
                  for_loop(warp, i32(0), i32(warps), i32(1), row, LoopKind::Serial,
                           ParallelUnit::GPUWarp),
                  LoopKind::Serial, ParallelUnit::GPUBlock);
}

// scheduleSpMMGPU: reorder(i,j,k), fuse(i,j,f), pos(f), split(fpos, block,
// fpos1, P*(B/W)), split(fpos1, warp, nnz, P), split(k, dense_val, thread,
// W), reorder(block, warp, thread, dense_val, nnz), bound(dense_val, D);
// thread Atomics.
Stmt spmm_nnz_split(const taco_hip_gen_spec& s, DType vt, const Csr& A,
                    const Dense& Bm, const Dense& C, std::vector<Stmt>& pre) {
  const int BS = s.block_size > 0 ? s.block_size : 256;
  const int W = s.warp_size > 0 ? s.warp_size : 32;
  const int P = s.nnz_per_thread > 0 ? s.nnz_per_thread : 8;
  const int D = s.bound > 0 ? s.bound : 4;
  if (BS % W) throw std::runtime_error("NNZ_SPLIT: block size not a multiple of warp size");
  const int warps = BS / W, per_tb = P * warps;
  Expr nnz = var("A_nnz", DType::Int32), nb = var("num_blocks", DType::Int32);
  Expr starts = var("i_blockStarts", DType::Int32, true);
  pre.push_back(nnz_decl(A, nnz));
  pre.push_back(launch("taco_hip_fill", {C.vals, mul(i64(C.dim1), i64(C.dim2)), zero(vt)}));
  for (auto& st : block_starts(A, nnz, nb, starts, per_tb)) pre.push_back(st);
  Expr block_v = var("block", DType::Int32), warp = var("warp", DType::Int32);
  Expr thread = var("thread", DType::Int32), dv = var("dense_val", DType::Int32);
  Expr nz = var("nnz", DType::Int32), f0 = var("fpos0", DType::Int32);
  Expr fA = var("fposA", DType::Int32), ip0 = var("i_pos0", DType::Int32);
  Expr ip = var("i_pos", DType::Int32), f = var("f", DType::Int32);
  Expr k = var("k", DType::Int32), t = var("tval", vt);
  auto cidx = [&](Expr row) { return add(mul(i64(row), i64(C.dim2)), i64(k)); };
  Stmt advance = block({
      store_add(C.vals, cidx(ip), t, true, ParallelUnit::GPUThread),
      assign(t, zero(vt)),
      while_loop(gte(fA, load(A.pos, add(ip, i32(1)))), assign(ip, add(ip, i32(1))))});
  Stmt entries = for_loop(
      nz, i32(0), i32(P), i32(1),
      block({decl(fA, add(f0, nz)),
             if_then(lt(fA, nnz),
                     block({decl(f, load(A.crd, fA)),
                            if_then(gte(fA, load(A.pos, add(ip, i32(1)))), advance),
                            assign(t, add(t, mul(load(A.vals, fA),
                                                 load(Bm.vals, add(mul(i64(f), i64(Bm.dim2)),
                                                                   i64(k))))))}))}));
  Stmt cols = for_loop(
      dv, i32(0), i32(D), i32(1),
      block({decl(k, add(mul(dv, i32(W)), thread)),
             if_then(lt(k, Bm.dim2),
                     block({decl(ip, ip0), decl(t, zero(vt)), entries,
                            store_add(C.vals, cidx(ip), t, true, ParallelUnit::GPUThread)}))}));
  // A chunk wholly inside [0, nnz): its P column indices and values are read
  // once into registers (shared by the lane's D columns), then for each
  // column all P B-row gathers are issued before the row walk.  Same
  // products, same order as the guarded loop.
  Expr cc = var("chunk_crd", DType::Int32, true), cv = var("chunk_vals", vt, true);
  Expr cb = var("chunk_b", vt, true);
  Stmt prefetch = for_loop(
      nz, i32(0), i32(P), i32(1),
      block({store(cc, nz, load(A.crd, add(f0, nz))), store(cv, nz, load(A.vals, add(f0, nz)))}),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, P);
  Stmt gather = for_loop(
      nz, i32(0), i32(P), i32(1),
      store(cb, nz, load(Bm.vals, add(mul(i64(load(cc, nz)), i64(Bm.dim2)), i64(k)))),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, P);
  Stmt walk = for_loop(
      nz, i32(0), i32(P), i32(1),
      block({decl(fA, add(f0, nz)),
             if_then(gte(fA, load(A.pos, add(ip, i32(1)))), advance),
             assign(t, add(t, mul(load(cv, nz), load(cb, nz))))}),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, P);
  Stmt full_cols = for_loop(
      dv, i32(0), i32(D), i32(1),
      block({decl(k, add(mul(dv, i32(W)), thread)),
             if_then(lt(k, Bm.dim2),
                     block({decl(ip, ip0), decl(t, zero(vt)), allocate(cb, i32(P)), gather,
                            walk,
                            store_add(C.vals, cidx(ip), t, true, ParallelUnit::GPUThread)}))}));
  Stmt full = block({allocate(cc, i32(P)), allocate(cv, i32(P)), prefetch, full_cols});
  Stmt thread_body = block({
      decl(f0, add(mul(block_v, i32(per_tb)), mul(warp, i32(P)))),
      if_then(lt(f0, nnz),
              block({decl(ip0, call("taco_hip_binarySearchBefore",
                                    {A.pos, load(starts, block_v),
                                     load(starts, add(block_v, i32(1))), f0},
                                    DType::Int32)),
                     if_then(lte(f0, sub(nnz, i32(P))), full, cols)}))});
  return for_loop(
      block_v, i32(0), nb, i32(1),
      for_loop(warp, i32(0), i32(warps), i32(1),
               for_loop(thread, i32(0), i32(W), i32(1), thread_body, LoopKind::Serial,
                        ParallelUnit::GPUThread, OutputRaceStrategy::Atomics),
               LoopKind::Serial, ParallelUnit::GPUWarp),
      LoopKind::Serial, ParallelUnit::GPUBlock);
}

// Merge-path SpMM (no reference schedule): teams of W lanes along k, D
// columns per lane (k = dv*W + lane, the bound of scheduleSpMMGPU), P merge
// items (row ends and entries) per team.  Inside a row segment the team
// loads up to W entries at once, one per lane, and broadcasts them
// (taco_hip_team_bcast); each lane gathers its columns of the B rows.  The
// team's first row may have started in the previous team and its last row
// may continue in the next: those are added with atomics, every row that
// starts and ends inside the team is stored.
Stmt spmm_merge(const taco_hip_gen_spec& s, DType vt, const Csr& A, const Dense& Bm,
                const Dense& C, std::vector<Stmt>& pre) {
  const int BS = s.block_size > 0 ? s.block_size : 256;
  const int W = s.warp_size > 0 ? s.warp_size : 64;
  const int P = s.nnz_per_thread > 0 ? s.nnz_per_thread : 256;
  const int D = s.bound > 0 ? s.bound : 1;
  if (BS % W || W > 64 || (W & (W - 1)))
    throw std::runtime_error("MERGE: lanes per team must be a power of two <= 64 dividing the block");
  const int teams = BS / W;
  Expr nnz = var("A_nnz", DType::Int32), total = var("merge_items", DType::Int64);
  Expr nb = var("num_blocks", DType::Int32);
  pre.push_back(nnz_decl(A, nnz));
  pre.push_back(decl(total, add(i64(A.dim1), i64(nnz))));
  pre.push_back(decl(nb, cast(DType::Int32, div(add(total, lit((int64_t)teams * P - 1, DType::Int64)),
                                                 lit((int64_t)teams * P, DType::Int64)))));
  pre.push_back(launch("taco_hip_fill", {C.vals, mul(i64(C.dim1), i64(C.dim2)), zero(vt)}));
  Expr block_v = var("block", DType::Int32), warp = var("warp", DType::Int32);
  Expr lane = var("lane", DType::Int32), dv = var("dense_val", DType::Int32);
  Expr d = var("d", DType::Int64), dend = var("dend", DType::Int64);
  Expr i = var("i", DType::Int32), i_first = var("i_first", DType::Int32);
  Expr fA = var("fA", DType::Int32), rend = var("rend", DType::Int32);
  Expr zstop = var("zstop", DType::Int32), zc = var("zc", DType::Int32);
  Expr cnt = var("cnt", DType::Int32), u = var("u", DType::Int32);
  Expr jl = var("jl", DType::Int32), al = var("al", vt);
  Expr j = var("j", DType::Int32), a = var("a", vt), k = var("k", DType::Int32);
  Expr acc = var("acc", vt, true);
  auto cidx = [&](Expr row) { return add(mul(i64(row), i64(C.dim2)), i64(k)); };
  auto each_col = [&](Stmt body) {
    return for_loop(dv, i32(0), i32(D), i32(1),
                    block({decl(k, add(mul(dv, i32(W)), lane)), if_then(lt(k, Bm.dim2), body)}),
                    LoopKind::Serial, ParallelUnit::NotParallel,
                    OutputRaceStrategy::IgnoreRaces, D);
  };
  // Eight broadcast entries at a time: their B rows are gathered together,
  // then added in entry order (the order of the one-at-a-time tail loop).
  constexpr int kBatch = 8;
  Expr u0 = var("u0", DType::Int32), q = var("q", DType::Int32);
  Expr jj = var("batch_j", DType::Int32, true), aa = var("batch_a", vt, true);
  Expr bb = var("batch_b", vt, true);
  Stmt batch_bcast = for_loop(
      q, i32(0), i32(kBatch), i32(1),
      block({store(jj, q, call("taco_hip_team_bcast", {jl, add(u0, q), i32(W)}, DType::Int32)),
             store(aa, q, call("taco_hip_team_bcast", {al, add(u0, q), i32(W)}, vt))}),
      LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces, kBatch);
  Stmt batch_cols = each_col(block({
      allocate(bb, i32(kBatch)),
      for_loop(q, i32(0), i32(kBatch), i32(1),
               store(bb, q, load(Bm.vals, add(mul(i64(load(jj, q)), i64(Bm.dim2)), i64(k)))),
               LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces,
               kBatch),
      for_loop(q, i32(0), i32(kBatch), i32(1),
               store(acc, dv, add(load(acc, dv), mul(load(aa, q), load(bb, q)))),
               LoopKind::Serial, ParallelUnit::NotParallel, OutputRaceStrategy::IgnoreRaces,
               kBatch)}));
  Stmt batches = for_loop(u0, i32(0), sub(cnt, i32(kBatch - 1)), i32(kBatch),
                          block({allocate(jj, i32(kBatch)), allocate(aa, i32(kBatch)),
                                 batch_bcast, batch_cols}));
  Stmt tail = for_loop(
      u, mul(div(cnt, i32(kBatch)), i32(kBatch)), cnt, i32(1),
      block({decl(j, call("taco_hip_team_bcast", {jl, u, i32(W)}, DType::Int32)),
             decl(a, call("taco_hip_team_bcast", {al, u, i32(W)}, vt)),
             each_col(store(acc, dv, add(load(acc, dv),
                                         mul(a, load(Bm.vals, add(mul(i64(j), i64(Bm.dim2)),
                                                                  i64(k)))))))}));
  Stmt gather = block({batches, tail});
  Stmt chunks = for_loop(
      zc, fA, zstop, i32(W),
      block({decl(cnt, min(i32(W), sub(zstop, zc))),
             decl(jl, ternary(lt(lane, cnt), load(A.crd, add(zc, lane)), i32(0))),
             decl(al, ternary(lt(lane, cnt), load(A.vals, add(zc, lane)), zero(vt))),
             gather}));
  Stmt row_end = block({
      each_col(block({if_then(eq(i, i_first),
                              store_add(C.vals, cidx(i), load(acc, dv), true,
                                        ParallelUnit::GPUThread),
                              store(C.vals, cidx(i), load(acc, dv))),
                      store(acc, dv, zero(vt))})),
      assign(i, add(i, i32(1))), assign(d, add(d, lit(1, DType::Int64)))});
  Stmt walk = while_loop(
      lt(d, dend),
      block({decl(rend, load(A.pos, add(i, i32(1)))),
             decl(zstop, cast(DType::Int32, min(add(i64(fA), sub(dend, d)), i64(rend)))),
             chunks,
             assign(d, add(d, i64(sub(zstop, fA)))), assign(fA, zstop),
             if_then(land(lt(d, dend), eq(fA, rend)), row_end)}));
  Stmt team_body = block({
      decl(d, mul(add(mul(i64(block_v), lit(teams, DType::Int64)), i64(warp)),
                  lit(P, DType::Int64))),
      if_then(lt(d, total),
              block({decl(dend, min(add(d, lit(P, DType::Int64)), total)),
                     decl(i, call("taco_hip_mergePathSearch", {A.pos, A.dim1, d}, DType::Int32)),
                     decl(i_first, i), decl(fA, cast(DType::Int32, sub(d, i64(i)))),
                     allocate(acc, i32(D), true), walk,
                     if_then(lt(i, A.dim1),
                             each_col(store_add(C.vals, cidx(i), load(acc, dv), true,
                                                ParallelUnit::GPUThread)))}))});
  return for_loop(
      block_v, i32(0), nb, i32(1),
      for_loop(warp, i32(0), i32(teams), i32(1),
               for_loop(lane, i32(0), i32(W), i32(1), team_body, LoopKind::Serial,
                        ParallelUnit::GPUThread, OutputRaceStrategy::Atomics),
               LoopKind::Serial, ParallelUnit::GPUWarp),
      LoopKind::Serial, ParallelUnit::GPUBlock);
}

std::vector<Stmt> lower_spmm(const taco_hip_gen_spec& s, DType vt) {
  Csr A = unpack_csr("A", vt);
  Dense Bm = unpack_dense("B", vt, 2), C = unpack_dense("C", vt, 2);
  std::vector<Stmt> pre = cat(cat(C.unpack, A.unpack), Bm.unpack);
  Stmt k;
  switch (s.family) {
    case TACO_HIP_GEN_DEFAULT: k = spmm_default(s, vt, A, Bm, C, pre); break;
    case TACO_HIP_GEN_WARP_ROWS: k = spmm_warp_rows(s, vt, A, Bm, C, pre); break;
    case TACO_HIP_GEN_NNZ_SPLIT: k = spmm_nnz_split(s, vt, A, Bm, C, pre); break;
    case TACO_HIP_GEN_MERGE: k = spmm_merge(s, vt, A, Bm, C, pre); break;
    default: throw std::runtime_error("SpMM: unsupported schedule family");
  }
  pre.push_back(k);
  pre.push_back(ret(i32(0)));
  std::vector<Expr> outs = {C.t}, ins = {A.t, Bm.t};
  Stmt compute = function("compute", block(pre), outs, ins);
  Dense Ca = unpack_dense("C", vt, 2);
  Stmt assemble = function(
      "assemble", block(cat(Ca.unpack, {assemble_dense(Ca, mul(i64(Ca.dim1), i64(Ca.dim2)))})),
      outs, ins);
  Stmt evaluate = function("evaluate", evaluate_body({C.t, A.t, Bm.t}), outs, ins);
  return {assemble, compute, evaluate};
}

// ---- SpGEMM (two-phase) ----

// One 64-lane wave per row of C, four waves per workgroup: the GPU form of
// taco's CPU assemble/compute (workspace scatter + sorted index list
// becomes a wave sort or LDS hash set; the pos prefix sum a device scan).
// Workgroups for the compute row loop (taco_hip_row_grid: one per 4 rows,
// at most 4 per CU).  A row takes a few thousand cycles, too short for one
// workgroup per four rows to keep the CUs occupied, so rows are strided
// over a grid the device holds at once.
Stmt row_grid_decl(const Csr& A, Expr nb) {
  return decl(nb, call("taco_hip_row_grid", {A.dim1}, DType::Int32));
}

// Assemble: one workgroup per four rows.  Compute: rows strided over a grid
// the device holds at once (nb workgroups, taco_hip_row_grid).
Stmt row_waves(const Csr& A, Expr i, Stmt body) {
  Expr r0 = var("r0", DType::Int32), w = var("w", DType::Int32);
  Expr lane = var("lane", DType::Int32);
  return for_loop(
      r0, i32(0), ceil_div(A.dim1, 4), i32(1),
      for_loop(w, i32(0), i32(4), i32(1),
               for_loop(lane, i32(0), i32(64), i32(1),
                        block({decl(i, add(mul(r0, i32(4)), w)),
                               if_then(lt(i, A.dim1), body)}),
                        LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::NoRaces),
               LoopKind::Serial, ParallelUnit::GPUWarp, OutputRaceStrategy::NoRaces),
      LoopKind::Serial, ParallelUnit::GPUBlock, OutputRaceStrategy::NoRaces);
}

Stmt strided_row_waves(const Csr& A, Expr i, Expr nb, Stmt body) {
  Expr r0 = var("r0", DType::Int32), w = var("w", DType::Int32);
  Expr lane = var("lane", DType::Int32);
  return for_loop(
      r0, i32(0), nb, i32(1),
      for_loop(w, i32(0), i32(4), i32(1),
               for_loop(lane, i32(0), i32(64), i32(1),
                        for_loop(i, add(mul(r0, i32(4)), w), A.dim1, mul(nb, i32(4)), body),
                        LoopKind::Serial, ParallelUnit::GPUThread, OutputRaceStrategy::NoRaces),
               LoopKind::Serial, ParallelUnit::GPUWarp, OutputRaceStrategy::NoRaces),
      LoopKind::Serial, ParallelUnit::GPUBlock, OutputRaceStrategy::NoRaces);
}

std::vector<Stmt> lower_spgemm(const taco_hip_gen_spec& s) {
  (void)s;
  const DType vt = DType::Float64;
  Csr A = unpack_csr("A", vt), B = unpack_csr("B", vt);
  Expr Ct = tensor_var("C");
  Expr C1 = var("C1_dimension", DType::Int32);
  Expr Cpos = var("C2_pos", DType::Int32, true), Ccrd = var("C2_crd", DType::Int32, true);
  Expr Cvals = var("C_vals", vt, true);
  auto cprop = [&](TensorProperty p, int mode, int idx, DType t, bool ptr) {
    return get_property(Ct, p, mode, idx, "", t, ptr);
  };
  std::vector<Expr> outs = {Ct}, ins = {A.t, B.t};

  // assemble
  Expr nnzc = var("C2_nnz", DType::Int32, true), status = var("status", DType::Int32, true);
  Expr qi = var("qi", DType::Int32), cnt = var("cnt", DType::Int32);
  Expr lane = var("lane", DType::Int32), total = var("C_nnz", DType::Int32);
  Expr i = var("i", DType::Int32);
  std::vector<Stmt> a = cat(A.unpack, B.unpack);
  a.push_back(decl(C1, cprop(TensorProperty::Dimension, 0, 0, DType::Int32, false)));
  a.push_back(decl(Cpos, lit(0)));
  a.push_back(decl(Ccrd, lit(0)));
  a.push_back(decl(Cvals, lit(0)));
  a.push_back(decl(nnzc, lit(0)));
  a.push_back(allocate_scratch(nnzc, add(A.dim1, i32(1)), 5));
  a.push_back(decl(status, lit(0)));
  a.push_back(allocate_scratch(status, i32(1), 6, true));
  a.push_back(comment("count query: distinct structural columns per row"));
  a.push_back(row_waves(A, qi,
                        block({decl(cnt, call("taco_hip_spgemm_row_symbolic<false>",
                                              {A.pos, A.crd, B.pos, B.crd, qi, i32(0),
                                               lit(0), status}, DType::Int32)),
                               if_then(eq(lane, i32(0)), store(nnzc, qi, cnt))})));
  a.push_back(allocate(Cpos, add(A.dim1, i32(1))));
  a.push_back(launch("taco_hip_device_exclusive_scan", {nnzc, A.dim1, Cpos}));
  a.push_back(launch("taco_hip_check_status", {status}));
  a.push_back(decl(total, load(Cpos, A.dim1)));
  a.push_back(allocate(Ccrd, total));
  a.push_back(allocate(Cvals, total));
  a.push_back(comment("insert sorted column indices"));
  a.push_back(row_waves(A, i,
                        evaluate(call("taco_hip_spgemm_row_symbolic<true>",
                                      {A.pos, A.crd, B.pos, B.crd, i, load(Cpos, i),
                                       Ccrd, status}, DType::Int32))));
  a.push_back(assign(cprop(TensorProperty::Indices, 1, 0, DType::Int32, true), Cpos));
  a.push_back(assign(cprop(TensorProperty::Indices, 1, 1, DType::Int32, true), Ccrd));
  a.push_back(assign(cprop(TensorProperty::Values, 0, 0, vt, true), Cvals));
  a.push_back(assign(cprop(TensorProperty::ValuesSize, 0, 0, DType::Int32, false), total));
  a.push_back(ret(i32(0)));
  Stmt assemble = function("assemble", block(a), outs, ins);

  // compute
  std::vector<Stmt> c = cat(A.unpack, B.unpack);
  c.push_back(decl(Cpos, cprop(TensorProperty::Indices, 1, 0, DType::Int32, true)));
  c.push_back(decl(Ccrd, cprop(TensorProperty::Indices, 1, 1, DType::Int32, true)));
  c.push_back(decl(Cvals, cprop(TensorProperty::Values, 0, 0, vt, true)));
  Expr nb = var("row_blocks", DType::Int32);
  c.push_back(row_grid_decl(A, nb));
  c.push_back(strided_row_waves(A, i, nb,
                        evaluate(call("taco_hip_spgemm_row_numeric",
                                      {A.pos, A.crd, A.vals, B.pos, B.crd, B.vals, i,
                                       Cpos, Ccrd, Cvals}, DType::Int32))));
  c.push_back(ret(i32(0)));
  Stmt compute = function("compute", block(c), outs, ins);
  Stmt evaluate = function("evaluate", evaluate_body({Ct, A.t, B.t}), outs, ins);
  return {assemble, compute, evaluate};
}

}  // namespace

std::vector<Stmt> lower_spec(const taco_hip_gen_spec& s) {
  DType vt;
  if (s.csize == 64)
    vt = DType::Float64;
  else if (s.csize == 32)
    vt = DType::Float32;
  else
    throw std::runtime_error("generated kernels support f64 and f32 values");
  switch (s.op) {
    case TACO_HIP_OP_SPMV: return lower_spmv(s, vt);
    case TACO_HIP_OP_SPMM: return lower_spmm(s, vt);
    case TACO_HIP_OP_SPGEMM:
      if (s.family != TACO_HIP_GEN_TWO_PHASE && s.family != TACO_HIP_GEN_DEFAULT)
        throw std::runtime_error("SpGEMM: only the two-phase schedule exists");
      if (vt != DType::Float64) throw std::runtime_error("SpGEMM: f64 values only");
      return lower_spgemm(s);
    default:
      throw std::runtime_error("unknown statement");
  }
}

}  // namespace taco_hip
