"""The reference's GPU schedules (test/tests-scheduling-eval.cpp:193-268),
written against taco_amd.api with the same commands and parameters.
WARP_SIZE is the reference's 32-lane warp (a lane group of a wave64)."""
from taco_amd.api import (BoundType, IndexVar, OutputRaceStrategy as ORS, ParallelUnit as PU,
                          TensorVar)

WARP_SIZE = 32


def schedule_spmv_gpu(stmt, A, precomputed, i, j, NNZ_PER_THREAD=8, BLOCK_SIZE=256):
    NNZ_PER_WARP = NNZ_PER_THREAD * WARP_SIZE
    NNZ_PER_TB = NNZ_PER_THREAD * BLOCK_SIZE
    f, fpos, fpos1, fpos2 = IndexVar("f"), IndexVar("fpos"), IndexVar("fpos1"), IndexVar("fpos2")
    block, warp, thread = IndexVar("block"), IndexVar("warp"), IndexVar("thread")
    thread_nz, thread_nz_pre = IndexVar("thread_nz"), IndexVar("thread_nz_pre")
    pre = TensorVar("precomputed")
    return (stmt.fuse(i, j, f)
            .pos(f, fpos, A[i, j])
            .split(fpos, block, fpos1, NNZ_PER_TB)
            .split(fpos1, warp, fpos2, NNZ_PER_WARP)
            .split(fpos2, thread, thread_nz, NNZ_PER_THREAD)
            .reorder([block, warp, thread, thread_nz])
            .precompute(precomputed, thread_nz, thread_nz_pre, pre)
            .unroll(thread_nz_pre, NNZ_PER_THREAD)
            .parallelize(block, PU.GPUBlock, ORS.IgnoreRaces)
            .parallelize(warp, PU.GPUWarp, ORS.IgnoreRaces)
            .parallelize(thread, PU.GPUThread, ORS.Atomics))


def schedule_spmv_rows_gpu(stmt, A, precomputed, i, j, ROWS_PER_WARP=1, BLOCK_SIZE=256):
    ROWS_PER_TB = ROWS_PER_WARP * BLOCK_SIZE
    block, warp, thread, thread_nz = (IndexVar("block"), IndexVar("warp"), IndexVar("thread"),
                                      IndexVar("thread_nz"))
    jpos, block_row, warp_row = IndexVar("jpos"), IndexVar("block_row"), IndexVar("warp_row")
    return (stmt.split(i, block, block_row, ROWS_PER_TB)
            .split(block_row, warp_row, warp, BLOCK_SIZE // WARP_SIZE)
            .pos(j, jpos, A[i, j])
            .split(jpos, thread_nz, thread, WARP_SIZE)
            .reorder([block, warp, warp_row, thread, thread_nz])
            .parallelize(block, PU.GPUBlock, ORS.IgnoreRaces)
            .parallelize(warp, PU.GPUWarp, ORS.IgnoreRaces)
            .parallelize(thread, PU.GPUThread, ORS.Temporary))


def schedule_spmv_thread_per_row_gpu(stmt, A, precomputed, i, j, BLOCK_SIZE=256):
    block, thread = IndexVar("block"), IndexVar("thread")
    return (stmt.split(i, block, thread, BLOCK_SIZE)
            .parallelize(block, PU.GPUBlock, ORS.NoRaces)
            .parallelize(thread, PU.GPUThread, ORS.NoRaces))


def schedule_spmv_merge_gpu(stmt, A, precomputed, i, j, ITEMS_PER_THREAD=8, BLOCK_SIZE=256):
    """The HIP backend's merge-path schedule: the fused (i,j) space split
    directly (rows and entries together), without pos()."""
    f, f1, item = IndexVar("f"), IndexVar("f1"), IndexVar("item")
    block, thread = IndexVar("block"), IndexVar("thread")
    return (stmt.fuse(i, j, f)
            .split(f, block, f1, ITEMS_PER_THREAD * BLOCK_SIZE)
            .split(f1, thread, item, ITEMS_PER_THREAD)
            .parallelize(block, PU.GPUBlock, ORS.IgnoreRaces)
            .parallelize(thread, PU.GPUThread, ORS.Atomics))


def schedule_spmm_gpu(stmt, A, precomputed, i, j, k, NNZ_PER_WARP=8, BLOCK_SIZE=256):
    NNZ_PER_TB = NNZ_PER_WARP * (BLOCK_SIZE // WARP_SIZE)
    f, fpos, block, fpos1, warp, nnz = (IndexVar("f"), IndexVar("fpos"), IndexVar("block"),
                                        IndexVar("fpos1"), IndexVar("warp"), IndexVar("nnz"))
    dvu, dv, thread = IndexVar("dense_val_unbounded"), IndexVar("dense_val"), IndexVar("thread")
    return (stmt.reorder([i, j, k])
            .fuse(i, j, f)
            .pos(f, fpos, A[i, j])
            .split(fpos, block, fpos1, NNZ_PER_TB)
            .split(fpos1, warp, nnz, NNZ_PER_WARP)
            .split(k, dvu, thread, WARP_SIZE)
            .reorder([block, warp, thread, dvu, nnz])
            .bound(dvu, dv, 4, BoundType.MaxExact)
            .parallelize(block, PU.GPUBlock, ORS.IgnoreRaces)
            .parallelize(warp, PU.GPUWarp, ORS.IgnoreRaces)
            .parallelize(thread, PU.GPUThread, ORS.Atomics))


def schedule_spmm_rows_gpu(stmt, A, precomputed, i, j, k, ROWS_PER_WARP=4, BLOCK_SIZE=256):
    """scheduleSpMMRowsGPU (tests-scheduling-eval.cpp:344-355): the same
    number of rows per warp, each thread of a warp a 32-column chunk of k."""
    ROWS_PER_TB = ROWS_PER_WARP * (BLOCK_SIZE // WARP_SIZE)
    i1, block, warp = IndexVar("i1"), IndexVar("block"), IndexVar("warp")
    warp_row, thread, thread_col = IndexVar("warp_row"), IndexVar("thread"), IndexVar("thread_col")
    return (stmt.split(i, block, i1, ROWS_PER_TB)
            .split(i1, warp, warp_row, ROWS_PER_WARP)
            .split(k, thread, thread_col, 32)
            .reorder([block, warp, warp_row, thread, thread_col, j])
            .parallelize(block, PU.GPUBlock, ORS.IgnoreRaces)
            .parallelize(warp, PU.GPUWarp, ORS.IgnoreRaces)
            .parallelize(thread, PU.GPUThread, ORS.Atomics))


def schedule_spmm_merge_gpu(stmt, A, precomputed, i, j, k, ITEMS_PER_TEAM=256, BLOCK_SIZE=256,
                            LANES=64, COLS_PER_LANE=2):
    """The HIP backend's merge-path SpMM: scheduleSpMMGPU's loop structure
    with the fused (i,j) space split directly (rows and entries together),
    without pos()."""
    f, f1, item = IndexVar("f"), IndexVar("f1"), IndexVar("item")
    block, warp = IndexVar("block"), IndexVar("warp")
    dvu, dv, thread = IndexVar("dense_val_unbounded"), IndexVar("dense_val"), IndexVar("thread")
    per_tb = ITEMS_PER_TEAM * (BLOCK_SIZE // LANES)
    return (stmt.reorder([i, j, k])
            .fuse(i, j, f)
            .split(f, block, f1, per_tb)
            .split(f1, warp, item, ITEMS_PER_TEAM)
            .split(k, dvu, thread, LANES)
            .reorder([block, warp, thread, dvu, item])
            .bound(dvu, dv, COLS_PER_LANE, BoundType.MaxExact)
            .parallelize(block, PU.GPUBlock, ORS.IgnoreRaces)
            .parallelize(warp, PU.GPUWarp, ORS.IgnoreRaces)
            .parallelize(thread, PU.GPUThread, ORS.Atomics))


def schedule_spmv_block_rows_gpu(stmt, A, precomputed, i, j, BLOCK_SIZE=256):
    """One workgroup per row: the row's positions split over the workgroup's
    threads with a Temporary and no warp level (a GPUBlockReduction)."""
    jpos, jpos0, thread = IndexVar("jpos"), IndexVar("jpos0"), IndexVar("thread")
    return (stmt.pos(j, jpos, A[i, j])
            .split(jpos, jpos0, thread, BLOCK_SIZE)
            .reorder([i, thread, jpos0])
            .parallelize(i, PU.GPUBlock, ORS.NoRaces)
            .parallelize(thread, PU.GPUThread, ORS.Temporary))
