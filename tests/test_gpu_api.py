"""The taco-API facade end to end, written like the reference's GPU tests
(test/tests-scheduling-eval.cpp:1210-1307): build tensors with insert/pack,
schedule with the reference's commands, compile/assemble/compute, and
compare with taco's results on the same inputs (the golden vectors the real
taco wrote) and with the default schedule."""
import os

import numpy as np
import pytest

import oracle
from taco_amd import api as taco
from taco_amd import workloads as W
from tests import schedules as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def csr_tensor(name, dims, pos, crd, vals, insert=True):
    A = taco.Tensor(name, dims, taco.CSR)
    if insert:  # the reference's way: insert every entry, then pack
        for r in range(dims[0]):
            for q in range(pos[r], pos[r + 1]):
                A.insert([r, int(crd[q])], float(vals[q]))
        A.pack()
    else:
        A.set_csr(pos, crd, vals)
    return A


@pytest.mark.parametrize("schedule", ["default", "spmv_gpu", "rows", "thread_per_row", "merge",
                                      "block_rows"])
def test_spmv_gpu_reference_inputs(schedule):
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    i, j = taco.IndexVar("i"), taco.IndexVar("j")
    A = csr_tensor("A", [102, 103], d["spmvgpu_pos"], d["spmvgpu_crd"], d["spmvgpu_vals"])
    x = taco.Tensor("x", [103], taco.Format([taco.Dense]))
    for q, v in enumerate(d["spmvgpu_x"]):
        x.insert([q], float(v))
    x.pack()
    y = taco.Tensor("y", [102], taco.Format([taco.Dense]))
    precomputed = A[i, j] * x[j]
    y[i] = precomputed
    stmt = y.get_assignment().concretize()
    if schedule == "spmv_gpu":
        stmt = S.schedule_spmv_gpu(stmt, A, precomputed, i, j)
    elif schedule == "rows":
        stmt = S.schedule_spmv_rows_gpu(stmt, A, precomputed, i, j)
    elif schedule == "thread_per_row":
        stmt = S.schedule_spmv_thread_per_row_gpu(stmt, A, precomputed, i, j)
    elif schedule == "merge":
        stmt = S.schedule_spmv_merge_gpu(stmt, A, precomputed, i, j)
    elif schedule == "block_rows":
        stmt = S.schedule_spmv_block_rows_gpu(stmt, A, precomputed, i, j)
    y.compile(stmt)
    y.assemble()
    y.compute()
    expected = taco.Tensor("expected", [102], taco.Format([taco.Dense]))
    expected[i] = A[i, j] * x[j]
    expected.evaluate()
    assert np.array_equal(y.to_numpy(), expected.to_numpy())
    assert np.array_equal(y.to_numpy(), d["spmvgpu_y"])
    assert taco.equals(y, expected)  # taco's own predicate, as its GPU tests use it


@pytest.mark.parametrize("schedule", ["spmm_gpu", "merge", "rows"])
def test_spmm_gpu_reference_inputs(schedule):
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    i, j, k = taco.IndexVar("i"), taco.IndexVar("j"), taco.IndexVar("k")
    A = csr_tensor("A", [102, 103], d["spmmgpu_pos"], d["spmmgpu_crd"], d["spmmgpu_vals"])
    B = taco.Tensor("B", [103, 128], taco.dense_matrix)
    B.set_values(d["spmmgpu_B"])
    C = taco.Tensor("C", [102, 128], taco.dense_matrix)
    precomputed = A[i, j] * B[j, k]
    C[i, k] = precomputed
    sched = {"spmm_gpu": S.schedule_spmm_gpu, "merge": S.schedule_spmm_merge_gpu,
             "rows": S.schedule_spmm_rows_gpu}[schedule]
    stmt = sched(C.get_assignment().concretize(), A, precomputed, i, j, k)
    C.compile(stmt)
    C.assemble()
    C.compute()
    assert np.array_equal(C.to_numpy(), d["spmmgpu_C"])


def test_spgemm_two_phase_matches_oracle():
    i, j, k = taco.IndexVar("i"), taco.IndexVar("j"), taco.IndexVar("k")
    Ah = W.csr_fixed(400, 300, 6, 51)
    Bh = W.csr_powerlaw(300, 350, seed=52, kmax=40)
    A = csr_tensor("A", [400, 300], Ah.pos, Ah.crd, Ah.vals, insert=False)
    B = csr_tensor("B", [300, 350], Bh.pos, Bh.crd, Bh.vals, insert=False)
    C = taco.Tensor("C", [400, 350], taco.CSR)
    C[i, k] = A[i, j] * B[j, k]
    C.compile()
    C.assemble()
    C.compute()
    pos, crd, vals = C.to_csr()
    rpos, rcrd = oracle.spgemm_assemble(Ah, Bh)
    assert np.array_equal(pos, rpos) and np.array_equal(crd, rcrd)
    ex, cond = oracle.exact_spgemm_values(Ah, Bh, rpos, rcrd)
    assert oracle.error_report(vals, ex, cond, 1e-12)["ok"]
    E = taco.Tensor("E", [400, 350], taco.CSR)
    E.set_csr(rpos, rcrd, oracle.spgemm_compute(Ah, Bh, rpos, rcrd))
    assert taco.equals(C, E)


def test_f32_spmv_through_the_api():
    Ah = W.csr_powerlaw(3000, 2500, seed=53, kmax=500)
    xh = W.uniform(2500, 53, W.STREAM_X)
    i, j = taco.IndexVar("i"), taco.IndexVar("j")
    A = taco.Tensor("A", [3000, 2500], taco.CSR, taco.Float32)
    A.set_csr(Ah.pos, Ah.crd, Ah.vals)
    x = taco.Tensor("x", [2500], taco.Format([taco.Dense]), taco.Float32)
    x.set_values(xh)
    y = taco.Tensor("y", [3000], taco.Format([taco.Dense]), taco.Float32)
    pre = A[i, j] * x[j]
    y[i] = pre
    y.compile(S.schedule_spmv_gpu(y.get_assignment().concretize(), A, pre, i, j))
    y.assemble()
    y.compute()
    A64 = W.Csr(3000, 2500, Ah.pos, Ah.crd, Ah.vals.astype(np.float32).astype(np.float64))
    ye, cond = oracle.exact_spmv(A64, xh.astype(np.float32).astype(np.float64))
    assert oracle.error_report(y.to_numpy(), ye, cond, 1e-5)["ok"]


needs_weights = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(taco.__file__), "data", "selector.npz")),
    reason="no trained selector weights in taco_amd/data")


@needs_weights
def test_learned_schedule_through_compute():
    """set_learned_schedule_enabled: an unscheduled SpMV and SpMM get the
    selector's schedule at compute(); the slab plan is kept on A and dropped
    when A's values change; every result passes the FP64 oracle."""
    from taco_amd import selector as Sel
    i, j, k = taco.IndexVar("i"), taco.IndexVar("j"), taco.IndexVar("k")
    Ah = W.csr_powerlaw(60000, 60000, seed=54, kmax=20000)
    xh = W.uniform(Ah.n, 54, W.STREAM_X)
    taco.set_learned_schedule_enabled(True)
    try:
        A = csr_tensor("A", [Ah.m, Ah.n], Ah.pos, Ah.crd, Ah.vals, insert=False)
        x = taco.Tensor("x", [Ah.n], taco.Format([taco.Dense]))
        x.set_values(xh)
        y = taco.Tensor("y", [Ah.m], taco.Format([taco.Dense]))
        y[i] = A[i, j] * x[j]
        y.compile()
        y.compute()
        picked = y._kernel.chosen()
        print(f"learned SpMV schedule: {picked}", flush=True)
        assert picked in Sel.CLASS_NAMES[:Sel.N_SPMV]
        ye, cond = oracle.exact_spmv(Ah, xh)
        assert oracle.error_report(y.to_numpy(), ye, cond, 1e-12)["ok"]
        # the slab plan through compute(): built once, reused, then dropped
        A._schedule_choice[(0, 0)] = (Sel.PLAN_CLASS, None)
        y.compute()
        plan = A._plan
        assert plan is not None
        assert oracle.error_report(y.to_numpy(), ye, cond, 1e-12)["ok"]
        y.compute()
        assert A._plan is plan
        vals2 = -2.0 * Ah.vals
        A.set_csr(Ah.pos, Ah.crd, vals2)
        assert A._plan is None and A._schedule_choice == {}
        A._schedule_choice[(0, 0)] = (Sel.PLAN_CLASS, None)
        y.compute()
        assert A._plan is not None and A._plan is not plan
        ye2, cond2 = oracle.exact_spmv(W.Csr(Ah.m, Ah.n, Ah.pos, Ah.crd, vals2), xh)
        assert oracle.error_report(y.to_numpy(), ye2, cond2, 1e-12)["ok"]
        # SpMM, K = 32
        Bh = W.uniform(Ah.n * 32, 55, W.STREAM_B).reshape(Ah.n, 32)
        B = taco.Tensor("B", [Ah.n, 32], taco.dense_matrix)
        B.set_values(Bh)
        C = taco.Tensor("C", [Ah.m, 32], taco.dense_matrix)
        C[i, k] = A[i, j] * B[j, k]
        C.compile()
        C.compute()
        assert C._kernel.chosen(32) in Sel.CLASS_NAMES[Sel.N_SPMV:]
        Ce, condc = oracle.exact_spmm(W.Csr(Ah.m, Ah.n, Ah.pos, Ah.crd, vals2), Bh)
        assert oracle.error_report(C.to_numpy(), Ce, condc, 1e-12)["ok"]
    finally:
        taco.set_learned_schedule_enabled(False)
