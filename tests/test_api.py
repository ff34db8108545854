"""The taco-API facade's host logic: statements and schedules recognised as
the reference writes them, and taco-style errors for the rest (no GPU)."""
import numpy as np
import pytest

from taco_amd import _lib
from taco_amd import api as taco
from tests import schedules as S


def spmv_setup(dtype=taco.Float64):
    i, j = taco.IndexVar("i"), taco.IndexVar("j")
    A = taco.Tensor("A", [50, 40], taco.CSR, dtype)
    x = taco.Tensor("x", [40], taco.Format([taco.Dense]), dtype)
    y = taco.Tensor("y", [50], taco.Format([taco.Dense]), dtype)
    pre = A[i, j] * x[j]
    y[i] = pre
    return A, x, y, i, j, pre


def recognise(y, stmt):
    kind, a, b = taco._statement_kind(stmt.assignment)
    return kind, taco._recognise(stmt, kind, a)


def test_default_schedule_is_taco_default():
    A, x, y, i, j, pre = spmv_setup()
    assert recognise(y, y.get_assignment().concretize()) == ("spmv", (_lib.GEN_DEFAULT, 256, 0, 0, 0, 0))


def test_reference_spmv_schedules():
    A, x, y, i, j, pre = spmv_setup()
    st = y.get_assignment().concretize()
    assert recognise(y, S.schedule_spmv_gpu(st, A, pre, i, j)) == \
        ("spmv", (_lib.GEN_NNZ_SPLIT, 256, 8, 32, 0, 0))
    assert recognise(y, S.schedule_spmv_gpu(st, A, pre, i, j, 4, 512)) == \
        ("spmv", (_lib.GEN_NNZ_SPLIT, 512, 4, 32, 0, 0))
    assert recognise(y, S.schedule_spmv_rows_gpu(st, A, pre, i, j)) == \
        ("spmv", (_lib.GEN_WARP_ROWS, 256, 0, 32, 1, 0))
    assert recognise(y, S.schedule_spmv_rows_gpu(st, A, pre, i, j, 2, 128)) == \
        ("spmv", (_lib.GEN_WARP_ROWS, 128, 0, 32, 2, 0))
    assert recognise(y, S.schedule_spmv_thread_per_row_gpu(st, A, pre, i, j, 128)) == \
        ("spmv", (_lib.GEN_DEFAULT, 128, 0, 0, 0, 0))
    assert recognise(y, S.schedule_spmv_merge_gpu(st, A, pre, i, j)) == \
        ("spmv", (_lib.GEN_MERGE, 256, 8, 0, 0, 0))
    assert recognise(y, S.schedule_spmv_block_rows_gpu(st, A, pre, i, j)) == \
        ("spmv", (_lib.GEN_BLOCK_ROWS, 256, 0, 0, 0, 0))
    assert recognise(y, S.schedule_spmv_block_rows_gpu(st, A, pre, i, j, 128)) == \
        ("spmv", (_lib.GEN_BLOCK_ROWS, 128, 0, 0, 0, 0))


def test_reference_spmm_schedule():
    i, j, k = taco.IndexVar("i"), taco.IndexVar("j"), taco.IndexVar("k")
    A = taco.Tensor("A", [30, 20], taco.CSR)
    B = taco.Tensor("B", [20, 64], taco.dense_matrix)
    C = taco.Tensor("C", [30, 64], taco.dense_matrix)
    pre = A[i, j] * B[j, k]
    C[i, k] = pre
    st = S.schedule_spmm_gpu(C.get_assignment().concretize(), A, pre, i, j, k)
    assert recognise(C, st) == ("spmm", (_lib.GEN_NNZ_SPLIT, 256, 8, 32, 0, 4))
    st = S.schedule_spmm_merge_gpu(C.get_assignment().concretize(), A, pre, i, j, k)
    assert recognise(C, st) == ("spmm", (_lib.GEN_MERGE, 256, 256, 64, 0, 2))
    st = S.schedule_spmm_rows_gpu(C.get_assignment().concretize(), A, pre, i, j, k)
    assert recognise(C, st) == ("spmm", (_lib.GEN_WARP_ROWS, 256, 0, 32, 4, 0))
    st = S.schedule_spmm_rows_gpu(C.get_assignment().concretize(), A, pre, i, j, k, 2, 128)
    assert recognise(C, st) == ("spmm", (_lib.GEN_WARP_ROWS, 128, 0, 32, 2, 0))
    st = S.schedule_spmm_merge_gpu(C.get_assignment().concretize(), A, pre, i, j, k, 64, 128,
                                   32, 4)
    assert recognise(C, st) == ("spmm", (_lib.GEN_MERGE, 128, 64, 32, 0, 4))


def test_spgemm_recognised():
    i, j, k = taco.IndexVar("i"), taco.IndexVar("j"), taco.IndexVar("k")
    A = taco.Tensor("A", [30, 20], taco.CSR)
    B = taco.Tensor("B", [20, 25], taco.CSR)
    C = taco.Tensor("C", [30, 25], taco.CSR)
    C[i, k] = A[i, j] * B[j, k]
    assert recognise(C, C.get_assignment().concretize())[0] == "spgemm"


def test_unsupported_schedule_and_statement_raise():
    A, x, y, i, j, pre = spmv_setup()
    st = y.get_assignment().concretize()
    o, n = taco.IndexVar("o"), taco.IndexVar("n")
    bad = st.split(j, o, n, 4).parallelize(o, taco.ParallelUnit.GPUBlock, taco.OutputRaceStrategy.NoRaces)
    with pytest.raises(taco.TacoError, match="schedule not supported"):
        recognise(y, bad)
    # rows schedule whose thread loop asks for atomics instead of a temporary
    wrong = S.schedule_spmv_rows_gpu(st, A, pre, i, j)
    wrong.commands[-1] = ("parallelize", wrong.commands[-1][1], taco.ParallelUnit.GPUThread,
                          taco.OutputRaceStrategy.Atomics)
    with pytest.raises(taco.TacoError):
        recognise(y, wrong)
    z = taco.Tensor("z", [40], taco.Format([taco.Dense]))
    w = taco.Tensor("w", [50], taco.Format([taco.Dense]))
    w[i] = x[j] * z[j]
    with pytest.raises(taco.TacoError, match="CSR"):
        taco._statement_kind(w.get_assignment())


def test_pack_and_state_errors():
    A = taco.Tensor("A", [3, 4], taco.CSR)
    A.insert([2, 1], 5.0)
    A.insert([0, 3], 1.0)
    A.insert([2, 0], 2.0)
    A.pack()
    pos, crd, vals = A.to_csr()
    assert pos.tolist() == [0, 1, 1, 3] and crd.tolist() == [3, 0, 1]
    assert vals.tolist() == [1.0, 2.0, 5.0]
    y = taco.Tensor("y", [3])
    with pytest.raises(taco.TacoError, match="compile"):
        y.compute()
    with pytest.raises(taco.TacoError, match="format"):
        taco.Tensor("T", [2, 2, 2], taco.Format([taco.Dense, taco.Sparse, taco.Sparse]))


def _dense(name, values, dtype=taco.Float64):
    t = taco.Tensor(name, [len(values)], taco.Format([taco.Dense]), dtype)
    t.set_values(np.array(values))
    return t


def _csr(name, shape, pos, crd, vals):
    t = taco.Tensor(name, shape, taco.CSR)
    t.set_csr(np.array(pos, np.int32), np.array(crd, np.int32), np.array(vals, np.float64))
    return t


def test_equals_scalar_rule():
    """scalarEquals (tensor.cpp:1038-1045): |(a - b) / a| <= 1e-5, a divides."""
    a = _dense("a", [1.0, -4.0, 0.0, 2.5])
    assert taco.equals(a, _dense("b", [1.0, -4.0, 0.0, 2.5]))
    assert taco.equals(a, _dense("b", [1.0 + 5e-6, -4.0 * (1 - 5e-6), 0.0, 2.5]))
    assert not taco.equals(a, _dense("b", [1.0 + 2e-5, -4.0, 0.0, 2.5]))
    # 0/0 is NaN and passes; x/0 is infinite and fails; b = 0 gives |1| > 1e-5
    assert not taco.equals(a, _dense("b", [1.0, -4.0, 1e-300, 2.5]))
    assert not taco.equals(_dense("a", [1e-300]), _dense("b", [0.0]))
    # a divides, so the rule is not symmetric near the bound:
    # 9.99995e-6 / 1 passes, 9.99995e-6 / (1 - 9.99995e-6) does not
    assert taco.equals(_dense("a", [1.0]), _dense("b", [1.0 - 9.99995e-6]))
    assert not taco.equals(_dense("a", [1.0 - 9.99995e-6]), _dense("b", [1.0]))


def test_equals_shape_and_type():
    a = _dense("a", [1.0, 2.0])
    assert not taco.equals(a, _dense("b", [1.0, 2.0, 0.0]))
    assert not taco.equals(a, _dense("b", [1.0, 2.0], taco.Float32))


def test_equals_merge_skips_stored_zeros():
    """equalsTyped (tensor.cpp:1056-1100): a component stored in one tensor
    only must be zero; explicit zeros are skipped."""
    a = _csr("a", [2, 3], [0, 2, 3], [0, 2, 1], [5.0, 0.0, 7.0])
    b = _csr("b", [2, 3], [0, 1, 2], [0, 1], [5.0, 7.0])
    assert taco.equals(a, b) and taco.equals(b, a)
    c = _csr("c", [2, 3], [0, 2, 3], [0, 2, 1], [5.0, 1.0, 7.0])
    assert not taco.equals(c, b) and not taco.equals(b, c)
    # the same values at different coordinates
    d = _csr("d", [2, 3], [0, 1, 2], [1, 1], [5.0, 7.0])
    assert not taco.equals(d, b)
    # trailing stored zeros are fine, trailing non-zeros are not
    e = _csr("e", [2, 3], [0, 1, 3], [0, 1, 2], [5.0, 7.0, 0.0])
    assert taco.equals(e, b)
    f = _csr("f", [2, 3], [0, 1, 3], [0, 1, 2], [5.0, 7.0, 3.0])
    assert not taco.equals(f, b)


def test_codegen_toggle_as_pytaco_exposes_it():
    """should_use_cuda_codegen / set_cuda_codegen_enabled (pytaco.cpp:14,
    cuda.h:15-28) forward to the HIP runtime; with GPU code generation off
    the facade, which has no CPU code path, refuses to compile."""
    was = taco.should_use_cuda_codegen()
    try:
        A, x, y, i, j, pre = spmv_setup()
        taco.set_cuda_codegen_enabled(False)
        assert not taco.should_use_cuda_codegen() and not taco.should_use_hip_codegen()
        with pytest.raises(taco.TacoError, match="code generation is off"):
            y.compile()
        taco.set_hip_codegen_enabled(True)
        assert taco.should_use_cuda_codegen()
    finally:
        taco.set_cuda_codegen_enabled(was)
