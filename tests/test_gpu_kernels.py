"""Parity of the hand-written HIP kernels against the CPU oracle.

Values: condition-scaled gate |got - exact| <= tol * sum_j |a_ij x_j| with
tol = 1e-12 (f64), 1e-5 (f32), 1e-2 (bf16) against the extended-precision
oracle; integer-valued inputs (the reference tests' data) must be bit-equal
to taco's outputs.  SpGEMM pos/crd must be bit-equal to taco's assemble.
"""
import os

import numpy as np
import pytest

import oracle
from taco_amd import _lib
from taco_amd import workloads as W
from tests import gpu_util as G
from tests import known_answers as KA

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"f64": 1e-12, "f32": 1e-5, "bf16": 1e-2}

SPMV_SCHEDULES = [
    ("thread_per_row", G.sched(_lib.THREAD_PER_ROW, 256)),
    ("group8", G.sched(_lib.GROUP_PER_ROW, 256, group=8)),
    ("group64x2", G.sched(_lib.GROUP_PER_ROW, 256, group=64, rows=2)),
    ("nnz_split8", G.sched(_lib.NNZ_SPLIT, 256, items=8)),
    ("nnz_split1", G.sched(_lib.NNZ_SPLIT, 64, items=1)),
    ("merge4", G.sched(_lib.MERGE_PATH, 256, items=4)),
    ("merge8", G.sched(_lib.MERGE_PATH, 256, items=8)),
    ("merge16", G.sched(_lib.MERGE_PATH, 256, items=16)),
    ("default", None),
]


def matrices():
    yield "uniform", W.csr_fixed(3000, 2500, 10, 1)
    yield "powerlaw", W.csr_powerlaw(30000, 30000, seed=2)
    counts = W.uniform_counts(5000, 0, 3, 3)  # many empty rows
    yield "empty_rows", W.csr_from_counts(5000, 4000, counts, 3)
    yield "one_long_row", W.csr_from_counts(3, 200000,
                                            np.array([0, 150000, 1], np.int32), 4)
    yield "single", W.Csr(1, 1, np.array([0, 1], np.int32),
                          np.array([0], np.int32), np.array([2.5]))
    yield "no_entries", W.Csr(7, 5, np.zeros(8, np.int32),
                              np.zeros(0, np.int32), np.zeros(0))


MATS = list(matrices())


@pytest.mark.parametrize("name,s", SPMV_SCHEDULES, ids=[n for n, _ in SPMV_SCHEDULES])
@pytest.mark.parametrize("mname,A", MATS, ids=[n for n, _ in MATS])
def test_spmv_f64(name, s, mname, A):
    print(f"spmv f64 {name} {mname} m={A.m} nnz={A.nnz}", flush=True)
    x = W.uniform(A.n, 7, W.STREAM_X)
    y = G.spmv(A, x, s)
    ye, cond = oracle.exact_spmv(A, x)
    rep = oracle.error_report(y, ye, cond, TOL["f64"])
    assert rep["ok"], rep


@pytest.mark.parametrize("name,s", SPMV_SCHEDULES, ids=[n for n, _ in SPMV_SCHEDULES])
def test_spmv_f32(name, s):
    A = W.csr_powerlaw(20000, 20000, seed=5)
    x = W.uniform(A.n, 5, W.STREAM_X)
    A32 = W.Csr(A.m, A.n, A.pos, A.crd, A.vals.astype(np.float32).astype(np.float64))
    x32 = x.astype(np.float32).astype(np.float64)
    y = G.spmv(A, x, s, dtype="f32")
    ye, cond = oracle.exact_spmv(A32, x32)
    rep = oracle.error_report(y, ye, cond, TOL["f32"])
    assert rep["ok"], rep


def test_spmv_reference_test_inputs_bit_exact():
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    for tag in ("spmvgpu", "spmvcpu"):
        A = W.Csr(102, 103, d[f"{tag}_pos"], d[f"{tag}_crd"], d[f"{tag}_vals"])
        for name, s in SPMV_SCHEDULES:
            y = G.spmv(A, d[f"{tag}_x"], s)
            assert np.array_equal(y, d[f"{tag}_y"]), (tag, name)


def test_spmv_known_answers():
    p, c, v = KA.coo_to_csr(5, 5, KA.MVM_B)
    A = W.Csr(5, 5, p, c, v)
    x = np.zeros(5)
    for k, val in KA.MVM_C.items():
        x[k] = val
    for name, s in SPMV_SCHEDULES:
        assert G.spmv(A, x, s).tolist() == [2, 0, 0, 8, 0], name


def test_spmv_vals_size_zero_reads_pos():
    """Descriptors without vals_size: nnz comes from pos[m] on the device."""
    import torch
    from taco_amd import descriptors as D
    A = W.csr_powerlaw(4000, 4000, seed=9)
    x = W.uniform(A.n, 9, 1)
    a = G.csr_desc(A)
    a.t.vals_size = 0
    xd = G.dev(x)
    y = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().taco_hip_spmv_compute(D.dense(y).ptr, a.ptr, D.dense(xd).ptr, None))
    ye, cond = oracle.exact_spmv(A, x)
    assert oracle.error_report(y.cpu().numpy(), ye, cond, 1e-12)["ok"]


def test_spmv_rejects_bad_shapes():
    import torch
    from taco_amd import descriptors as D
    A = W.csr_fixed(10, 8, 2, 1)
    a = G.csr_desc(A)
    xd = torch.zeros(9, dtype=torch.float64, device="cuda")
    y = torch.zeros(10, dtype=torch.float64, device="cuda")
    rc = _lib.lib().taco_hip_spmv_compute(D.dense(y).ptr, a.ptr, D.dense(xd).ptr, None)
    assert rc != 0
    assert b"x has 9 entries" in _lib.lib().taco_hip_last_error()


SPMM_SCHEDULES = [
    ("thread_per_row", G.sched(_lib.THREAD_PER_ROW, 256)),
    ("group_rows", G.sched(_lib.GROUP_PER_ROW, 256)),
    ("group_rows4", G.sched(_lib.GROUP_PER_ROW, 256, rows=4)),
    ("nnz_split", G.sched(_lib.NNZ_SPLIT, 256, items=8)),
    ("merge32", G.sched(_lib.MERGE_PATH, 256, items=32)),
    ("merge7", G.sched(_lib.MERGE_PATH, 256, items=7)),
    ("default", None),
]


@pytest.mark.parametrize("name,s", SPMM_SCHEDULES, ids=[n for n, _ in SPMM_SCHEDULES])
@pytest.mark.parametrize("k", [64, 16, 100])
def test_spmm_f64(name, s, k):
    A = W.csr_powerlaw(6000, 5000, seed=11, kmax=4000)
    B = W.uniform(A.n * k, 11, W.STREAM_B).reshape(A.n, k)
    C = G.spmm(A, B, s)
    Ce, cond = oracle.exact_spmm(A, B)
    rep = oracle.error_report(C, Ce, cond, TOL["f64"])
    assert rep["ok"], rep


@pytest.mark.parametrize("name,s", SPMM_SCHEDULES, ids=[n for n, _ in SPMM_SCHEDULES])
def test_spmm_f32(name, s):
    A = W.csr_powerlaw(6000, 5000, seed=12, kmax=4000)
    B = W.uniform(A.n * 64, 12, W.STREAM_B).reshape(A.n, 64)
    A32 = W.Csr(A.m, A.n, A.pos, A.crd, A.vals.astype(np.float32).astype(np.float64))
    B32 = B.astype(np.float32).astype(np.float64)
    C = G.spmm(A, B, s, dtype="f32")
    Ce, cond = oracle.exact_spmm(A32, B32)
    rep = oracle.error_report(C, Ce, cond, TOL["f32"])
    assert rep["ok"], rep


BF16_SCHEDULES = SPMM_SCHEDULES[1:] + [("mfma", G.sched(_lib.MFMA))]


@pytest.mark.parametrize("name,s", BF16_SCHEDULES, ids=[n for n, _ in BF16_SCHEDULES])
def test_spmm_bf16(name, s):
    A = W.csr_powerlaw(6000, 5000, seed=13, kmax=4000)
    B = W.uniform(A.n * 64, 13, W.STREAM_B).reshape(A.n, 64)
    C = G.spmm(A, B, s, dtype="bf16")
    # Against the FP64 oracle on the original (unrounded) inputs.
    Ce, cond = oracle.exact_spmm(A, B)
    rep = oracle.error_report(C, Ce, cond, TOL["bf16"])
    assert rep["ok"], rep
    # And tightly against the oracle on the bf16-rounded inputs (f32 accumulate).
    Ab = W.Csr(A.m, A.n, A.pos, A.crd, G.bf16_values(A.vals))
    Cb, condb = oracle.exact_spmm(Ab, G.bf16_values(B))
    rep = oracle.error_report(C, Cb, condb, 1e-5)
    assert rep["ok"], rep


def test_spmm_reference_test_inputs_bit_exact():
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    A = W.Csr(102, 103, d["spmmgpu_pos"], d["spmmgpu_crd"], d["spmmgpu_vals"])
    for name, s in SPMM_SCHEDULES:
        C = G.spmm(A, d["spmmgpu_B"], s)
        assert np.array_equal(C, d["spmmgpu_C"]), name


def test_spmm_mfma_integer_exact():
    """Small-integer bf16 operands: every product and sum is exact in f32."""
    A = W.csr_powerlaw(3000, 3000, seed=14, kmax=3000)
    A.vals = np.round(A.vals * 4)
    B = np.round(W.uniform(A.n * 64, 14, 2).reshape(A.n, 64) * 4)
    C = G.spmm(A, B, G.sched(_lib.MFMA), dtype="bf16")
    assert np.array_equal(C, oracle.spmm(A, B))


def spgemm_cases():
    yield "fixed8", W.csr_fixed(4000, 3500, 8, 21), W.csr_fixed(3500, 3000, 8, 22)
    yield "long_rows", W.csr_powerlaw(1500, 1200, seed=23, kmax=300), \
        W.csr_fixed(1200, 2000, 3, 24)
    yield "empty", W.Csr(4, 3, np.zeros(5, np.int32), np.zeros(0, np.int32),
                         np.zeros(0)), W.csr_fixed(3, 6, 2, 25)
    # A rows longer than 64 entries over mostly empty B rows: few candidates
    yield "long_a_empty_b", W.csr_fixed(600, 2000, 80, 26), W.csr_from_counts(
        2000, 900, np.where(np.arange(2000) % 50 == 0, 3, 0).astype(np.int32), 27)
    # short A rows over B rows that are often empty and sometimes longer than 64
    counts = np.random.default_rng(29).integers(0, 12, 2500).astype(np.int32)
    counts[np.random.default_rng(30).random(2500) < 0.4] = 0
    counts[::97] = 70
    yield "gappy_b", W.csr_fixed(3000, 2500, 6, 28), W.csr_from_counts(2500, 3000, counts, 31)
    # C rows of more than 1,024 distinct columns (beyond the LDS row set):
    # sorted in global memory when assembling, accumulated in C's values
    # when computing; mixed with short, hashed and empty rows
    ca = np.random.default_rng(32).integers(0, 12, 2000).astype(np.int32)
    ca[::101] = 1500
    ca[5::257] = 0
    yield "over_row_cap", W.csr_from_counts(2000, 30000, ca, 33), W.csr_fixed(30000, 40000, 4, 34)
    # Short rows sort (column, encounter) keys in 32 bits while every column
    # is below 2^26 and in 64 bits otherwise: the first 1,000 rows of A name
    # B rows whose columns are below 2^20, the rest B rows spread over 2^27
    # columns.
    yield "wide_columns", _stack_rows(W.csr_fixed(1000, 1500, 8, 35),
                                      W.csr_fixed(1000, 1500, 8, 36), 1500), \
        _stack_rows(W.csr_fixed(1500, 1 << 20, 8, 37), W.csr_fixed(1500, (1 << 27) + 5, 8, 38))


def _stack_rows(top, bottom, bottom_col_offset=0):
    """The rows of `top` above those of `bottom`, whose columns move right by
    bottom_col_offset; the width is the larger of the two."""
    pos = np.concatenate([top.pos, bottom.pos[1:] + top.pos[-1]]).astype(np.int32)
    crd = np.concatenate([top.crd, bottom.crd + bottom_col_offset]).astype(np.int32)
    n = max(top.n, bottom.n + bottom_col_offset)
    return W.Csr(top.m + bottom.m, n, pos, crd, np.concatenate([top.vals, bottom.vals]))


@pytest.mark.parametrize("name,A,B", list(spgemm_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_spgemm(name, A, B):
    pos, crd, vals = G.spgemm(A, B)
    opos, ocrd = oracle.spgemm_assemble(A, B)
    assert np.array_equal(pos, opos) and np.array_equal(crd, ocrd)
    ovals = oracle.spgemm_compute(A, B, opos, ocrd)
    ex, cond = oracle.exact_spgemm_values(A, B, opos, ocrd)
    rep = oracle.error_report(vals, ex, cond, TOL["f64"])
    assert rep["ok"], rep
    if name == "over_row_cap":
        assert np.diff(pos).max() > 4000  # the global-memory row path ran
    if name in ("fixed8", "long_a_empty_b", "wide_columns"):
        # Short rows sum duplicates sequentially in taco's encounter order,
        # with separately rounded products: equal to taco's values bit for bit.
        assert np.array_equal(vals, ovals)


@pytest.mark.parametrize("seed", range(12))
def test_spgemm_random_shapes(seed):
    """Seeded sweep of row-length mixes around the 64-candidate boundary: A
    rows of 0..100 entries, B rows that are often empty or longer than 64."""
    rng = np.random.default_rng(100 + seed)
    m, k, n = (int(v) for v in rng.integers(1, 1500, 3))
    ca = rng.integers(0, 12, m).astype(np.int32)
    ca[rng.random(m) < 0.05] = rng.integers(60, 101)
    cb = rng.integers(0, 10, k).astype(np.int32)
    cb[rng.random(k) < rng.random()] = 0
    cb[rng.random(k) < 0.02] = 70
    A = W.csr_from_counts(m, k, np.minimum(ca, k), 200 + seed)
    B = W.csr_from_counts(k, n, np.minimum(cb, n), 300 + seed)
    print(f"spgemm sweep seed={seed} m={m} k={k} n={n} nnzA={A.nnz} nnzB={B.nnz}", flush=True)
    pos, crd, vals = G.spgemm(A, B)
    opos, ocrd = oracle.spgemm_assemble(A, B)
    assert np.array_equal(pos, opos) and np.array_equal(crd, ocrd)
    ex, cond = oracle.exact_spgemm_values(A, B, opos, ocrd)
    assert oracle.error_report(vals, ex, cond, TOL["f64"])["ok"]


def test_spgemm_known_answer():
    p, c, v = KA.coo_to_csr(5, 5, KA.SPMM_B)
    A = W.Csr(5, 5, p, c, v)
    p, c, v = KA.coo_to_csr(5, 5, KA.SPMM_C)
    B = W.Csr(5, 5, p, c, v)
    pos, crd, vals = G.spgemm(A, B)
    assert pos.tolist() == KA.SPMM_EXPECTED_POS
    assert crd.tolist() == KA.SPMM_EXPECTED_CRD
    assert vals.tolist() == [30.0, 180.0]


def test_spgemm_golden():
    d = np.load(os.path.join(GOLD, "seeded.npz"))
    A = W.Csr(len(d["g_apos"]) - 1, len(d["g_bpos"]) - 1, d["g_apos"],
              d["g_acrd"], d["g_avals"])
    B = W.Csr(len(d["g_bpos"]) - 1, int(d["g_bn"][0]), d["g_bpos"], d["g_bcrd"],
              d["g_bvals"])
    pos, crd, vals = G.spgemm(A, B)
    assert np.array_equal(pos, d["g_cpos"]) and np.array_equal(crd, d["g_ccrd"])
    ex, cond = oracle.exact_spgemm_values(A, B, pos, crd)
    assert oracle.error_report(vals, ex, cond, 1e-12)["ok"]
