"""Modules emitted by CodeGen_HIP against the oracle and the hand-written
kernels (same results within the value tolerance; bit-equal on the
reference tests' integer data; SpGEMM patterns bit-equal)."""
import os

import numpy as np
import pytest
import torch

import oracle
from taco_amd import _lib, codegen, descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G
from tests import known_answers as KA

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SPMV = ["spmv_default_f64", "spmv_rows_f64", "spmv_rows64_f64", "spmv_nnz_f64", "spmv_merge_f64",
        "spmv_block_f64"]
SPMM = ["spmm_default_f64", "spmm_nnz_f64", "spmm_merge_f64", "spmm_rows_f64"]


def mats():
    yield "powerlaw", W.csr_powerlaw(20000, 20000, seed=31)
    counts = W.uniform_counts(3000, 0, 4, 32)
    yield "empty_rows", W.csr_from_counts(3000, 2000, counts, 32)
    yield "long_row", W.csr_from_counts(3, 100000, np.array([5, 90000, 0], np.int32), 33)


MATS = list(mats())
KERNELS = {}


def kernel(name):
    if name not in KERNELS:
        KERNELS[name] = codegen.Kernel(codegen.standard(name))
    return KERNELS[name]


def run_spmv(name, A, x):
    a = G.csr_desc(A)
    xd = G.dev(x)
    y = torch.full((A.m,), float("nan"), dtype=torch.float64, device="cuda")
    yd = D.dense(y)
    kernel(name).compute(yd, a, D.dense(xd))
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("name", SPMV)
@pytest.mark.parametrize("mname,A", MATS, ids=[m for m, _ in MATS])
def test_generated_spmv(name, mname, A):
    print(f"generated {name} on {mname}", flush=True)
    x = W.uniform(A.n, 3, W.STREAM_X)
    y = run_spmv(name, A, x)
    ye, cond = oracle.exact_spmv(A, x)
    rep = oracle.error_report(y, ye, cond, 1e-12)
    assert rep["ok"], rep
    hand = G.spmv(A, x, G.sched(_lib.MERGE_PATH, 256, 8))
    assert oracle.error_report(y, hand, cond, 2e-12)["ok"]


@pytest.mark.parametrize("block", [32, 64, 128, 1024])
def test_generated_block_rows_sizes(block):
    """The block reduction's wave-only (<= 64) and LDS (> 64) forms."""
    name = f"spmv_block{block}"
    if name not in KERNELS:
        KERNELS[name] = codegen.Kernel(codegen.spec(_lib.OP_SPMV, _lib.GEN_BLOCK_ROWS,
                                                    block_size=block))
    for mname, A in MATS:
        print(f"generated block rows {block} on {mname}", flush=True)
        x = W.uniform(A.n, 5, W.STREAM_X)
        y = run_spmv(name, A, x)
        ye, cond = oracle.exact_spmv(A, x)
        assert oracle.error_report(y, ye, cond, 1e-12)["ok"], (mname, block)


@pytest.mark.parametrize("name", SPMV)
def test_generated_spmv_reference_inputs_bit_exact(name):
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    for tag in ("spmvgpu", "spmvcpu"):
        A = W.Csr(102, 103, d[f"{tag}_pos"], d[f"{tag}_crd"], d[f"{tag}_vals"])
        assert np.array_equal(run_spmv(name, A, d[f"{tag}_x"]), d[f"{tag}_y"]), tag


def test_generated_evaluate_allocates_result():
    p, c, v = KA.coo_to_csr(5, 5, KA.MVM_B)
    A = W.Csr(5, 5, p, c, v)
    x = np.zeros(5)
    for k, val in KA.MVM_C.items():
        x[k] = val
    a = G.csr_desc(A)
    xd = G.dev(x)
    yd = D.Desc([5], [D.DENSE], 64)
    yd._sz = __import__("ctypes").c_int32(5)
    kernel("spmv_default_f64").evaluate(yd, a, D.dense(xd))
    out = np.empty(5)
    L = _lib.lib()
    import ctypes
    _lib.check(L.taco_hip_memcpy(out.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_void_p(yd.vals_ptr()), 40, 1))
    L.taco_hip_unified_free(ctypes.c_void_p(yd.vals_ptr()))
    assert out.tolist() == [2, 0, 0, 8, 0]


@pytest.mark.parametrize("name", SPMM)
@pytest.mark.parametrize("k", [64, 128, 20])
def test_generated_spmm(name, k):
    A = W.csr_powerlaw(5000, 4000, seed=34, kmax=3000)
    B = W.uniform(A.n * k, 34, W.STREAM_B).reshape(A.n, k)
    a = G.csr_desc(A)
    Bd = G.dev(B)
    C = torch.full((A.m, k), float("nan"), dtype=torch.float64, device="cuda")
    kernel(name).compute(D.dense(C), a, D.dense(Bd))
    torch.cuda.synchronize()
    Ce, cond = oracle.exact_spmm(A, B)
    rep = oracle.error_report(C.cpu().numpy(), Ce, cond, 1e-12)
    assert rep["ok"], rep


@pytest.mark.parametrize("name", ["spmv_default_f32", "spmv_nnz_f32"])
def test_generated_spmv_f32(name):
    """f32 modules, including the nnz-split full-chunk path, against the FP64
    oracle on the f32-rounded operands (tolerance 1e-5)."""
    for mname, A in MATS:
        print(f"generated {name} on {mname}", flush=True)
        A32 = W.Csr(A.m, A.n, A.pos, A.crd, A.vals.astype(np.float32).astype(np.float64))
        x = W.uniform(A.n, 6, W.STREAM_X).astype(np.float32).astype(np.float64)
        a = G.csr_desc(A32, "f32")
        xd = G.dev(x, torch.float32)
        y = torch.full((A.m,), float("nan"), dtype=torch.float32, device="cuda")
        kernel(name).compute(D.dense(y, "f32"), a, D.dense(xd, "f32"))
        torch.cuda.synchronize()
        ye, cond = oracle.exact_spmv(A32, x)
        rep = oracle.error_report(y.cpu().numpy().astype(np.float64), ye, cond, 1e-5)
        assert rep["ok"], (mname, rep)


@pytest.mark.parametrize("k", [32, 128])
def test_generated_spmm_f32(k):
    A = W.csr_powerlaw(5000, 4000, seed=35, kmax=3000)
    A32 = W.Csr(A.m, A.n, A.pos, A.crd, A.vals.astype(np.float32).astype(np.float64))
    B = W.uniform(A.n * k, 35, W.STREAM_B).reshape(A.n, k).astype(np.float32).astype(np.float64)
    a = G.csr_desc(A32, "f32")
    Bd = G.dev(B, torch.float32)
    C = torch.full((A.m, k), float("nan"), dtype=torch.float32, device="cuda")
    kernel("spmm_nnz_f32").compute(D.dense(C, "f32"), a, D.dense(Bd, "f32"))
    torch.cuda.synchronize()
    Ce, cond = oracle.exact_spmm(A32, B)
    rep = oracle.error_report(C.cpu().numpy().astype(np.float64), Ce, cond, 1e-5)
    assert rep["ok"], rep


def test_generated_spmm_reference_inputs_bit_exact():
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    A = W.Csr(102, 103, d["spmmgpu_pos"], d["spmmgpu_crd"], d["spmmgpu_vals"])
    for name in SPMM:
        a = G.csr_desc(A)
        Bd = G.dev(d["spmmgpu_B"])
        C = torch.zeros((102, 128), dtype=torch.float64, device="cuda")
        kernel(name).compute(D.dense(C), a, D.dense(Bd))
        torch.cuda.synchronize()
        assert np.array_equal(C.cpu().numpy(), d["spmmgpu_C"]), name


def run_spgemm(A, B):
    import ctypes
    a, b = G.csr_desc(A), G.csr_desc(B)
    c = D.csr_empty((A.m, B.n))
    k = kernel("spgemm_f64")
    k.assemble(c, a, b)
    k.compute(c, a, b)
    torch.cuda.synchronize()
    L = _lib.lib()
    nnz = c.t.vals_size
    pos, crd, vals = np.empty(A.m + 1, np.int32), np.empty(nnz, np.int32), np.empty(nnz)
    for h, p in ((pos, c.level_ptr(1, 0)), (crd, c.level_ptr(1, 1)), (vals, c.vals_ptr())):
        if h.nbytes:
            _lib.check(L.taco_hip_memcpy(h.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), h.nbytes, 1))
    for p in (c.level_ptr(1, 0), c.level_ptr(1, 1), c.vals_ptr()):
        L.taco_hip_unified_free(ctypes.c_void_p(p))
    return pos, crd, vals


def test_generated_spgemm():
    A = W.csr_fixed(3000, 2500, 8, 35)
    B = W.csr_powerlaw(2500, 2000, seed=36, kmax=100)
    pos, crd, vals = run_spgemm(A, B)
    opos, ocrd = oracle.spgemm_assemble(A, B)
    assert np.array_equal(pos, opos) and np.array_equal(crd, ocrd)
    ex, cond = oracle.exact_spgemm_values(A, B, opos, ocrd)
    assert oracle.error_report(vals, ex, cond, 1e-12)["ok"]
    hp, hc, hv = G.spgemm(A, B)
    assert np.array_equal(hp, pos) and np.array_equal(hc, crd) and np.array_equal(hv, vals)


def test_generated_spgemm_rows_strided_and_hashed():
    """More rows than the compute grid holds at once (4 workgroups of 4
    waves per CU), so rows are strided over the resident grid, and rows with
    more than 64 candidates, so one wave reuses its LDS hash accumulator
    across consecutive long rows."""
    A = W.csr_powerlaw(20000, 15000, seed=37, kmax=100)
    B = W.csr_fixed(15000, 12000, 8, 38)
    assert (np.diff(A.pos) * 8 > 64).sum() > 100
    pos, crd, vals = run_spgemm(A, B)
    opos, ocrd = oracle.spgemm_assemble(A, B)
    assert np.array_equal(pos, opos) and np.array_equal(crd, ocrd)
    ex, cond = oracle.exact_spgemm_values(A, B, opos, ocrd)
    assert oracle.error_report(vals, ex, cond, 1e-12)["ok"]
    hp, hc, hv = G.spgemm(A, B)
    assert np.array_equal(hp, pos) and np.array_equal(hc, crd) and np.array_equal(hv, vals)


@pytest.mark.parametrize("name", ["spmm_merge16_f64", "spmm_merge32_f64"])
@pytest.mark.parametrize("k", [20, 128])
def test_generated_merge_spmm_narrow_teams(name, k):
    """Merge-path SpMM with lane teams narrower than a wave: several teams
    share one wave and broadcast their entries through LDS permutes."""
    A = W.csr_powerlaw(5000, 4000, seed=39, kmax=3000)
    B = W.uniform(A.n * k, 39, W.STREAM_B).reshape(A.n, k)
    C = torch.full((A.m, k), float("nan"), dtype=torch.float64, device="cuda")
    kernel(name).compute(D.dense(C), G.csr_desc(A), D.dense(G.dev(B)))
    torch.cuda.synchronize()
    Ce, cond = oracle.exact_spmm(A, B)
    rep = oracle.error_report(C.cpu().numpy(), Ce, cond, 1e-12)
    assert rep["ok"], rep


def test_generated_spgemm_known_answer():
    p, c, v = KA.coo_to_csr(5, 5, KA.SPMM_B)
    A = W.Csr(5, 5, p, c, v)
    p, c, v = KA.coo_to_csr(5, 5, KA.SPMM_C)
    B = W.Csr(5, 5, p, c, v)
    pos, crd, vals = run_spgemm(A, B)
    assert pos.tolist() == KA.SPMM_EXPECTED_POS and crd.tolist() == KA.SPMM_EXPECTED_CRD
    assert vals.tolist() == [30.0, 180.0]
