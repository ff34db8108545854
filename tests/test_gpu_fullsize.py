"""Parity at the benchmark shapes: the 2M-row power-law SpMV (slab plan and
merge path) and SpMM against the FP64 oracle, and the 500k SpGEMM pattern
bit for bit against taco's two-phase CPU code (the oracle restatement)."""
import numpy as np
import pytest
import torch

import oracle
from taco_amd import _lib, descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def powerlaw():
    A = W.powerlaw_2m()
    x = W.uniform(A.n, 2, W.STREAM_X)
    ye, cond = oracle.exact_spmv(A, x)
    return A, x, ye, cond


def test_powerlaw_spmv_slab_plan(powerlaw):
    A, x, ye, cond = powerlaw
    print(f"slab plan on the 2M power-law matrix, nnz={A.nnz}", flush=True)
    a = G.csr_desc(A)
    plan = _lib.SpmvPlan(a)
    y = torch.full((A.m,), float("nan"), dtype=torch.float64, device="cuda")
    xd = G.dev(x)
    for _ in range(3):  # repeated computes on one plan
        plan.compute(D.dense(y), D.dense(xd))
        torch.cuda.synchronize()
        rep = oracle.error_report(y.cpu().numpy(), ye, cond, 1e-12)
        assert rep["ok"], rep


def test_powerlaw_spmv_merge_path(powerlaw):
    A, x, ye, cond = powerlaw
    y = G.spmv(A, x, G.sched(_lib.MERGE_PATH, 256, 8))
    assert oracle.error_report(y, ye, cond, 1e-12)["ok"]


def test_powerlaw_spmm_f32_k64_first_rows():
    """SpMM on the full matrix; checked on a row sample (the exact oracle of
    all 128M outputs would take minutes on the CPU)."""
    A = W.powerlaw_2m()
    K = 64
    B = W.uniform(A.n * K, 2, W.STREAM_B).reshape(A.n, K)
    a = G.csr_desc(A, "f32")
    Bd = G.dev(B, torch.float32)
    C = torch.full((A.m, K), float("nan"), dtype=torch.float32, device="cuda")
    s = _lib.Schedule()
    L = _lib.lib()
    L.taco_hip_default_schedule(1, A.m, A.nnz, K, s)
    _lib.check(L.taco_hip_spmm_compute(D.dense(C, "f32").ptr, a.ptr, D.dense(Bd, "f32").ptr, s))
    torch.cuda.synchronize()
    rows = np.unique(np.concatenate([np.arange(0, A.m, 9973), np.argsort(np.diff(A.pos))[-50:]]))
    sub_pos = np.concatenate([[0], np.cumsum(np.diff(A.pos)[rows])]).astype(np.int32)
    sub_crd = np.concatenate([A.crd[A.pos[r]:A.pos[r + 1]] for r in rows])
    sub_val = np.concatenate([A.vals[A.pos[r]:A.pos[r + 1]] for r in rows])
    As = W.Csr(len(rows), A.n, sub_pos, sub_crd, sub_val.astype(np.float32).astype(np.float64))
    Ce, cond = oracle.exact_spmm(As, B.astype(np.float32).astype(np.float64))
    got = C[torch.from_numpy(rows).cuda()].cpu().numpy()
    rep = oracle.error_report(got, Ce, cond, 1e-5)
    assert rep["ok"], rep


def test_spgemm_500k_pattern_bit_exact():
    A, B = W.spgemm_500k()
    print("spgemm 500k", flush=True)
    pos, crd, vals = G.spgemm(A, B)
    opos, ocrd = oracle.spgemm_assemble(A, B)
    assert np.array_equal(pos, opos) and np.array_equal(crd, ocrd)
    ovals = oracle.spgemm_compute(A, B, opos, ocrd)
    ex, cond = oracle.exact_spgemm_values(A, B, opos, ocrd)
    assert oracle.error_report(vals, ex, cond, 1e-12)["ok"]
    assert oracle.error_report(ovals, ex, cond, 1e-12)["ok"]
