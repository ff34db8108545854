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


@pytest.fixture(scope="module")
def powerlaw_spmm():
    """The 2M power-law A with a dense 2M x 64 B, and the row sample the
    SpMM checks use: every 9973rd row plus the 50 longest rows."""
    A = W.powerlaw_2m()
    K = 64
    B = W.uniform(A.n * K, 2, W.STREAM_B).reshape(A.n, K)
    rows = np.unique(np.concatenate([np.arange(0, A.m, 9973), np.argsort(np.diff(A.pos))[-50:],
                                     [A.m - 1]]))
    return A, B, rows


def sample_rows(A, rows, vals):
    sub_pos = np.concatenate([[0], np.cumsum(np.diff(A.pos)[rows])]).astype(np.int32)
    sub_crd = np.concatenate([A.crd[A.pos[r]:A.pos[r + 1]] for r in rows])
    sub_val = np.concatenate([vals[A.pos[r]:A.pos[r + 1]] for r in rows])
    return W.Csr(len(rows), A.n, sub_pos, sub_crd, sub_val)


# (dtype, schedule, tolerance against the FP64 oracle of the unrounded
# inputs); bf16 is also held to the f32 tolerance on its rounded inputs.
FULL_SPMM = [("f64", "default", 1e-12), ("f32", "default", 1e-5), ("bf16", "mfma", 1e-2)]


@pytest.mark.parametrize("dtype,which,tol", FULL_SPMM, ids=[f[0] for f in FULL_SPMM])
def test_powerlaw_spmm_k64_row_sample(powerlaw_spmm, dtype, which, tol):
    """SpMM on the full 2M-row matrix with K = 64, in each precision form the
    benchmark times; checked on a row sample (the exact oracle of all 128M
    outputs would take minutes on the CPU)."""
    A, B, rows = powerlaw_spmm
    K = B.shape[1]
    print(f"full-size SpMM {dtype} ({which}), nnz={A.nnz}", flush=True)
    L = _lib.lib()
    s = _lib.Schedule()
    if which == "mfma":
        s = G.sched(_lib.MFMA)
    else:
        L.taco_hip_default_schedule(1, A.m, A.nnz, K, s)
    a = G.csr_desc(A, dtype)
    if dtype == "bf16":
        bdesc_t = G.bf16_bits(B).to("cuda")
        C = torch.full((A.m, K), float("nan"), dtype=torch.float32, device="cuda")
        bdesc, cdesc = D.dense(bdesc_t, "bf16"), D.dense(C, "f32")
    else:
        tdt = torch.float64 if dtype == "f64" else torch.float32
        bdesc_t = G.dev(B, tdt)
        C = torch.full((A.m, K), float("nan"), dtype=tdt, device="cuda")
        bdesc, cdesc = D.dense(bdesc_t, dtype), D.dense(C, dtype)
    _lib.check(L.taco_hip_spmm_compute(cdesc.ptr, a.ptr, bdesc.ptr, s))
    torch.cuda.synchronize()
    got = C[torch.from_numpy(rows).cuda()].cpu().numpy().astype(np.float64)
    del C, bdesc_t, a
    # the oracle of the inputs as given ...
    Ce, cond = oracle.exact_spmm(sample_rows(A, rows, A.vals), B)
    rep = oracle.error_report(got, Ce, cond, tol)
    assert rep["ok"], rep
    # ... and, for the reduced precisions, of the inputs as the GPU saw them
    if dtype != "f64":
        rnd = G.bf16_values if dtype == "bf16" else (lambda v: v.astype(np.float32).astype(np.float64))
        Cr, condr = oracle.exact_spmm(sample_rows(A, rows, rnd(A.vals)), rnd(B))
        rep = oracle.error_report(got, Cr, condr, 1e-5)
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


def test_spmm_mfma_b_beyond_4gib():
    """bf16 SpMM on the matrix cores with B spanning more than 4 GiB (2^26+64
    rows of 64 columns, 8.6 GB): row offsets of B no longer fit 32 bits and
    the kernel takes its 64-bit addressing.  Half of A's columns lie past
    row 2^25, where a 32-bit byte offset would wrap.  The oracle runs on the
    rows of B that A touches."""
    K = 64
    n = (1 << 26) + 64
    A = W.csr_fixed(3000, n, 8, 71)
    print("spmm bf16, B of", n, "rows", flush=True)
    Bd = torch.empty((n, K), dtype=torch.bfloat16, device="cuda").uniform_(-1.0, 1.0)
    a = G.csr_desc(A, "bf16")
    C = torch.full((A.m, K), float("nan"), dtype=torch.float32, device="cuda")
    bdesc, cdesc = D.dense(Bd.view(torch.int16), "bf16"), D.dense(C, "f32")
    _lib.check(_lib.lib().taco_hip_spmm_compute(cdesc.ptr, a.ptr, bdesc.ptr,
                                                G.sched(_lib.MFMA)))
    torch.cuda.synchronize()
    got = C.cpu().numpy().astype(np.float64)
    uniq = np.unique(A.crd)
    assert uniq[-1] >= (1 << 25)
    rows = Bd[torch.from_numpy(uniq.astype(np.int64)).cuda()].double().cpu().numpy()
    del bdesc, Bd
    torch.cuda.empty_cache()
    Ar = W.Csr(A.m, len(uniq), A.pos, np.searchsorted(uniq, A.crd).astype(np.int32),
               G.bf16_values(A.vals))
    Ce, cond = oracle.exact_spmm(Ar, rows)
    rep = oracle.error_report(got, Ce, cond, 1e-5)
    assert rep["ok"], rep
