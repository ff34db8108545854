"""Slab-tiled SpMV plan against the extended-precision oracle."""
import numpy as np
import pytest
import torch

import oracle
from taco_amd import _lib, descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def run(A, x, parts=0, slab=0, unroll=0, threads=0, form=2):
    a = G.csr_desc(A)
    plan = _lib.SpmvPlan(a, parts, slab)
    plan.configure(threads, form)
    y = torch.full((A.m,), float("nan"), dtype=torch.float64, device="cuda")
    xd = G.dev(x)
    plan.compute(D.dense(y), D.dense(xd), unroll)
    torch.cuda.synchronize()
    return y.cpu().numpy(), plan.info()


CASES = [
    ("powerlaw", lambda: W.csr_powerlaw(30000, 30000, seed=41), 0, 4096),
    ("powerlaw_many_parts", lambda: W.csr_powerlaw(30000, 30000, seed=42), 3000, 1000),
    ("one_slab", lambda: W.csr_powerlaw(20000, 20000, seed=43), 64, 0),
    ("long_rows", lambda: W.csr_from_counts(10, 300000, np.array([0, 250000, 3, 0, 0, 40000, 1, 0, 0, 7], np.int32), 44), 16, 50000),
    ("empty_rows", lambda: W.csr_from_counts(50000, 9000, W.uniform_counts(50000, 0, 2, 45), 45), 40, 2048),
    ("no_entries", lambda: W.Csr(9, 4, np.zeros(10, np.int32), np.zeros(0, np.int32), np.zeros(0)), 4, 0),
    ("leading_trailing_empty", lambda: W.csr_from_counts(
        12000, 7000, np.concatenate([np.zeros(3000, np.int32), W.uniform_counts(6000, 0, 40, 46),
                                     np.zeros(3000, np.int32)]), 46), 16, 1024),
]


# (entries per thread, threads, kernel form: 0 per-thread runs, 1 wave
# segmented sums, 2 interleaved entries)
FORMS = [(1, 256, 1), (2, 256, 1), (4, 512, 1), (8, 256, 1), (1, 256, 0), (4, 256, 0),
         (8, 512, 0), (1, 256, 2), (4, 256, 2), (8, 256, 2), (8, 512, 2), (0, 0, 2)]


@pytest.mark.parametrize("name,make,parts,slab", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("unroll,threads,form", FORMS,
                         ids=[f"u{u}t{t}f{f}" for u, t, f in FORMS])
def test_slab_spmv(name, make, parts, slab, unroll, threads, form):
    A = make()
    print(f"slab spmv {name} m={A.m} nnz={A.nnz} parts={parts} slab={slab} unroll={unroll} "
          f"threads={threads} form={form}", flush=True)
    x = W.uniform(A.n, 5, W.STREAM_X)
    y, info = run(A, x, parts, slab, unroll, threads, form)
    ye, cond = oracle.exact_spmv(A, x)
    rep = oracle.error_report(y, ye, cond, 1e-12)
    assert rep["ok"], (rep, info)


def test_slab_block_boundaries():
    """Rows whose runs start and end at every lane position of a 64-entry
    block and across the 16-lane rows of the wave (run lengths 1..70)."""
    counts = np.concatenate([np.arange(1, 71), np.arange(70, 0, -1), [0, 0, 129, 1, 63, 65]])
    A = W.csr_from_counts(len(counts), 5000, counts.astype(np.int32), 47)
    x = W.uniform(A.n, 6, W.STREAM_X)
    ye, cond = oracle.exact_spmv(A, x)
    for unroll, threads in [(1, 256), (2, 512)]:
        y, info = run(A, x, 4, 1024, unroll, threads, 1)
        rep = oracle.error_report(y, ye, cond, 1e-12)
        assert rep["ok"], (unroll, threads, rep, info)


def test_slab_plan_rejects_f32():
    A = W.csr_fixed(100, 100, 3, 1)
    with pytest.raises(_lib.TacoHipError, match="f64"):
        _lib.SpmvPlan(G.csr_desc(A, "f32"))


def test_slab_plan_library_defaults():
    """The plan as the library builds it by default: one part per CU,
    segments ordered by column, interleaved entries, 8 per thread, 512
    threads."""
    A = W.csr_powerlaw(200000, 180000, seed=48, kmax=50000)
    x = W.uniform(A.n, 7, W.STREAM_X)
    a = G.csr_desc(A)
    plan = _lib.SpmvPlan(a)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert plan.info()["parts"] == cus
    assert plan.shape() == {"threads": 512, "form": 2, "unroll": 8}
    y = torch.full((A.m,), float("nan"), dtype=torch.float64, device="cuda")
    plan.compute(D.dense(y), D.dense(G.dev(x)))
    torch.cuda.synchronize()
    ye, cond = oracle.exact_spmv(A, x)
    rep = oracle.error_report(y.cpu().numpy(), ye, cond, 1e-12)
    assert rep["ok"], rep


def test_slab_plan_default_parts_grow_with_rows():
    """Rows beyond half the LDS capacity of one part per CU: the default
    takes the next multiple of the CU count (here two parts per CU, at 256
    threads), still within the FP64 oracle's tolerance."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    m = cus * 9984 + 1000
    A = W.csr_from_counts(m, 50000, W.uniform_counts(m, 0, 4, 49), 49)
    x = W.uniform(A.n, 8, W.STREAM_X)
    print(f"slab plan default parts m={m} nnz={A.nnz}", flush=True)
    plan = _lib.SpmvPlan(G.csr_desc(A))
    assert plan.info()["parts"] == 2 * cus
    assert plan.shape()["threads"] == 256
    y = torch.full((A.m,), float("nan"), dtype=torch.float64, device="cuda")
    plan.compute(D.dense(y), D.dense(G.dev(x)))
    torch.cuda.synchronize()
    ye, cond = oracle.exact_spmv(A, x)
    assert oracle.error_report(y.cpu().numpy(), ye, cond, 1e-12)["ok"]
