"""The exhaustive autotuner on one matrix: every candidate passes the output
check, the ranking is sorted, and the winner's result meets the oracle."""
import math

import pytest
import torch

import oracle
from taco_amd import autotune, descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def test_tune_spmv_ranks_checked_candidates():
    A = W.csr_powerlaw(30000, 30000, seed=61, kmax=5000)
    x = W.uniform(A.n, 61, W.STREAM_X)
    print(f"tune_spmv m={A.m} nnz={A.nnz}", flush=True)
    a, xd = G.csr_desc(A), D.dense(G.dev(x))
    y = torch.full((A.m,), math.nan, dtype=torch.float64, device="cuda")
    res = autotune.tune_spmv(a, xd, D.dense(y), reps=2)
    assert len(res) == len(autotune.SPMV_CANDIDATES)
    assert all(math.isfinite(t) for _, t in res), res
    assert [t for _, t in res] == sorted(t for _, t in res)
    ye, cond = oracle.exact_spmv(A, x)
    for name, _ in res[:3]:  # the fastest three, run again on their own
        s = dict(autotune.SPMV_CANDIDATES)[name]
        y.fill_(math.nan)
        if s is None:  # the slab plan
            plan, build_ms = autotune.slab_plan(a)
            assert plan is not None and build_ms > 0
            plan.compute(D.dense(y), xd)
            torch.cuda.synchronize()
            yy = y.cpu().numpy()
        else:
            yy = G.spmv(A, x, s)
        assert oracle.error_report(yy, ye, cond, 1e-12)["ok"], name


def test_slab_plan_candidate_charges_its_build():
    """The plan's candidate time is its per-call time plus build time /
    PLAN_CALLS; on a f32 matrix the plan does not apply (inf)."""
    A = W.csr_powerlaw(20000, 20000, seed=63, kmax=2000)
    x = W.uniform(A.n, 63, W.STREAM_X)
    a, xd = G.csr_desc(A), D.dense(G.dev(x))
    y = torch.full((A.m,), math.nan, dtype=torch.float64, device="cuda")
    only = [c for c in autotune.SPMV_CANDIDATES if c[0] in ("thread_per_row", "slab_plan")]
    res = dict(autotune.tune_spmv(a, xd, D.dense(y), reps=3, candidates=only))
    plan, build_ms = autotune.slab_plan(a)
    per_call = autotune._time(lambda: plan.compute(D.dense(y), xd), 5)
    assert res["slab_plan"] > build_ms / autotune.PLAN_CALLS
    assert abs(res["slab_plan"] - (per_call + build_ms / autotune.PLAN_CALLS)) < 0.5 * res["slab_plan"]
    a32 = G.csr_desc(A, "f32")
    assert autotune.slab_plan(a32) == (None, float("inf"))


@pytest.mark.parametrize("K", [16, 64])
def test_tune_spmm_ranks_checked_candidates(K):
    A = W.csr_powerlaw(8000, 6000, seed=62, kmax=2000)
    B = W.uniform(A.n * K, 62, W.STREAM_B).reshape(A.n, K)
    print(f"tune_spmm K={K} nnz={A.nnz}", flush=True)
    a, bd = G.csr_desc(A), D.dense(G.dev(B))
    C = torch.full((A.m, K), math.nan, dtype=torch.float64, device="cuda")
    res = autotune.tune_spmm(a, bd, D.dense(C), K, reps=2)
    assert len(res) == len(autotune.spmm_candidates(K))
    assert all(math.isfinite(t) for _, t in res), res
    Ce, cond = oracle.exact_spmm(A, B)
    # the last candidate run left its output in C; all passed the check
    assert oracle.error_report(C.cpu().numpy(), Ce, cond, 1e-12)["ok"]
