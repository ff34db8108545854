"""taco_hip_spmv_compute_noalloc, the SpMV entry point that allocates
nothing: called through the C ABI, its results equal taco_hip_spmv_compute's
bit for bit under the deterministic merge-path schedule (whose carries live
in the static scratch) and pass the FP64 oracle, on two calls in a row, the
second with a smaller matrix; the other schedules pass through; a problem
beyond the scratch limit is rejected before anything is launched."""
import math

import numpy as np
import pytest
import torch

import oracle
from taco_amd import _lib, descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def run(entry, a, xd, m, sched, dtype):
    tdt = torch.float64 if dtype == "f64" else torch.float32
    y = torch.full((m,), math.nan, dtype=tdt, device="cuda")
    _lib.check(entry(D.dense(y, dtype).ptr, a.ptr, xd.ptr, sched), "spmv entry")
    torch.cuda.synchronize()
    return y.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_equals_allocating_entry(dtype):
    L = _lib.lib()
    tdt = torch.float64 if dtype == "f64" else torch.float32
    for m, seed, items in ((400_000, 71, 8), (30_000, 72, 4)):  # second call smaller
        print(f"no-alloc SpMV {dtype} m={m} items={items}", flush=True)
        A = W.csr_powerlaw(m, m, seed=seed, kmax=20000)
        x = W.uniform(A.n, seed, W.STREAM_X)
        a = G.csr_desc(A, dtype)
        xd = D.dense(G.dev(x, tdt), dtype)
        s = G.sched(_lib.MERGE_PATH, 256, items)
        want = run(L.taco_hip_spmv_compute, a, xd, A.m, s, dtype)
        got = run(L.taco_hip_spmv_compute_noalloc, a, xd, A.m, s, dtype)
        assert np.array_equal(got, want)
        xr = x if dtype == "f64" else x.astype(np.float32).astype(np.float64)
        Ar = A if dtype == "f64" else W.Csr(A.m, A.n, A.pos, A.crd,
                                            A.vals.astype(np.float32).astype(np.float64))
        ye, cond = oracle.exact_spmv(Ar, xr)
        rep = oracle.error_report(got, ye, cond, TOL[dtype])
        assert rep["ok"], rep


def test_noalloc_other_schedules_pass_through():
    L = _lib.lib()
    A = W.csr_powerlaw(50_000, 40_000, seed=73, kmax=5000)
    x = W.uniform(A.n, 73, W.STREAM_X)
    a, xd = G.csr_desc(A), D.dense(G.dev(x))
    ye, cond = oracle.exact_spmv(A, x)
    for s in (G.sched(_lib.THREAD_PER_ROW, 256), G.sched(_lib.GROUP_PER_ROW, 256, group=16),
              G.sched(_lib.NNZ_SPLIT, 256, 8), None):
        got = run(L.taco_hip_spmv_compute_noalloc, a, xd, A.m, s, "f64")
        assert oracle.error_report(got, ye, cond, 1e-12)["ok"]


def test_noalloc_rejects_problems_beyond_its_scratch():
    """45M rows + entries at 2 items per thread need 87,891 workgroup
    carries of 12 bytes, above the 1 MiB scratch: an error, nothing
    launched, and the entry point still works afterwards."""
    L = _lib.lib()
    A = W.csr_fixed(1_000_000, 1_000_000, 44, 74)
    x = W.uniform(A.n, 74, W.STREAM_X)
    a, xd = G.csr_desc(A), D.dense(G.dev(x))
    y = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    rc = L.taco_hip_spmv_compute_noalloc(D.dense(y).ptr, a.ptr, xd.ptr,
                                         G.sched(_lib.MERGE_PATH, 256, 2))
    assert rc != 0
    assert "limit is 1048576 bytes" in L.taco_hip_last_error().decode()
    torch.cuda.synchronize()
    assert not y.any()  # nothing ran
    got = run(L.taco_hip_spmv_compute_noalloc, a, xd, A.m, G.sched(_lib.MERGE_PATH, 256, 8),
              "f64")
    ye, cond = oracle.exact_spmv(A, x)
    assert oracle.error_report(got, ye, cond, 1e-12)["ok"]
