"""Hand-written SpMV/SpMM calls captured in a HIP graph (through torch's
graph capture) and replayed, with eager work between the replays: every
replay matches the FP64 oracle.  The schedules that accumulate with atomics
clear their output with a kernel, not a memset, so that they can be
captured."""
import numpy as np
import pytest
import torch

import oracle
from taco_amd import _lib, descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def _capture(fn):
    """fn run once on a side stream (so that workspaces are sized outside
    the capture), then captured there; the library's stream is restored."""
    L = _lib.lib()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    try:
        with torch.cuda.stream(side):
            _lib.check(L.taco_hip_set_stream(side.cuda_stream))
            fn()
            side.synchronize()
            with torch.cuda.graph(g, stream=side):
                fn()
    finally:
        torch.cuda.current_stream().wait_stream(side)
        _lib.check(L.taco_hip_set_stream(torch.cuda.current_stream().cuda_stream))
    return g


def _replays(g, out, check, n=4):
    for r in range(n):
        out.fill_(float("nan"))          # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check(out.cpu().numpy().astype(np.float64), r)


def test_spmv_nnz_split_graph_replays():
    A = W.csr_powerlaw(30000, 30000, seed=61)
    x = W.uniform(A.n, 61, W.STREAM_X)
    a = G.csr_desc(A)
    xd = G.dev(x)
    y = torch.empty(A.m, dtype=torch.float64, device="cuda")
    s = G.sched(_lib.NNZ_SPLIT, 256, items=8)
    L = _lib.lib()
    g = _capture(lambda: _lib.check(L.taco_hip_spmv_compute(D.dense(y).ptr, a.ptr,
                                                            D.dense(xd).ptr, s)))
    ye, cond = oracle.exact_spmv(A, x)

    def check(got, r):
        rep = oracle.error_report(got, ye, cond, 1e-12)
        assert rep["ok"], (r, rep)

    _replays(g, y, check)


@pytest.mark.parametrize("dtype,K", [("f64", 16), ("bf16", 64)], ids=["nnz_split_f64", "mfma_bf16"])
def test_spmm_graph_replays(dtype, K):
    A = W.csr_powerlaw(20000, 18000, seed=62, kmax=5000)
    B = W.uniform(A.n * K, 62, W.STREAM_B).reshape(A.n, K)
    a = G.csr_desc(A, dtype)
    if dtype == "bf16":
        bdesc = D.dense(G.bf16_bits(B).to("cuda"), "bf16")
        C = torch.empty((A.m, K), dtype=torch.float32, device="cuda")
        cdesc = D.dense(C, "f32")
        s, tol = G.sched(_lib.MFMA), 1e-2
    else:
        bdesc = D.dense(G.dev(B, torch.float64), "f64")
        C = torch.empty((A.m, K), dtype=torch.float64, device="cuda")
        cdesc = D.dense(C, "f64")
        s, tol = G.sched(_lib.NNZ_SPLIT, 256, items=8), 1e-12
    L = _lib.lib()
    g = _capture(lambda: _lib.check(L.taco_hip_spmm_compute(cdesc.ptr, a.ptr, bdesc.ptr, s)))
    Ce, cond = oracle.exact_spmm(A, B)

    def check(got, r):
        rep = oracle.error_report(got, Ce, cond, tol)
        assert rep["ok"], (r, rep)

    _replays(g, C, check)
