"""The batched workload at its benchmark size: 4,096 CSR instances (rows
log-uniform in [1e3, 2e5], row lengths uniform in [2, 64], uniform / banded
/ 8x8-block structure, f32, SpMM width 32), about 3.85e9 entries in all, so
entry offsets pass 2^31 and the kernel's 64-bit instance offsets are needed.
One hipGraph launch each of SpMV and SpMM over the whole batch; a seeded
sample of instances — among them the one whose entries straddle 2^31, the
first one past it and the last one — is checked against the FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
from taco_amd import batch as BT

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
SEED = 4  # the benchmark's workload seed


@pytest.fixture(scope="module")
def full():
    print("generating the 4,096-instance batch", flush=True)
    w = BT.BatchedWorkload(ninst=4096, rmin=1000, rmax=200000, kmin=2, kmax=64,
                           seed=SEED, width=32)
    print(f"  {w.total_rows} rows, {w.total_nnz} entries", flush=True)
    yield w
    del w
    torch.cuda.empty_cache()


def sample(w):
    big = 2 ** 31
    straddle = int(np.searchsorted(w.ent_off, big, side="right")) - 1
    first_past = int(np.argmax(w.ent_off[:-1] >= big))
    rng = np.random.default_rng(11)
    picks = {0, straddle, first_past, w.ninst - 1,
             int(np.argmax(w.m)), *rng.choice(w.ninst, 5, replace=False).tolist()}
    # one instance of every structure
    for s in (0, 1, 2):
        picks.add(int(np.flatnonzero(w.structure == s)[0]))
    return sorted(picks), straddle, first_past


def test_batch_is_the_benchmark_size(full):
    assert full.ninst == 4096
    assert full.total_nnz > 2 ** 31 and full.total_nnz > 3.8e9
    assert set(full.structure.tolist()) == {0, 1, 2}
    assert full.m.min() >= 1000 and full.m.max() <= 200000
    insts, straddle, first_past = sample(full)
    assert full.ent_off[straddle] < 2 ** 31 < full.ent_off[straddle + 1]
    assert full.ent_off[first_past] >= 2 ** 31 and full.ent_off[first_past - 1] < 2 ** 31


@pytest.mark.parametrize("lanes", [0, 1], ids=["stream", "thread_per_row"])
def test_full_batch_spmv_graph(full, lanes):
    insts, _, _ = sample(full)
    full.y.fill_(float("nan"))
    b = BT.Batch(full, 0, np.full(full.ninst, lanes, np.int32))
    b.capture()
    b.launch_graph()
    torch.cuda.synchronize()
    del b
    assert not torch.isnan(full.y).any()
    for i in insts:
        A, x, _ = full.instance(i)
        ye, cond = oracle.exact_spmv(A, x)
        rep = oracle.error_report(full.y_of(i), ye, cond, TOL32)
        assert rep["ok"], (i, int(full.ent_off[i]), rep)


def test_full_batch_spmm_graph(full):
    insts, _, _ = sample(full)
    full.C.fill_(float("nan"))
    b = BT.Batch(full, 1, BT.default_schedule(1, full.ninst))
    b.capture()
    b.launch_graph()
    torch.cuda.synchronize()
    del b
    assert not torch.isnan(full.C).any()
    for i in insts:
        A, _, B = full.instance(i)
        Ce, cond = oracle.exact_spmm(A, B)
        rep = oracle.error_report(full.c_of(i), Ce, cond, TOL32)
        assert rep["ok"], (i, int(full.ent_off[i]), rep)
