"""Batched uneven SpMV/SpMM in one persistent launch: every instance against
the oracle, identical results for every schedule, hipGraph replays equal to
eager launches, and the autotuner's per-instance choice."""
import numpy as np
import pytest
import torch

import oracle
from taco_amd import batch as BT
from taco_amd import autotune

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


@pytest.fixture(scope="module")
def wl():
    w = BT.BatchedWorkload(ninst=48, rmin=100, rmax=4000, seed=7, width=32, keep_host=True)
    assert set(w.structure.tolist()) == {0, 1, 2}
    return w


def check_spmv(wl, insts):
    for i in insts:
        A, x, _ = wl.instance(i)
        ye, cond = oracle.exact_spmv(A, x)
        rep = oracle.error_report(wl.y_of(i), ye, cond, TOL32)
        assert rep["ok"], (i, rep)


def check_spmm(wl, insts):
    for i in insts:
        A, _, B = wl.instance(i)
        Ce, cond = oracle.exact_spmm(A, B)
        rep = oracle.error_report(wl.c_of(i), Ce, cond, TOL32)
        assert rep["ok"], (i, rep)


def test_generated_patterns_match_host(wl):
    for i in (0, 7, 31):
        pos, crd = BT.instance_pattern(i, int(wl.m[i]), int(wl.structure[i]), wl.seed,
                                       wl.kmin, wl.kmax)
        A, _, _ = wl.instance(i)
        assert np.array_equal(A.pos, pos) and np.array_equal(A.crd, crd)


@pytest.mark.parametrize("lanes", BT.SPMV_LANES)
def test_batched_spmv_every_schedule(wl, lanes):
    wl.y.zero_()
    b = BT.Batch(wl, 0, np.full(wl.ninst, lanes, np.int32), rows_per_task=128)
    b.run()
    torch.cuda.synchronize()
    check_spmv(wl, range(wl.ninst))


@pytest.mark.parametrize("lanes", BT.SPMM_LANES)
def test_batched_spmm_every_schedule(wl, lanes):
    wl.C.zero_()
    b = BT.Batch(wl, 1, np.full(wl.ninst, lanes, np.int32), rows_per_task=64)
    b.run()
    torch.cuda.synchronize()
    check_spmm(wl, range(0, wl.ninst, 3))


def test_mixed_schedules_and_graph_replay(wl):
    rng = np.random.default_rng(3)
    sched = rng.choice(BT.SPMV_LANES, size=wl.ninst).astype(np.int32)
    b = BT.Batch(wl, 0, sched)
    wl.y.zero_()
    b.run()
    torch.cuda.synchronize()
    eager = wl.y.clone()
    check_spmv(wl, range(wl.ninst))
    b.capture()
    for rep in range(4):
        wl.y.fill_(float("nan"))
        b.launch_graph()
        torch.cuda.synchronize()
        # eager work between replays, then compare every replay
        assert torch.equal(wl.y, eager), rep
        b.run()
        torch.cuda.synchronize()


def test_batched_equals_unbatched(wl):
    """Per-instance results of the batch equal a one-instance batch."""
    b = BT.Batch(wl, 0, np.full(wl.ninst, 8, np.int32))
    wl.y.zero_()
    b.run()
    torch.cuda.synchronize()
    full = wl.y.clone()
    single = BT.BatchedWorkload(ninst=1, rmin=int(wl.m[0]), rmax=int(wl.m[0]), seed=wl.seed,
                                width=32, keep_host=True)
    # instance 0 of both workloads has the same seed-derived pattern and values
    assert np.array_equal(single.instance(0)[0].crd, wl.instance(0)[0].crd)
    sb = BT.Batch(single, 0, np.full(1, 8, np.int32))
    sb.run()
    torch.cuda.synchronize()
    assert np.array_equal(single.y_of(0), full[:int(wl.m[0])].cpu().numpy().astype(np.float64))


def test_run_range_writes_only_its_instances(wl):
    b = BT.Batch(wl, 0, np.full(wl.ninst, 4, np.int32))
    wl.y.zero_()
    b.run_range(5, 9)
    torch.cuda.synchronize()
    check_spmv(wl, range(5, 9))
    r0, r1 = int(wl.row_off[5]), int(wl.row_off[9])
    assert not wl.y[:r0].any() and not wl.y[r1:].any()
    b.run_range(9, 9)  # empty range: no launch, no error
    t = b.time_instances(2)
    assert t.shape == (wl.ninst,) and np.all(np.isfinite(t)) and np.all(t > 0)


def test_autotune_batch(wl):
    best, times, cands, rejected = autotune.tune_batch(wl, 0, reps=2)
    assert rejected == []
    assert best.shape == (wl.ninst,) and set(best.tolist()) <= set(cands)
    race = autotune.tune_batch.last_race
    assert set(race) == {"per_instance"} | {f"all_{l}" for l in cands}
    per_inst = np.array(cands, np.int32)[np.argmin(times, axis=0)]
    # the choice is the fastest whole launch: the per-instance minimum or
    # one candidate for every instance
    winner = min(race, key=race.get)
    if winner == "per_instance":
        assert np.array_equal(best, per_inst)
    else:
        assert np.all(best == int(winner.split("_")[1]))
    b = BT.Batch(wl, 0, best)
    wl.y.zero_()
    b.run()
    torch.cuda.synchronize()
    check_spmv(wl, range(wl.ninst))


@pytest.mark.parametrize("ninst", [1, 3, 11])
def test_batch_fewer_instances_than_queues(ninst):
    """Per-XCD queues with empty queues: every task still runs exactly once."""
    w = BT.BatchedWorkload(ninst=ninst, rmin=70, rmax=300, seed=20 + ninst, width=32,
                           keep_host=True)
    for lanes in (0, 4):
        w.y.fill_(float("nan"))
        b = BT.Batch(w, 0, np.full(w.ninst, lanes, np.int32), rows_per_task=16)
        b.run()
        torch.cuda.synchronize()
        check_spmv(w, range(w.ninst))
        del b
