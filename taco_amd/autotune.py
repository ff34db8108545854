"""Exhaustive autotuner for the hot-path schedules.

Two forms:
  * ``tune_batch`` — per-instance schedules of a batched launch.  For every
    candidate lane count, each instance is launched alone ``reps`` times
    between two HIP events (all instances queued back to back, one wait per
    candidate).  Each candidate's whole-batch output is first checked
    against the reference candidate (one lane per row, the sequential
    summation order); a candidate whose result differs beyond the value
    tolerance is never selected.  Instances alone do not see the L2 and
    queue contention of the whole launch, so the per-instance choice is
    then raced, as whole-batch launches, against every accepted candidate
    applied to all instances; the fastest whole launch wins.
  * ``tune_spmv`` / ``tune_spmm`` — one matrix, every schedule in the
    candidate space timed with HIP events; returns the ranking.  The first
    candidate (one thread per row, the sequential summation order) gives
    the reference output; every other candidate's output is checked against
    it before it is timed, and one that differs beyond the value tolerance
    is ranked last with time inf.

The candidate spaces are the schedule families the kernels implement; they
are also the classes the learned selector predicts.  The SpMV space includes
the slab plan (taco_hip_spmv_plan_*, f64): a one-time reordering of A that is
then reused by every call on the same matrix.  Its build time is charged to
the calls that reuse it, PLAN_CALLS of them: its candidate time is the
per-call time plus build time / PLAN_CALLS.
"""
import numpy as np

from . import _lib
from . import batch as BT


def _sched(algo, block=0, items=0, group=0, rows=0):
    s = _lib.Schedule()
    s.algo, s.block_size, s.items_per_thread = algo, block, items
    s.group_size, s.rows_per_group = group, rows
    return s


# Calls that share one slab plan's build cost (an iterative solver's
# products with one matrix).
PLAN_CALLS = 100

# The SpMV schedule space (also the selector's SpMV classes); None is the
# slab plan.
SPMV_CANDIDATES = [
    ("thread_per_row", _sched(_lib.THREAD_PER_ROW, 256)),
    ("group2", _sched(_lib.GROUP_PER_ROW, 256, group=2)),
    ("group4", _sched(_lib.GROUP_PER_ROW, 256, group=4)),
    ("group8", _sched(_lib.GROUP_PER_ROW, 256, group=8)),
    ("group16", _sched(_lib.GROUP_PER_ROW, 256, group=16)),
    ("group32", _sched(_lib.GROUP_PER_ROW, 256, group=32)),
    ("group64", _sched(_lib.GROUP_PER_ROW, 256, group=64)),
    ("nnz_split4", _sched(_lib.NNZ_SPLIT, 256, items=4)),
    ("nnz_split8", _sched(_lib.NNZ_SPLIT, 256, items=8)),
    ("slab_plan", None),
    ("merge4", _sched(_lib.MERGE_PATH, 256, items=4)),
    ("merge8", _sched(_lib.MERGE_PATH, 256, items=8)),
    ("merge16", _sched(_lib.MERGE_PATH, 256, items=16)),
]

# The SpMM schedule space for a dense width K (lane teams of up to 64).
def spmm_candidates(K):
    g = 8
    while g < K and g < 64:
        g *= 2
    out = [("thread_per_row", _sched(_lib.THREAD_PER_ROW, 256))]
    for rows in (1, 2, 4):
        out.append((f"rows{rows}", _sched(_lib.GROUP_PER_ROW, 256, group=g, rows=rows)))
    for items in (4, 8, 16):
        out.append((f"nnz_split{items}", _sched(_lib.NNZ_SPLIT, 256, items=items, group=g)))
    for items in (64, 128, 256, 512):
        out.append((f"merge{items}", _sched(_lib.MERGE_PATH, 256, items=items, group=g)))
    return out


def _time(fn, reps, warmup=2):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def accepts(out, ref, rtol):
    """True when out matches ref to rtol of ref's largest magnitude."""
    if ref.numel() == 0:
        return True
    scale = max(float(ref.abs().max().item()), 1e-30)
    diff = (out.double() - ref.double()).abs().max().item()
    return bool(diff <= rtol * scale)  # NaN fails


def _tune(runs, result, reps, rtol):
    """runs: (name, fn, extra_ms) with fn None for a candidate that cannot
    run; extra_ms is added to its per-call time."""
    import torch
    out, ref = [], None
    for name, fn, extra in runs:
        if fn is None:
            out.append((name, float("inf")))
            continue
        fn()
        torch.cuda.synchronize()
        if ref is None:
            ref = result.clone()
        elif not accepts(result, ref, rtol):
            out.append((name, float("inf")))
            continue
        out.append((name, _time(fn, reps) + extra))
    return sorted(out, key=lambda t: t[1])


def slab_plan(a_desc):
    """(plan, build ms), or (None, inf) when the plan does not apply (f32
    values, or a shape outside its limits)."""
    import time
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    try:
        plan = _lib.SpmvPlan(a_desc)
    except _lib.TacoHipError:
        return None, float("inf")
    torch.cuda.synchronize()
    return plan, (time.perf_counter() - t0) * 1e3


def spmv_runner(s, a_desc, x_desc, y_desc, plan=None):
    """The call that runs SpMV schedule s (None: the slab plan `plan`)."""
    L = _lib.lib()
    if s is None:
        if plan is None:
            return None
        return lambda: plan.compute(y_desc, x_desc)
    return lambda: _lib.check(L.taco_hip_spmv_compute(y_desc.ptr, a_desc.ptr, x_desc.ptr, s))


def tune_spmv(a_desc, x_desc, y_desc, reps=5, candidates=None, rtol=1e-10):
    """Times every SpMV schedule on one matrix; returns [(name, ms)] sorted
    (inf: output rejected, or the plan does not apply).  y_desc must describe
    a device tensor.  The slab plan's time includes its build time /
    PLAN_CALLS."""
    runs = []
    for name, s in candidates or SPMV_CANDIDATES:
        plan, extra = (slab_plan(a_desc) if s is None else (None, 0.0))
        runs.append((name, spmv_runner(s, a_desc, x_desc, y_desc, plan), extra / PLAN_CALLS
                     if s is None else 0.0))
    return _tune(runs, y_desc.keep[0], reps, rtol)


def tune_spmm(a_desc, b_desc, c_desc, K, reps=3, candidates=None, rtol=1e-4):
    """Times every SpMM schedule on one matrix; returns [(name, ms)] sorted
    (inf: output rejected).  c_desc must describe a device tensor."""
    L = _lib.lib()
    runs = [(name, lambda s=s: _lib.check(
        L.taco_hip_spmm_compute(c_desc.ptr, a_desc.ptr, b_desc.ptr, s)), 0.0)
        for name, s in candidates or spmm_candidates(K)]
    return _tune(runs, c_desc.keep[0], reps, rtol)


def tune_batch(wl, op, reps=3, rtol=1e-4, check=True):
    """Exhaustive per-instance schedule search for a batched launch.

    Returns (chosen lanes per instance, times[candidate, instance] in ns,
    candidate lane counts, candidates rejected by the output check).  The
    whole-batch times of the raced schedules are in tune_batch.last_race
    ({name: ms})."""
    import torch
    cands = list(BT.SPMV_LANES if op == 0 else BT.SPMM_LANES)
    ref_lanes = cands[0]
    b = BT.Batch(wl, op, np.full(wl.ninst, ref_lanes, np.int32))
    out = wl.y if op == 0 else wl.C
    b.run()
    torch.cuda.synchronize()
    ref = out.clone()
    scale = float(ref.abs().max().item()) if ref.numel() else 1.0
    times = np.zeros((len(cands), wl.ninst))
    rejected = []
    for ci, lanes in enumerate(cands):
        b.set_schedules(np.full(wl.ninst, lanes, np.int32))
        b.run()
        torch.cuda.synchronize()
        times[ci] = b.time_instances(reps)
        if check:
            diff = float((out - ref).abs().max().item())
            if diff > rtol * max(scale, 1e-30):
                rejected.append(lanes)
                times[ci] = np.inf
    best = np.array(cands, np.int32)[np.argmin(times, axis=0)]
    race = {"per_instance": best}
    for lanes in cands:
        if lanes not in rejected:
            race[f"all_{lanes}"] = np.full(wl.ninst, lanes, np.int32)
    race_ms = {}
    for name, sched in race.items():
        b.set_schedules(sched)
        race_ms[name] = _time(b.run, reps, warmup=1)
    tune_batch.last_race = race_ms
    chosen = race[min(race_ms, key=race_ms.get)]
    b.set_schedules(chosen)
    return chosen, times, cands, rejected
