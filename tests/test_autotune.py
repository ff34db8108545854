"""The autotuner's output check (host logic, CPU tensors)."""
import math

import torch

from taco_amd import autotune


def test_accepts_within_tolerance_only():
    ref = torch.tensor([1.0, -4.0, 2.0], dtype=torch.float64)
    assert autotune.accepts(ref.clone(), ref, 1e-10)
    assert autotune.accepts(ref + 3e-11, ref, 1e-10)       # 3e-11 <= 1e-10 * 4
    assert not autotune.accepts(ref + 5e-10, ref, 1e-10)
    bad = ref.clone()
    bad[1] = math.nan
    assert not autotune.accepts(bad, ref, 1e-10)
    assert autotune.accepts(torch.zeros(0), torch.zeros(0), 1e-10)
    # float32 outputs are compared in float64
    assert autotune.accepts(ref.float(), ref.float(), 1e-6)


def test_rejected_candidate_ranks_last(monkeypatch):
    """A candidate whose output differs is never timed and ranks last."""
    out = torch.zeros(4, dtype=torch.float64)
    timed = []

    def run(v):
        def fn():
            out.fill_(v)
        return fn

    monkeypatch.setattr(autotune, "_time", lambda fn, reps: timed.append(fn) or 1.0)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda: None)
    runs = [("ref", run(2.0), 0.0), ("same", run(2.0), 0.25), ("wrong", run(2.5), 0.0),
            ("unavailable", None, 0.0)]
    res = autotune._tune(runs, out, 3, 1e-10)
    assert [n for n, _ in res] == ["ref", "same", "wrong", "unavailable"]
    assert res[1][1] == 1.25  # the extra per-call cost (a plan's build share) is added
    assert res[-1][1] == math.inf and res[-2][1] == math.inf and len(timed) == 2


def test_plan_candidate_replaces_the_sixteen_entry_split():
    names = [n for n, _ in autotune.SPMV_CANDIDATES]
    assert len(names) == 13 and "slab_plan" in names and "nnz_split16" not in names
    assert dict(autotune.SPMV_CANDIDATES)["slab_plan"] is None
    assert autotune.PLAN_CALLS == 100
