"""Builds the selector's timing database and trains it on the GPU.

    python -m taco_amd.train_selector timings --seed 7 --start 0 --count 2000 --out DIR
    python -m taco_amd.train_selector merge --db 'DIR/timings_s7_*.npz' --out taco_amd/data/selector_train.npz
    python -m taco_amd.train_selector train --db taco_amd/data/selector_train.npz --out taco_amd/data/selector.npz
    python -m taco_amd.train_selector evaluate --db taco_amd/data/selector_heldout.npz --weights taco_amd/data/selector.npz
    python -m taco_amd.train_selector retime --db taco_amd/data/selector_train.npz --seed 7 --count 20000 --cls slab_plan --out DIR/selector_train.npz

`timings` generates problems (taco_amd.selector.make_problem(seed, i) for i
in [start, start+count)), computes their features on the GPU and times every
schedule class with the exhaustive autotuner, writing one npz per call
(features F[N, 32], times T[N, 24] in ms, op, K, rows, nnz, structure).
`retime` times one class again on every problem of a merged database (used
when a class is added or changed) and writes the database with that column
replaced.  `train` fits the MLP with the library's HIP training kernels.  `evaluate`
reports learned vs exhaustive vs default totals on a database.
"""
import argparse
import glob
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cmd_timings(args):
    import torch
    from . import _lib, selector as S
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().taco_hip_set_stream(torch.cuda.current_stream().cuda_stream))
    idx = [args.start + j for j in range(args.count)
           if args.op < 0 or S.problem_op(args.seed, args.start + j) == args.op]
    n = len(idx)
    F = np.zeros((n, S.N_FEATURES), np.float32)
    T = np.full((n, S.N_CLASSES), np.inf)
    meta = {k: np.zeros(n, np.int64) for k in ("op", "K", "m", "nnz", "structure", "index")}
    tag = "" if args.op < 0 else f"_op{args.op}"
    out = f"{args.out}/timings_s{args.seed}_{args.start}_{args.start + args.count}{tag}.npz"
    t0 = time.time()
    for j, i in enumerate(idx):
        A, op, K, d = S.make_problem(args.seed, i, rmin=args.rmin, rmax=args.rmax)
        prob = S.DeviceProblem(A, op, K, seed=i)
        F[j] = S.features(prob.a, op, K)
        T[j] = S.time_classes(prob, reps=args.reps)
        meta["op"][j], meta["K"][j], meta["m"][j], meta["nnz"][j] = op, K, A.m, A.nnz
        meta["structure"][j] = S.STRUCTURES.index(d["structure"])
        meta["index"][j] = i
        del prob, A
        if (j + 1) % args.progress == 0 or j + 1 == n:
            log(f"[{j + 1}/{n}] {time.time() - t0:.0f}s; last: {d} best "
                f"{S.CLASS_NAMES[int(np.argmin(T[j]))]} {np.min(T[j]):.3f} ms")
            # the problems timed so far, rewritten at every progress line
            np.savez(out, F=F[:j + 1], T=T[:j + 1], seed=args.seed, start=args.start,
                     rmin=args.rmin, rmax=args.rmax, **{k: v[:j + 1] for k, v in meta.items()})
    log(f"wrote {out} in {time.time() - t0:.0f}s")


def problem_key(seed, index, rmin=10_000, rmax=2_000_000):
    """(op, K, rows, structure) of make_problem(seed, index) from its first
    random draws, without generating the matrix."""
    from . import selector as S
    rng = np.random.default_rng([seed, index])
    op = int(rng.integers(0, 2))
    K = int(rng.choice(S.SPMM_WIDTHS)) if op == 1 else 0
    hi = rmax if op == 0 else max(rmin, min(rmax, (1 << 25) // K))
    m = int(np.exp(rng.uniform(np.log(rmin), np.log(hi))))
    return op, K, m, int(rng.integers(0, 4))


def cmd_retime(args):
    """Times one class again for every problem of a merged database and
    writes the database with that column replaced.  The database holds no
    problem indices: each row is matched to make_problem(seed, i) by its
    (op, K, rows, structure), and rows that match several indices by their
    entry count."""
    import collections
    import torch
    from . import _lib, selector as S
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().taco_hip_set_stream(torch.cuda.current_stream().cuda_stream))
    c = S.CLASS_NAMES.index(args.cls)
    op_c = 0 if c < S.N_SPMV else 1
    db = dict(np.load(args.db, allow_pickle=False))
    n = len(db["op"])
    keys = collections.defaultdict(list)
    for i in range(args.count):
        keys[problem_key(args.seed, i)].append(i)
    T = db["T"].astype(np.float64)
    todo = [r for r in range(n) if int(db["op"][r]) == op_c]
    log(f"{args.db}: {n} problems, retiming class {args.cls} on {len(todo)}")
    t0 = time.time()
    used = set()
    for q, r in enumerate(todo):
        k = (int(db["op"][r]), int(db["K"][r]), int(db["m"][r]), int(db["structure"][r]))
        cands = [i for i in keys[k] if i not in used]
        A = None
        for i in cands:
            A, op, K, d = S.make_problem(args.seed, i)
            if A.nnz == int(db["nnz"][r]) or len(cands) == 1:
                break
        if A is None or A.nnz != int(db["nnz"][r]):
            raise SystemExit(f"row {r}: no problem of seed {args.seed} matches {k}")
        used.add(i)
        prob = S.DeviceProblem(A, op, K, seed=i)
        T[r, c] = retime_class(prob, c, args.reps)
        del prob, A
        if (q + 1) % args.progress == 0 or q + 1 == len(todo):
            log(f"[{q + 1}/{len(todo)}] {time.time() - t0:.0f}s; last: {d} "
                f"{args.cls} {T[r, c]:.4f} ms, best {S.CLASS_NAMES[int(np.argmin(T[r]))]}")
    db["T"] = T.astype(np.float32)
    np.savez_compressed(args.out, **db)
    log(f"wrote {args.out}")


def retime_class(prob, c, reps):
    """Time of class c on prob (ms, inf if it does not apply or its output
    differs from the sequential schedule beyond the value tolerance), as
    selector.time_classes measures it."""
    import torch
    from . import selector as S
    lo, _ = S.class_range(prob.op)
    rtol = 1e-10 if prob.op == 0 else 1e-4
    prob.runner(lo)()
    torch.cuda.synchronize()
    ref = prob.out.clone()
    scale = max(float(ref.abs().max().item()) if ref.numel() else 0.0, 1e-30)
    fn = prob.runner(c)
    if fn is None:
        return float("inf")
    fn()
    torch.cuda.synchronize()
    if ref.numel() and float((prob.out - ref).abs().max().item()) > rtol * scale:
        return float("inf")
    e = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    e[0].record()
    for r in range(reps):
        fn()
        e[r + 1].record()
    torch.cuda.synchronize()
    t = float(np.median([e[r].elapsed_time(e[r + 1]) for r in range(reps)]))
    if c == S.PLAN_CLASS and prob.op == 0:
        from . import autotune
        t += prob.plan_build_ms / autotune.PLAN_CALLS
    return t


def load_db(patterns):
    parts = []
    for pat in patterns:
        for f in sorted(glob.glob(pat)):
            parts.append(np.load(f, allow_pickle=False))
    if not parts:
        raise SystemExit(f"no timing files match {patterns}")
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("F", "T", "op", "K", "m", "nnz",
                                                               "structure")}
    cat["T"] = cat["T"].astype(np.float64)
    ok = np.isfinite(cat["T"].min(axis=1))
    return {k: v[ok] for k, v in cat.items()}


def cmd_merge(args):
    """Concatenates timing files into one compact database (float32)."""
    db = load_db(args.db)
    if args.op >= 0:
        keep = db["op"] == args.op
        db = {k: v[keep] for k, v in db.items()}
    np.savez_compressed(args.out, F=db["F"].astype(np.float32), T=db["T"].astype(np.float32),
                        op=db["op"].astype(np.int8), K=db["K"].astype(np.int16),
                        m=db["m"].astype(np.int32), nnz=db["nnz"].astype(np.int64),
                        structure=db["structure"].astype(np.int8))
    log(f"wrote {args.out}: {len(db['F'])} problems")


def cmd_train(args):
    import torch
    from . import _lib, selector as S
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().taco_hip_set_stream(torch.cuda.current_stream().cuda_stream))
    db = load_db(args.db)
    log(f"training on {len(db['F'])} problems")
    t0 = time.time()
    params, mu, sigma, hist = S.train(db["F"], db["T"], epochs=args.epochs, batch=args.batch,
                                      lr=args.lr, weight_decay=args.wd, seed=args.seed, log=log)
    secs = time.time() - t0
    sel_choice = None
    S.save(args.out, params, mu, sigma,
           {"problems": int(len(db["F"])), "epochs": args.epochs, "batch": args.batch,
            "lr": args.lr, "train_seconds": secs})
    sel = S.Selector(args.out)
    sel_choice, _ = sel.choose_from_features(db["F"])
    ev = S.evaluate(sel_choice, db["T"], db["op"])
    log(f"trained in {secs:.1f}s; on the training set: {json.dumps(ev)}")
    print(json.dumps({"train_seconds": secs, "final_loss": hist[-1][1], "train_eval": ev}))


def cmd_evaluate(args):
    import torch
    from . import _lib, selector as S
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().taco_hip_set_stream(torch.cuda.current_stream().cuda_stream))
    db = load_db(args.db)
    sel = S.Selector(args.weights)
    choice, _ = sel.choose_from_features(db["F"])
    out = {"problems": int(len(db["F"]))}
    out.update(S.evaluate(choice, db["T"], db["op"]))
    for op in (0, 1):
        k = db["op"] == op
        if k.any():
            out[("spmv", "spmm")[op]] = S.evaluate(choice[k], db["T"][k], db["op"][k])
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("timings")
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--start", type=int, default=0)
    t.add_argument("--count", type=int, default=100)
    t.add_argument("--rmin", type=int, default=10_000)
    t.add_argument("--rmax", type=int, default=2_000_000)
    t.add_argument("--reps", type=int, default=3)
    t.add_argument("--progress", type=int, default=50)
    t.add_argument("--op", type=int, default=-1, help="only problems of this op (0/1)")
    t.add_argument("--out", default=".")
    r = sub.add_parser("train")
    r.add_argument("--db", nargs="+", required=True)
    r.add_argument("--out", required=True)
    r.add_argument("--epochs", type=int, default=300)
    r.add_argument("--batch", type=int, default=256)
    r.add_argument("--lr", type=float, default=2e-3)
    r.add_argument("--wd", type=float, default=1e-5)
    r.add_argument("--seed", type=int, default=0)
    g = sub.add_parser("merge")
    g.add_argument("--db", nargs="+", required=True)
    g.add_argument("--out", required=True)
    g.add_argument("--op", type=int, default=-1, help="keep only problems of this op")
    t2 = sub.add_parser("retime")
    t2.add_argument("--db", required=True)
    t2.add_argument("--seed", type=int, required=True)
    t2.add_argument("--count", type=int, required=True, help="problems drawn for the database")
    t2.add_argument("--cls", required=True, help="class name to time again")
    t2.add_argument("--reps", type=int, default=3)
    t2.add_argument("--progress", type=int, default=250)
    t2.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--db", nargs="+", required=True)
    e.add_argument("--weights", required=True)
    args = ap.parse_args()
    {"timings": cmd_timings, "merge": cmd_merge, "train": cmd_train,
     "evaluate": cmd_evaluate, "retime": cmd_retime}[args.cmd](args)


if __name__ == "__main__":
    main()
